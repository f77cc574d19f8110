"""batched_sample_jumps / batched_trajectory_rngs (src/batched_trajectories.jl:609-669):
the decision rule restated exactly on crafted draws, the validation of the
reference, and statistical parity of the sampled decisions (the reference's
MersenneTwister streams cannot be reproduced, so the distribution is what is
pinned: jump probability 1 - exp(-lambda h), channel k with rates[k]/lambda)."""
import numpy as np
import pytest

import pimi
from pimi.trajectories import _select_jump_channel


class _Fixed:
    """A stream returning given uniforms in order."""
    def __init__(self, values):
        self.v = list(values)

    def random(self):
        return self.v.pop(0)


def test_decision_rule_on_given_draws():
    rates = np.array([[0.2, 0.0, 1.0], [0.3, 0.0, 0.0], [0.5, 0.0, 0.0]])  # columns: 3 channels, none, one
    h = 0.1
    p = -np.expm1(-1.0 * h)
    ch = np.zeros(3, dtype=np.int64)
    # column 0: first draw below p -> jump; second draw u * lambda = 0.45 falls in channel 2 (0.2 <= 0.45 < 0.5)
    # column 1: zero intensity -> no jump and no draw; column 2: first draw above p -> no jump
    rngs = [_Fixed([p * 0.5, 0.45]), _Fixed([]), _Fixed([p * 1.01])]
    pimi.batched_sample_jumps(ch, rates, h, rngs)
    assert list(ch) == [2, 0, 0]
    assert rngs[0].v == [] and rngs[2].v == []        # one draw for no jump, two for a jump
    # channel selection: cumulative search, trailing zero channels never chosen
    assert _select_jump_channel([0.2, 0.3, 0.5], 0.0) == 1
    assert _select_jump_channel([0.2, 0.3, 0.5], 0.2) == 2
    assert _select_jump_channel([0.2, 0.3, 0.0], 0.5) == 2   # roundoff guard: last positive channel


def test_validation():
    rng = pimi.batched_trajectory_rngs(1, 2)
    ch = np.zeros(2, dtype=np.int64)
    with pytest.raises(ValueError):
        pimi.batched_sample_jumps(ch, np.array([[0.1, -0.2]]), 0.1, rng)
    with pytest.raises(ValueError):
        pimi.batched_sample_jumps(ch, np.array([[0.1, np.nan]]), 0.1, rng)
    with pytest.raises(ValueError):
        pimi.batched_sample_jumps(ch, np.array([[0.1, 0.2]]), 0.0, rng)
    with pytest.raises(ValueError):
        pimi.batched_sample_jumps(ch, np.array([[0.1, 0.2]]), True, rng)
    with pytest.raises(ValueError):
        pimi.batched_sample_jumps(np.zeros(2, dtype=bool), np.array([[0.1, 0.2]]), 0.1, rng)
    with pytest.raises(ValueError):
        pimi.batched_sample_jumps(np.zeros(3, dtype=np.int64), np.array([[0.1, 0.2]]), 0.1, rng)
    with pytest.raises(ValueError):
        pimi.batched_trajectory_rngs(1, 0)


def test_streams_are_per_trajectory_and_reproducible():
    a = pimi.batched_trajectory_rngs(20261204, 4)
    b = pimi.batched_trajectory_rngs(20261204, 4)
    xa = [r.random() for r in a]
    assert xa == [r.random() for r in b]
    assert len(set(xa)) == 4
    # a column's decision depends only on its own stream: regrouping columns keeps it
    rates = np.array([[0.3, 1.2, 0.7, 2.0], [0.6, 0.1, 0.2, 0.4]])
    c1 = pimi.batched_sample_jumps(np.zeros(4, dtype=np.int64), rates, 0.5, pimi.batched_trajectory_rngs(5, 4))
    r = pimi.batched_trajectory_rngs(5, 4)
    perm = [2, 0, 3, 1]
    c2 = pimi.batched_sample_jumps(np.zeros(4, dtype=np.int64), rates[:, perm], 0.5, [r[k] for k in perm])
    assert list(c2) == [c1[k] for k in perm]


def test_statistical_parity_of_decisions():
    """200,000 decisions: jump frequency 1 - exp(-lambda h) and channel shares
    rates/lambda, each within 5 standard errors."""
    rates = np.array([[0.5], [1.5], [0.0], [2.0]])
    h = 0.2
    M = 200_000
    R = np.repeat(rates, M, axis=1)
    ch = pimi.batched_sample_jumps(np.zeros(M, dtype=np.int64), R, h, pimi.batched_trajectory_rngs(7, M))
    lam = rates.sum()
    p = -np.expm1(-lam * h)
    jumps = ch > 0
    assert abs(jumps.mean() - p) <= 5 * np.sqrt(p * (1 - p) / M)
    nj = jumps.sum()
    for k in range(4):
        q = rates[k, 0] / lam
        share = np.mean(ch[jumps] == k + 1)
        assert abs(share - q) <= 5 * np.sqrt(max(q * (1 - q), 1e-12) / nj) + (1e-12 if q == 0 else 0)
    assert not np.any(ch == 3)    # a zero-rate channel is never chosen
