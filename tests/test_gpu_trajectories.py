"""Conditional (no-jump) trajectory cohorts on the device against the oracle.

The oracle has no trajectory code of its own; its pieces give every
reference quantity: the fused plan without gains applies -i[H,X] - {E,X}/2,
single-channel plans apply each jump's gain, and the channels' L^dagger L
blocks give the intensities sum_nu sqrt(f) Re tr(E_nu X_nu)
(src/batched_trajectories.jl:198-265, 402-427).
"""
import copy

import numpy as np
import pytest

import pimi
from oracle import lowering as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu
SP2 = pimi.spin_matrices(2)
N_ = 6


def _setup():
    model = pimi.qubit_ensemble_model(N_, hamiltonian=SP2.jz + SP2.jx, emission=0.1, dephasing=0.2, pumping=0.02,
                                      collective_emission=0.1)
    comp = pimi.compile(model)
    sp = L.spin_matrices(2)
    b = L.Basis(N_, 2)
    g = L.OneBodyGeometry(b)
    plan = L.compile_model(b, L.qubit_ensemble_terms(hamiltonian=sp["jz"] + sp["jx"], emission=0.1, dephasing=0.2,
                                                     pumping=0.02, collective_emission=0.1), g)
    nogain = copy.copy(plan)
    nogain.collective_gains, nogain.onebody_gains = [], []
    # channels in term order: local jm, jz, jp, collective jm
    gains, qs = [], []
    for k, (rate, op) in enumerate(((0.1, sp["jm"]), (0.2, sp["jz"]), (0.02, sp["jp"]))):
        p = copy.copy(plan)
        p.H = p.Q = None
        p.collective_gains, p.onebody_gains = [], [plan.onebody_gains[k]]
        gains.append(O.OraclePlan(p))
        LdL = np.asarray(op).conj().T @ np.asarray(op)
        qs.append((rate, [L.collective_sparse_block(g, LdL, s).dense() for s in range(len(b.dims))]))
    p = copy.copy(plan)
    p.H = p.Q = None
    p.onebody_gains, p.collective_gains = [], [plan.collective_gains[0]]
    gains.append(O.OraclePlan(p))
    jm = sp["jm"]
    qs.append((0.1, [(lambda K: K.conj().T @ K)(L.collective_sparse_block(g, jm, s).dense())
                     for s in range(len(b.dims))]))
    ti, tv = L.trace_functional(b)
    return comp, O.OraclePlan(nogain), gains, qs, b, ti, tv


def _intensity(b, tv, qblocks, scale, x):
    lam, k = 0.0, 0
    for s, m in enumerate(b.dims):
        X = x[b.offsets[s]:b.offsets[s + 1]].reshape(m, m, order="F")
        lam += tv[k] * np.trace(qblocks[s] @ X).real
        k += m
    return scale * lam


def _states(comp, b, n_):
    e = pimi.computational_product_state(comp.basis, 2).data
    gr = pimi.computational_product_state(comp.basis, 1).data
    idx, tv = comp.trace_functional()
    mm = np.zeros(b.n, dtype=complex)
    mm[idx] = tv / (tv @ tv)
    cols = [e, gr, 0.3 * e + 0.7 * gr, mm, 0.5 * mm + 0.5 * e]
    return np.stack(cols[:n_], axis=1)


def test_conditional_action_and_intensities_against_oracle():
    comp, nog, gains, qs, b, ti, tv = _setup()
    X = _states(comp, b, 5)
    Y, lam = pimi.batched_conditional_action(comp, X)
    rates = pimi.batched_channel_intensities(comp, X)
    for c in range(X.shape[1]):
        x = X[:, c]
        lam_ref = sum(_intensity(b, tv, q, r, x) for r, q in qs)
        assert abs(lam[c] - max(lam_ref, 0)) <= 1e-13 * max(1.0, abs(lam_ref))
        for k, (r, q) in enumerate(qs):
            assert abs(rates[k, c] - max(_intensity(b, tv, q, r, x), 0)) <= 1e-13 * max(1.0, lam_ref)
        assert abs(rates[:, c].sum() - lam[c]) <= 1e-13 * max(1.0, lam[c])
        y_ref = nog.apply(x) + max(lam_ref, 0) * x
        assert np.abs(Y[:, c] - y_ref).max() <= 1e-13 * max(np.abs(y_ref).max(), 1.0)


def test_conditional_rk4_against_oracle_and_cohort_equals_columns():
    comp, nog, gains, qs, b, ti, tv = _setup()
    X = _states(comp, b, 4)
    h = 0.05

    def f(x):
        return nog.apply(x) + max(sum(_intensity(b, tv, q, r, x) for r, q in qs), 0) * x

    out = pimi.batched_conditional_rk4(comp, X, h)
    for c in range(X.shape[1]):
        x = X[:, c]
        k1 = f(x)
        k2 = k1.copy()
        k1 = f(x + (h / 2) * k1)
        k2 += 2 * k1
        k1 = f(x + (h / 2) * k1)
        k2 += 2 * k1
        k1 = f(x + h * k1)
        xr = x + (h / 6) * (k2 + k1)
        xr = xr / np.dot(tv, xr[ti])
        assert np.abs(out[:, c] - xr).max() <= 1e-12 * np.abs(xr).max()
        assert abs(np.dot(tv, out[ti, c]) - 1) <= 1e-13
        single = pimi.batched_conditional_rk4(comp, X[:, c:c + 1], h)
        np.testing.assert_array_equal(single[:, 0], out[:, c])


def test_apply_jumps_against_oracle():
    comp, nog, gains, qs, b, ti, tv = _setup()
    X = _states(comp, b, 5)
    channels = [1, 3, 0, 2, 4]   # the all-ground column jumps by pumping (emission has zero probability there)
    out = pimi.batched_apply_jumps(comp, X, channels)
    for c, ch in enumerate(channels):
        if ch == 0:
            np.testing.assert_array_equal(out[:, c], X[:, c])
            continue
        gx = gains[ch - 1].apply(X[:, c])
        ref = gx / np.dot(tv, gx[ti])
        assert np.abs(out[:, c] - ref).max() <= 1e-13 * np.abs(ref).max()
    with pytest.raises(pimi.PIMIError):
        pimi.batched_apply_jumps(comp, X, [1, 2, 3, 4, 5])
