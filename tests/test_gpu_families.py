"""Rate families and parameter scans on the device
(src/compiled_families.jl:154-300, src/scans.jl:189-300,1457-1480):
a specialization applies the same operator as a model compiled with those
rates (bitwise) and as the oracle (to 1e-13), following
test/test_compiled_families.jl:38-62; a batched scan of the local-only
driven-qubit family reproduces the closed-form product steady state at every
point."""
import os
import sys

import numpy as np
import pytest

import pimi
from oracle import lowering as L
from oracle import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_closed_form as CF  # noqa: E402

pytestmark = pytest.mark.gpu
SP = pimi.spin_matrices(2)


def _prototype(N_=2, rates=(0.17, 0.8, 0.2)):
    return pimi.PIModel(pimi.PIBasis(N_, 2), (pimi.LocalHamiltonian(SP.jx, rate=rates[0]),
                                              pimi.LocalJump(SP.jm, rate=rates[1]),
                                              pimi.LocalJump(SP.jp, rate=rates[2])))


def _oracle(N_, rates):
    sp = L.spin_matrices(2)
    return O.compile_plan(N_, 2, [L.Term("local_h", sp["jx"], rates[0]), L.Term("local_jump", sp["jm"], rates[1]),
                                  L.Term("local_jump", sp["jp"], rates[2])])[0]


@pytest.mark.parametrize("N_", [2, 9])
def test_specialization_applies_the_bound_model(N_):
    print(f"family specialize N={N_}", flush=True)
    fam = pimi.compile_family(_prototype(N_))
    rates = (0.23, 0.71, 0.19)
    spec = pimi.specialize(fam, rates)
    assert spec.rates == rates and spec.family is fam
    direct = pimi.compile(_prototype(N_, rates))
    ref = _oracle(N_, rates)
    n = len(spec.basis)
    x = (np.arange(n) + 1.0) / (n + 1)
    x = x * (1 + 0.3j)
    y = spec.apply(x)
    np.testing.assert_array_equal(y, direct.apply(x))
    assert np.abs(y - ref.apply(x)).max() <= 1e-13 * np.abs(ref.apply(x)).max()
    Ld = O.dense_liouvillian(ref)
    ya = spec.apply(x, adjoint=True)
    assert np.abs(ya - Ld.conj().T @ x).max() <= 1e-13 * np.abs(Ld).max() * np.abs(x).max() * n
    X = np.stack([x, 2 * x], axis=1)
    np.testing.assert_array_equal(spec.apply(X)[:, 1], spec.apply(2 * x))
    assert pimi.specialize(fam).rates == fam.default_rates


def test_batched_scan_matches_closed_form():
    """Family of local-only driven qubits (rates of emission and dephasing
    vary; drive 2 omega, detuning 1): <sum_i n_i> = N sigma_ee at every point."""
    print("family batched scan", flush=True)
    N_, omega = 40, 0.6
    proto = pimi.Models.driven_qubits(N_, drive=2 * omega, detuning=1.0, emission=0.1, dephasing=0.04, pumping=0.0)
    fam = pimi.compile_family(proto, rate_indices=(1, 2))  # local jm, local jz
    points = [(0.1, 0.04), (0.2, 0.004), (0.05, 0.2), (0.3, 0.01), (0.12, 0.06), (0.08, 0.001)]
    plan = pimi.ParameterScanPlan(points, fam, rtol=1e-10)
    res = pimi.parameter_scan(plan)
    assert len(res) == len(points) and res.metadata["batched"]
    H = 2 * omega * CF.jx + 1.0 * CF.jz
    n_op = None
    for k, pt in enumerate(res):
        assert pt.index == k and pt.parameter == points[k] and pt.status == "success" and pt.converged
        assert pt.residual <= 1e-10
        if n_op is None:
            n_op = pimi.compile(proto).collective_operator(SP.jp @ SP.jm)
        val = pimi.expect(n_op, pt.output).real
        want = CF.steady(N_, H, [(points[k][0], CF.jm), (points[k][1], CF.jz)])[0]
        print(f"  point {k}: <n> {val:.12f} closed form {want:.12f}", flush=True)
        assert abs(val - want) <= 1e-7 * want
    sub = pimi.parameter_scan(plan, indices=[4, 1])
    assert [p.index for p in sub] == [1, 4]


def test_scan_failure_policies():
    print("family scan failures", flush=True)
    fam = pimi.compile_family(_prototype(4))
    plan = pimi.ParameterScanPlan([(0.2, 0.5, 0.1), (0.2, np.inf, 0.1)], fam, rtol=1e-9)
    with pytest.raises(ValueError):
        pimi.parameter_scan(plan)
    res = pimi.parameter_scan(plan, on_error="record")
    assert res[0].status == "success" and res[1].status == "failed" and "finite" in res[1].message
    with pytest.raises(RuntimeError):
        pimi.parameter_scan(plan, on_error="throw")
