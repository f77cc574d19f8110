"""libpimi host side against the oracle (no GPU needed).

The product's C++ lowering (exact __int128 rationals) must reproduce the
oracle's exact-rational lowering bit for bit: CG contractions, branch scales,
collective blocks and fused Hamiltonian blocks.  The fused loss blocks contain
K'K Gram products whose summation order differs (<= 1 ulp).  Also checks that
the C ABI library loads and exports every symbol include/pimi.h declares.
"""
import os
import re

import numpy as np
import pytest

import pimi
import pimi._native as N
import planutil as PU
from oracle import lowering as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SP2 = pimi.spin_matrices(2)
SP3 = pimi.spin_matrices(3)


def _compare(model, terms, N_, d):
    c = pimi.compile(model, host_only=True)
    e = c.export_fused()
    b = L.Basis(N_, d)
    g = L.OneBodyGeometry(b)
    plan = L.compile_model(b, terms, g)
    H, Q, K = PU.exported_blocks(e, b.dims)
    for s in range(len(b.dims)):
        np.testing.assert_array_equal(H[s], plan.H[s].dense())
        ref_q = plan.Q[s].dense()
        assert np.abs(Q[s] - ref_q).max() <= 4e-16 * max(np.abs(ref_q).max(), 1.0)
        for gi in range(len(K)):
            np.testing.assert_array_equal(K[gi][s], plan.collective_gains[gi][1][s].dense())
    br = PU.exported_branches(e, b.dims)
    assert len(br) == len(plan.onebody_gains)
    for (sa, ga), (sb, gb) in zip(br, plan.onebody_gains):
        assert sa == sb
        assert len(ga) == len(gb)
        for a, o in zip(ga, gb):
            assert a[:3] == o[:3]          # out sector, in sector, branch scale (exact)
            assert a[4] == o[4]            # support choice
            np.testing.assert_array_equal(a[3], o[3])
    idx, val = c.trace_functional()
    ti, tv = L.trace_functional(b)
    np.testing.assert_array_equal(idx, ti)
    np.testing.assert_array_equal(val, tv)
    return c, g


@pytest.mark.parametrize("N_", [1, 3, 8, 20, 45])
def test_qubit_lowering_bitwise(N_):
    model = pimi.qubit_ensemble_model(N_, hamiltonian=SP2.jz + SP2.jx, emission=0.1, dephasing=0.2, pumping=0.02,
                                      collective_emission=0.1)
    terms = L.qubit_ensemble_terms(hamiltonian=L.spin_matrices(2)["jz"] + L.spin_matrices(2)["jx"], emission=0.1,
                                   dephasing=0.2, pumping=0.02, collective_emission=0.1)
    c, g = _compare(model, terms, N_, 2)
    for X in (SP2.jx, SP2.jz, SP2.jm):
        np.testing.assert_array_equal(c.collective_operator(X).data, L.collective_operator(g, X))


@pytest.mark.parametrize("N_", [2, 4, 6])
def test_qutrit_lowering_bitwise(N_):
    B = pimi.PIBasis(N_, 3)
    model = pimi.PIModel(B, [pimi.LocalHamiltonian(1.0 * SP3.jz + 0.7 * SP3.jx), pimi.LocalJump(SP3.jm, 0.1),
                             pimi.CollectiveJump(SP3.jm, 0.05)])
    s3 = L.spin_matrices(3)
    terms = [L.Term("local_h", 1.0 * s3["jz"] + 0.7 * s3["jx"]), L.Term("local_jump", s3["jm"], 0.1),
             L.Term("collective_jump", s3["jm"], 0.05)]
    _compare(model, terms, N_, 3)


def test_c1_model_lowering_bitwise():
    _compare(pimi.Models.driven_qubits(8, drive=1.0, detuning=1.0, emission=0.1, dephasing=0.2, pumping=0.0),
             L.driven_qubits_terms(drive=1.0, detuning=1.0, emission=0.1, dephasing=0.2, pumping=0.0), 8, 2)


def test_plan_info_and_basis():
    c = pimi.compile(pimi.qubit_ensemble_model(120, hamiltonian=SP2.jz + SP2.jx, emission=0.1), host_only=True)
    assert c.info["n"] == 302621 == len(pimi.PIBasis(120, 2))
    assert c.info["nsectors"] == 61 and c.info["max_block"] == 121


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "pimi.h")).read()
    declared = set(re.findall(r"\b(pimi_[a-z_0-9]+)\s*\(", hdr))
    lib = N.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(N.EXPORTS)


def test_errors_are_reported():
    B = pimi.PIBasis(3, 2)
    with pytest.raises(ValueError):
        pimi.PIModel(B, [pimi.LocalJump(np.eye(3))])
    model = pimi.qubit_ensemble_model(3, emission=0.1)
    c = pimi.compile(model, host_only=True)
    with pytest.raises(pimi.PIMIError) as err:
        c.workspace()
    assert err.value.status == N.PIMI_EINVAL
    with pytest.raises(ValueError):
        pimi.GMRESAlgorithm(krylovdim=0)
    with pytest.raises(ValueError):
        pimi.ExpvAlgorithm(safety=1.5)


def test_saved_times_follow_julia_ranges():
    ts = pimi.saved_times((0.0, 4.0), 0.1)
    assert len(ts) == 41 and ts[3] == 0.3 and ts[-1] == 4.0
    ts = pimi.saved_times((0.0, 1.0), 0.3)
    assert list(ts) == [0.0, 0.3, 0.6, 0.9, 1.0]
    with pytest.raises(ValueError):
        pimi.saved_times((0.0, 1.0), [0.0, 0.5])


@pytest.mark.parametrize("N_,d", [(12, 2), (5, 3)])
def test_sector_preconditioner_factors_agree_with_lapack(N_, d):
    """The product's Householder + implicit-QL eigen-solver diagonalizes each
    fused Hamiltonian block with a unitary factor (qutrit blocks are complex
    and not tridiagonal: jy), and for qubits, whose blocks have simple
    spectra, gives the same preconditioner operator as numpy's LAPACK route
    used by the oracle.  (Qutrit irreps have repeated weights, so there the
    operator depends on the basis chosen inside degenerate eigenspaces.)"""
    from oracle import oracle as O
    s_o = L.spin_matrices(d)
    H_o = s_o["jz"] + 0.7 * s_o["jx"] + (0.3 * s_o["jy"] if d == 3 else 0)
    ref, _ = O.compile_plan(N_, d, [L.Term("local_h", H_o), L.Term("local_jump", s_o["jm"], 0.1),
                                    L.Term("local_jump", s_o["jz"], 0.2)])
    U, inv = ref.sector_preconditioner()
    b = ref.basis
    rng = np.random.default_rng(1)
    r = rng.standard_normal(b.n) + 1j * rng.standard_normal(b.n)

    def apply(U, inv, r):
        out = np.empty_like(r)
        for s, m in enumerate(b.dims):
            o0, o1 = b.offsets[s], b.offsets[s + 1]
            Us = U[o0:o1].reshape(m, m, order="F")
            D = inv[o0:o1].reshape(m, m, order="F")
            R = r[o0:o1].reshape(m, m, order="F")
            out[o0:o1] = (Us @ ((Us.conj().T @ R @ Us) * D) @ Us.conj().T).reshape(-1, order="F")
        return out
    sp = pimi.spin_matrices(d)
    H = sp.jz + 0.7 * sp.jx + (0.3 * sp.jy if d == 3 else 0)
    c = pimi.compile(pimi.PIModel(pimi.PIBasis(N_, d), [pimi.LocalHamiltonian(H), pimi.LocalJump(sp.jm, 0.1),
                                                        pimi.LocalJump(sp.jz, 0.2)]), host_only=True)
    Up = np.zeros(b.n, dtype=np.complex128)
    Ip = np.zeros(b.n, dtype=np.complex128)
    N.check(N.lib().pimi_sector_preconditioner_factors(c._plan, N.ptr(Up), N.ptr(Ip)))
    for s, m in enumerate(b.dims):   # unitary factors that diagonalize the Hamiltonian block
        Us = Up[b.offsets[s]:b.offsets[s + 1]].reshape(m, m, order="F")
        assert np.abs(Us.conj().T @ Us - np.eye(m)).max() <= 1e-13
        Hs = ref.plan.H[s].dense()
        Dg = Us.conj().T @ (0.5 * (Hs + Hs.conj().T)) @ Us
        assert np.abs(Dg - np.diag(np.diag(Dg))).max() <= 1e-13 * max(np.abs(Hs).max(), 1.0)
    if d == 3:
        return
    a, o = apply(Up, Ip, r), apply(U, inv, r)
    assert np.abs(a - o).max() <= 1e-12 * np.abs(o).max()


def test_workload_generators_are_seeded():
    from pimi.workloads import batched_instances, served_requests
    a, b = batched_instances(64), batched_instances(64)
    assert a == b and len(a) == 64
    assert {i.N for i in batched_instances(512)} == {20, 40, 60, 80, 100, 120}
    assert all(0 <= i.omega <= 2 and 1e-3 <= i.gamma_phi <= 1e-1 for i in a)
    r = served_requests(200.0, 300)
    assert r == served_requests(200.0, 300)
    arr = np.array([q.arrival for q in r])
    assert np.all(np.diff(arr) > 0)
    assert abs(len(r) / arr[-1] - 200.0) < 40.0          # Poisson rate
    frac = np.mean([q.kind == "steady" for q in r])
    assert 0.5 < frac < 0.7
    assert all(16 <= q.instance.N <= 160 for q in r)
