"""Pins of the oracle against mathematics that does not depend on it.

* combinatorial identities and CG isometries (reference: test/test_combinatorics.jl,
  test/test_cgc.jl);
* brute-force d^N Hilbert-space Lindbladians, linked to PI coordinates only by
  physical pairings with products of collective operators (a spanning set by
  Schur--Weyl duality);
* invariants: trace and Hermiticity preservation (test/test_liouvillian_properties.jl);
* closed-form product-state dynamics and steady states of local-only models
  (tests/golden/closed_form.json, written by tests/golden/make_closed_form.py);
* dense-matrix routes for the Krylov solvers (test/test_krylov.jl,
  test/test_krylov_extensions.jl).
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg as sla

import bruteforce as bf
from oracle import lowering as L
from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "closed_form.json")))


@pytest.mark.parametrize("N,d", [(1, 2), (5, 2), (12, 2), (3, 3), (6, 3), (4, 4), (3, 5)])
def test_basis_dimension_identities(N, d):
    b = L.Basis(N, d)
    assert b.n == math.comb(N + d * d - 1, N)
    for p, pats in zip(b.sectors, b.patterns):
        assert len(pats) == L.unitary_group_dimension(p)
        assert sorted(pats) == list(pats)
    # Schur-Weyl: sum_lambda f^lambda dim(U_lambda) = d^N
    assert sum(L.symmetric_group_dimension(p) * m for p, m in zip(b.sectors, b.dims)) == d ** N
    # sectors in descending lexicographic order
    assert b.sectors == sorted(b.sectors, reverse=True)


@pytest.mark.parametrize("N,d", [(6, 2), (4, 3), (3, 4)])
def test_branch_weight_is_multiplicity_ratio(N, d):
    from fractions import Fraction
    for lam in L.partitions(N, d):
        for r in L.removable_corners(lam):
            mu = L.remove_corner(lam, r)
            w = L.one_box_branch_weight(lam, mu)
            assert w == Fraction(N * L.symmetric_group_dimension(mu), L.symmetric_group_dimension(lam))


@pytest.mark.parametrize("N,d", [(3, 2), (4, 2), (3, 3), (4, 3), (2, 4), (3, 4)])
def test_cg_isometry(N, d):
    """For each parent shape and child: sum over (child pattern, label) of
    CG products is the identity on the parent patterns."""
    b = L.Basis(N, d)
    for s, lam in enumerate(b.sectors):
        par = b.patterns[s]
        for mu in b.removed[s]:
            ch = L.gt_patterns(mu)
            M = np.array([[L.cgc(c, j, p, d) for c in ch for j in range(d)] for p in par])
            np.testing.assert_allclose(M @ M.T, np.eye(len(par)), atol=1e-14)


def test_qubit_collective_blocks_are_spin_matrices():
    """For qubits each sector is a spin-j irrep: collective jz/jx blocks have
    the spectra of spin-j matrices and satisfy [Jx, Jy] = i Jz."""
    b = L.Basis(9, 2)
    g = L.OneBodyGeometry(b)
    sp = L.spin_matrices(2)
    for s, p in enumerate(b.sectors):
        j = (p[0] - p[1]) / 2
        Jz = L.collective_block_dense(g, sp["jz"], s)
        Jx = L.collective_block_dense(g, sp["jx"], s)
        Jy = L.collective_block_dense(g, sp["jy"], s)
        np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(Jz)), np.arange(-j, j + 1), atol=1e-13)
        np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(Jx)), np.arange(-j, j + 1), atol=1e-12)
        np.testing.assert_allclose(Jx @ Jy - Jy @ Jx, 1j * Jz, atol=1e-12)


def _bruteforce_check(N, d, terms_pi, terms_full, degree, steps=4):
    b = L.Basis(N, d)
    g = L.OneBodyGeometry(b)
    P = O.OraclePlan(L.compile_model(b, terms_pi, g))
    Lf = bf.lindbladian(N, d, terms_full)
    coll_pi = {(i, j): L.collective_operator(g, bf.unit(d, i, j)) for i in range(d) for j in range(d)}
    coll_f = {(i, j): bf.collective(bf.unit(d, i, j), N, d) for i in range(d) for j in range(d)}
    Dh = d ** N
    ident = L.collective_operator(g, np.eye(d) / N)
    Opi, Of = [], []
    for w in bf.collective_monomials(d, degree):
        a, A = ident.copy(), np.eye(Dh, dtype=complex)
        for ij in w:
            a = L.pi_product(b, a, coll_pi[ij])
            A = A @ coll_f[ij]
        Opi.append(a)
        Of.append(A)
    G = np.array(Opi).conj()
    assert np.linalg.matrix_rank(G, 1e-9) == b.n  # the pairings determine PI vectors

    def to_pi(R):
        return np.linalg.lstsq(G, np.array([np.trace(O_.conj().T @ R) for O_ in Of]), rcond=None)[0]

    rng = np.random.default_rng(N * 10 + d)
    psi = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    phi = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    fp, ff = psi, phi
    for _ in range(N - 1):
        fp, ff = np.kron(fp, psi), np.kron(ff, phi)
    rho = np.outer(fp, ff.conj())
    x, v = to_pi(rho), bf.vec(rho)
    for _ in range(steps):
        x, v = P.apply(x), Lf @ v
        ref = to_pi(bf.unvec(v, Dh))
        assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()


def test_bruteforce_qubits_all_term_kinds():
    sp = L.spin_matrices(2)
    H = 1.0 * sp["jz"] + 0.7 * sp["jx"]
    tp = [L.Term("collective_h", H), L.Term("local_jump", sp["jm"], 0.1), L.Term("local_jump", sp["jz"], 0.2),
          L.Term("local_jump", sp["jp"], 0.02), L.Term("collective_jump", sp["jm"], 0.1),
          L.Term("collective_jump", sp["jz"], 0.03)]
    tf = [("collective_h", H, 1.0), ("local_jump", sp["jm"], 0.1), ("local_jump", sp["jz"], 0.2),
          ("local_jump", sp["jp"], 0.02), ("collective_jump", sp["jm"], 0.1), ("collective_jump", sp["jz"], 0.03)]
    _bruteforce_check(4, 2, tp, tf, 4)


def test_bruteforce_qutrits_with_direct_term():
    s3 = L.spin_matrices(3)
    N = 3
    b = L.Basis(N, 3)
    g = L.OneBodyGeometry(b)
    jz = L.collective_operator(g, s3["jz"])
    tp = [L.Term("local_h", 1.0 * s3["jz"] + 0.7 * s3["jx"]),
          L.Term("direct_h", L.physical_blocks(b, L.pi_product(b, jz, jz)), 0.3 / N),
          L.Term("local_jump", s3["jm"], 0.1)]
    Jzf = bf.collective(s3["jz"], N, 3)
    tf = [("local_h", 1.0 * s3["jz"] + 0.7 * s3["jx"], 1.0), ("full_h", Jzf @ Jzf, 0.3 / N),
          ("local_jump", s3["jm"], 0.1)]
    _bruteforce_check(N, 3, tp, tf, 3)


def _c2(N):
    sp = L.spin_matrices(2)
    return O.compile_plan(N, 2, L.qubit_ensemble_terms(hamiltonian=sp["jz"] + sp["jx"], emission=0.1,
                                                       dephasing=0.2, pumping=0.02, collective_emission=0.1))


def _hermitian_adjoint(b, x):
    out = np.empty_like(x)
    for s, m in enumerate(b.dims):
        o0, o1 = b.offsets[s], b.offsets[s + 1]
        out[o0:o1] = x[o0:o1].reshape(m, m, order="F").conj().T.reshape(-1, order="F")
    return out


@pytest.mark.parametrize("N", [6, 17])
def test_invariants_trace_and_hermiticity(N):
    P, _ = _c2(N)
    rng = np.random.default_rng(N)
    x = rng.standard_normal(P.n) + 1j * rng.standard_normal(P.n)
    y = P.apply(x)
    # trace preservation: t . (L x) = 0
    assert abs(np.dot(P.t_val, y[P.t_idx])) <= 1e-13 * np.linalg.norm(y) * np.linalg.norm(P.t_val)
    # Hermiticity preservation: L(x^dagger) = (L x)^dagger
    np.testing.assert_allclose(P.apply(_hermitian_adjoint(P.basis, x)), _hermitian_adjoint(P.basis, y),
                               rtol=0, atol=1e-12 * np.abs(y).max())


def test_threaded_apply_is_bitwise_serial():
    P, _ = _c2(30)
    x = np.random.default_rng(3).standard_normal(P.n) + 0j
    assert np.array_equal(P.apply(x), P.apply(x, nthreads=5))


def test_c1_dynamics_closed_form():
    gold = GOLD["c1"]
    P, g = O.compile_plan(8, 2, L.driven_qubits_terms(drive=1.0, detuning=1.0, emission=0.1, dephasing=0.2,
                                                      pumping=0.0))
    sp = L.spin_matrices(2)
    obs = [L.collective_operator(g, sp["jp"] @ sp["jm"]),
           L.pi_product(P.basis, L.collective_operator(g, sp["jp"]), L.collective_operator(g, sp["jm"]))]
    rho0 = L.computational_product_state(P.basis, 2)
    out, _, info = P.solve_dynamics(rho0, np.array(gold["t"]), obs, krylovdim=30, atol=0.0, rtol=1e-12)
    assert info["converged"]
    np.testing.assert_allclose(out[:, 0].real, gold["n"], rtol=1e-10, atol=0)
    np.testing.assert_allclose(out[:, 1].real, gold["JpJm"], rtol=1e-10, atol=0)
    x, sinfo = P.steady_state(krylovdim=30, maxiter=5000, atol=0.0, rtol=1e-12)
    assert sinfo["converged"]
    assert abs(np.vdot(obs[0], x).real - gold["steady_n"]) <= 1e-9 * gold["steady_n"]
    assert abs(np.vdot(obs[1], x).real - gold["steady_JpJm"]) <= 1e-9 * gold["steady_JpJm"]


def test_local_steady_closed_form_preconditioned():
    gold = GOLD["local_steady_20"]
    sp = L.spin_matrices(2)
    P, g = O.compile_plan(20, 2, L.qubit_ensemble_terms(hamiltonian=sp["jz"] + sp["jx"], emission=0.1,
                                                        dephasing=0.2, pumping=0.02))
    x, info = P.steady_state(krylovdim=300, maxiter=300, atol=0.0, rtol=1e-11, preconditioner="sector",
                             criterion="backward")
    assert info["converged"]
    n_op = L.collective_operator(g, sp["jp"] @ sp["jm"])
    jj = L.pi_product(P.basis, L.collective_operator(g, sp["jp"]), L.collective_operator(g, sp["jm"]))
    assert abs(np.vdot(n_op, x).real - gold["steady_n"]) <= 1e-8 * gold["steady_n"]
    assert abs(np.vdot(jj, x).real - gold["steady_JpJm"]) <= 1e-8 * gold["steady_JpJm"]


def test_steady_state_matches_dense_null_space():
    P, _ = _c2(6)
    n = P.n
    M = np.array([P.apply(np.eye(n, dtype=complex)[:, k]) for k in range(n)]).T
    w, V = np.linalg.eig(M)
    v = V[:, np.argmin(np.abs(w))]
    v = v / np.dot(P.t_val, v[P.t_idx])
    x, info = P.steady_state(krylovdim=30, maxiter=3000, atol=0.0, rtol=1e-12)
    assert info["converged"]
    assert np.linalg.norm(x - v) <= 1e-9 * np.linalg.norm(v)


def test_expv_matches_dense_exponential():
    P, _ = _c2(5)
    n = P.n
    M = np.array([P.apply(np.eye(n, dtype=complex)[:, k]) for k in range(n)]).T
    b = np.random.default_rng(2).standard_normal(n) + 0j
    y, info = P.expv(b, 2.5, krylovdim=20, atol=0.0, rtol=1e-12)
    assert info["converged"]
    ref = sla.expm(2.5 * M) @ b
    assert np.linalg.norm(y - ref) <= 1e-10 * np.linalg.norm(ref)


@pytest.mark.parametrize("scale", [1e-3, 0.3, 3.0, 40.0])
def test_small_expm_against_scipy(scale):
    A = scale * (np.random.default_rng(5).standard_normal((9, 9)) + 1j * np.random.default_rng(6).standard_normal((9, 9)))
    np.testing.assert_allclose(O.expm(A), sla.expm(A), rtol=0, atol=1e-12 * np.abs(sla.expm(A)).max())


def test_reference_gmres30_stalls_at_n20():
    """The reference's default restarted GMRES(30) does not reach 1e-8 on the
    N=20 member of the benchmark model family (it plateaus near 4e-5)."""
    P, _ = _c2(20)
    x, info = P.steady_state(krylovdim=30, maxiter=3000, atol=0.0, rtol=1e-8)
    assert not info["converged"]
    assert info["normalized_residual"] > 1e-6


def test_sector_weighting_pins_weakly_driven_observable():
    """Weakly driven local qubits (N=40, Omega=0.045): the steady state is a
    product state, so <sum_i n_i> = N sigma_ee in closed form.  Iterating on the
    PI coefficients, a residual 4e-11 relative to |x| still leaves <n> off by
    ~1e-4 (errors in sectors of large multiplicity f are amplified by
    sqrt(f) ~ 1e5 in traces); iterating on u = sqrt(f) x per sector reaches
    the closed form to ~1e-8 at the same tolerance."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_closed_form as CF
    N_, om, gphi = 40, 0.045, 0.071
    sp = L.spin_matrices(2)
    P, g = O.compile_plan(N_, 2, L.qubit_ensemble_terms(hamiltonian=2 * om * sp["jx"] + sp["jz"], emission=0.1,
                                                       dephasing=4 * gphi))
    exact = CF.steady(N_, 2 * om * CF.jx + CF.jz, [(0.1, CF.jm), (4 * gphi, CF.jz)])[0]
    nop = L.collective_operator(g, sp["jp"] @ sp["jm"])
    errs = {}
    for w in (None, "sector"):
        x, info = P.steady_state(krylovdim=400, maxiter=400, atol=0.0, rtol=1e-10, preconditioner="sector",
                                 criterion="backward", nthreads=4, weighting=w)
        assert info["converged"]
        assert info["normalized_residual"] <= 1e-10 * np.linalg.norm(x)
        errs[w] = abs(np.vdot(nop, x).real - exact) / exact
    assert errs["sector"] <= 1e-7, errs
    assert errs[None] >= 100 * errs["sector"], errs


def _small_model():
    sp = L.spin_matrices(2)
    return O.compile_plan(4, 2, L.qubit_ensemble_terms(hamiltonian=sp["jz"] + 0.7 * sp["jx"], emission=0.3,
                                                        dephasing=0.1, pumping=0.05, collective_emission=0.2))[0]


def test_rk4_matches_dense_exponential_at_fourth_order():
    """The oracle's RK4 (evolve!, src/evolution.jl:75-118) against expm(L t):
    accurate at fine steps, and the error falls 16x per halving of h (a wrong
    stage weight or stage point drops the order and fails the ratio)."""
    plan = _small_model()
    Ld = O.dense_liouvillian(plan)
    rng = np.random.default_rng(5)
    x0 = rng.standard_normal(plan.n) + 1j * rng.standard_normal(plan.n)
    exact = sla.expm(Ld * 1.5) @ x0
    errs = [np.linalg.norm(O.rk4(plan, x0, 0.0, 1.5, s) - exact) / np.linalg.norm(exact) for s in (40, 80, 640)]
    assert errs[2] <= 3e-10, errs
    assert 13.0 <= errs[0] / errs[1] <= 19.0, errs


def test_dense_liouvillian_spectrum_is_a_lindblad_spectrum():
    """Trace preservation gives one eigenvalue 0 (the steady state) and the
    rest decays: max real part 0, all others negative."""
    Ld = O.dense_liouvillian(_small_model())
    ev = np.linalg.eigvals(Ld)
    ev = ev[np.argsort(-ev.real)]
    assert abs(ev[0]) <= 1e-12 * np.abs(ev).max()
    assert ev[1].real < -1e-6


# ---------------------------------------------------------------------------
# operator scale (oracle_operator_scale, restating src/krylov.jl:156-169 with
# a splitmix64 + Box-Muller probe stream in place of Julia's MersenneTwister)
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def _splitmix64_stream(seed):
    s = seed
    while True:
        s = (s + 0x9E3779B97F4A7C15) & _M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        yield z ^ (z >> 31)


def _probe_vectors(n, count, seed=0x51CA1E):
    """The documented probe stream: element e of a probe uses two uniforms
    u, v in (0, 1) from consecutive outputs; z = sqrt(-2 log u) e^{2 pi i v} / sqrt(2)
    (a standard complex Gaussian)."""
    g = _splitmix64_stream(seed)
    out = []
    for _ in range(count):
        z = np.empty(n, dtype=complex)
        for e in range(n):
            u = ((next(g) >> 11) + 0.5) / 2.0 ** 53
            v = ((next(g) >> 11) + 0.5) / 2.0 ** 53
            z[e] = math.sqrt(-2.0 * math.log(u)) * complex(math.cos(2 * math.pi * v), math.sin(2 * math.pi * v))
        out.append(z / math.sqrt(2.0))
    return out


@pytest.mark.parametrize("N,model", [(4, "c2"), (6, "c2"), (10, "c2"), (8, "c1"), (3, "qutrit")])
def test_operator_scale_bounded_by_dense_norms(N, model):
    """s = max over three probes of |L z|/|z| is a lower estimate of the
    spectral norm: s <= |L|_2 for every probe, and for Gaussian probes
    E|L z|^2/|z|^2 ~ |L|_F^2 / n, so s >= |L|_F / (2 sqrt(n)).  The same
    number follows from the dense matrix and an independent implementation
    of the probe stream."""
    if model == "c2":
        P, _ = _c2(N)
    elif model == "c1":
        P, _ = O.compile_plan(8, 2, L.driven_qubits_terms(drive=1.0, detuning=1.0, emission=0.1, dephasing=0.2,
                                                          pumping=0.0))
    else:
        b = L.Basis(N, 3)
        g = L.OneBodyGeometry(b)
        s3 = L.spin_matrices(3)
        P = O.OraclePlan(L.compile_model(b, [L.Term("local_h", s3["jz"] + 0.7 * s3["jx"]),
                                             L.Term("local_jump", s3["jm"], 0.1)], g))
    M = O.dense_liouvillian(P)
    n = P.n
    s = P.operator_scale()
    two = np.linalg.norm(M, 2)
    fro = np.linalg.norm(M, "fro")
    assert s <= two * (1 + 1e-13), (s, two)
    assert s >= 0.5 * fro / math.sqrt(n), (s, fro / math.sqrt(n))
    # independent recomputation: dense products and the stream restated here
    ratios = [np.linalg.norm(M @ z) / np.linalg.norm(z) for z in _probe_vectors(n, 3)]
    assert abs(s - max(ratios)) <= 1e-13 * s, (s, ratios)
    # it scales with the operator: the same probes on 2.5 L
    P2s = max(np.linalg.norm(2.5 * M @ z) / np.linalg.norm(z) for z in _probe_vectors(n, 3))
    assert abs(P2s - 2.5 * s) <= 1e-13 * P2s


# ---------------------------------------------------------------------------
# sector preconditioner (OraclePlan.sector_preconditioner)
# ---------------------------------------------------------------------------
def _apply_sector_pc(P, x, scale=1.0):
    U, inv = P.sector_preconditioner()
    b = P.basis
    y = np.zeros_like(x)
    for s, m in enumerate(b.dims):
        o0, o1 = b.offsets[s], b.offsets[s + 1]
        Us = U[o0:o1].reshape(m, m, order="F")
        R = x[o0:o1].reshape(m, m, order="F")
        y[o0:o1] = (scale * Us @ ((Us.conj().T @ R @ Us) * inv[o0:o1].reshape(m, m, order="F"))
                    @ Us.conj().T).reshape(-1, order="F")
    return y


def _sector_G(P, s):
    m = P.basis.dims[s]
    H = P.plan.H[s].dense() if P.plan.H is not None else np.zeros((m, m), complex)
    Q = P.plan.Q[s].dense() if P.plan.Q is not None else np.zeros((m, m), complex)
    return -1j * H - 0.5 * Q


@pytest.mark.parametrize("N", [6, 11])
def test_sector_preconditioner_inverts_diagonal_part_exactly(N):
    """With H = sum jz and jumps whose L^dagger L are diagonal in the
    collective jz basis (local emission, dephasing, pumping, collective
    emission: J+J- = J^2 - Jz^2 + Jz), Q is diagonal in the eigenbasis of
    every fused H block, so the preconditioner is the exact inverse of the
    sector-diagonal map X -> G X + X G^H, G = -iH - Q/2 (checked against dense
    G built from the plan's fused blocks, on random X), except on the null
    directions where G X + X G^H vanishes."""
    sp = L.spin_matrices(2)
    P, _ = O.compile_plan(N, 2, L.qubit_ensemble_terms(hamiltonian=0.9 * sp["jz"], emission=0.1, dephasing=0.2,
                                                       pumping=0.02, collective_emission=0.1))
    rng = np.random.default_rng(N)
    x = rng.standard_normal(P.n) + 1j * rng.standard_normal(P.n)
    b = P.basis
    ax = np.zeros_like(x)
    for s, m in enumerate(b.dims):
        o0, o1 = b.offsets[s], b.offsets[s + 1]
        G = _sector_G(P, s)
        X = x[o0:o1].reshape(m, m, order="F")
        ax[o0:o1] = (G @ X + X @ G.conj().T).reshape(-1, order="F")
    back = _apply_sector_pc(P, ax)
    assert np.linalg.norm(back - x) <= 1e-12 * np.linalg.norm(x)
    # and the forward direction: A(P(x)) = x
    px = _apply_sector_pc(P, x)
    fwd = np.zeros_like(x)
    for s, m in enumerate(b.dims):
        o0, o1 = b.offsets[s], b.offsets[s + 1]
        G = _sector_G(P, s)
        X = px[o0:o1].reshape(m, m, order="F")
        fwd[o0:o1] = (G @ X + X @ G.conj().T).reshape(-1, order="F")
    assert np.linalg.norm(fwd - x) <= 1e-12 * np.linalg.norm(x)


def test_sector_preconditioner_approximates_drive_and_is_nonsingular():
    """With a transverse drive Q is no longer diagonal in the H eigenbasis;
    the preconditioner then inverts the map with Q replaced by its diagonal
    there, so it stays a close, invertible approximation: |P A x - x| small
    against |x| for the C2 model at N=8, and the eigenvalues of P A stay
    near 1."""
    P, _ = _c2(8)
    b = P.basis
    rng = np.random.default_rng(3)
    errs = []
    for s, m in enumerate(b.dims):
        G = _sector_G(P, s)
        A = np.kron(np.eye(m), G) + np.kron(G.conj(), np.eye(m))  # vec(G X + X G^H)
        Ppc = np.zeros((m * m, m * m), dtype=complex)
        e = np.zeros(P.n, dtype=complex)
        o0 = b.offsets[s]
        for k in range(m * m):
            e[:] = 0
            e[o0 + k] = 1
            Ppc[:, k] = _apply_sector_pc(P, e)[o0:o0 + m * m]
        ev = np.linalg.eigvals(Ppc @ A)
        # measured: |eigenvalues| in [0.72, 1.28], relative error <= 0.15 over the sectors
        assert 0.5 < np.min(np.abs(ev)) and np.max(np.abs(ev)) < 1.5, (s, ev)
        X = rng.standard_normal(m * m) + 1j * rng.standard_normal(m * m)
        errs.append(np.linalg.norm(Ppc @ (A @ X) - X) / np.linalg.norm(X))
    assert max(errs) < 0.3


def test_sector_diagonal_apply_is_the_block_diagonal_of_L():
    """oracle_apply_diag (the inner solver's D) equals the sector-block
    diagonal of the dense L built from the full apply (itself pinned by brute
    force): each sector's m^2 x m^2 diagonal block, and nothing else."""
    P, _ = _c2(7)
    M = O.dense_liouvillian(P)
    b = P.basis
    D = np.zeros_like(M)
    for s in range(len(b.dims)):
        o0, o1 = b.offsets[s], b.offsets[s + 1]
        D[o0:o1, o0:o1] = M[o0:o1, o0:o1]
    rng = np.random.default_rng(11)
    for _ in range(3):
        x = rng.standard_normal(P.n) + 1j * rng.standard_normal(P.n)
        y = P.apply_diag(x, nthreads=3)
        assert np.linalg.norm(y - D @ x) <= 1e-13 * np.linalg.norm(D @ x)
    assert np.linalg.norm(M - D) > 0.02 * np.linalg.norm(M)   # the off-diagonal gains (4 % here) are left out


@pytest.mark.parametrize("N,backward", [(12, True), (24, True), (8, False)])
def test_oracle_flexible_gmres_reaches_the_preconditioned_solution(N, backward):
    """The oracle's inner-outer flexible GMRES (the product's sector_inner
    solver) reaches the state of plain sector-preconditioned GMRES, in fewer
    outer steps, and the dense null space at small N."""
    P, _ = _c2(N)
    crit = "backward" if backward else "reference"
    x1, i1 = P.steady_state(krylovdim=400, maxiter=400, atol=0.0, rtol=1e-13, preconditioner="sector",
                            criterion=crit, nthreads=4)
    x2, i2 = P.steady_state(krylovdim=200, maxiter=400, atol=0.0, rtol=1e-13, preconditioner="sector_inner",
                            inner_steps=6, criterion=crit, nthreads=4)
    assert i1["converged"] and i2["converged"]
    assert i2["iterations"] < i1["iterations"]
    assert np.linalg.norm(x2 - x1) <= 1e-10 * np.linalg.norm(x1)
    if N <= 8:
        M = O.dense_liouvillian(P)
        w, V = np.linalg.eig(M)
        v = V[:, np.argmin(np.abs(w))]
        v = v / np.dot(P.t_val, v[P.t_idx])
        assert np.linalg.norm(x2 - v) <= 1e-10 * np.linalg.norm(v)
