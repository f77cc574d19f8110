"""iid_state (src/operators.jl:146-250) on the host: the PI coefficients of
rho^{otimes N} from the one-box Clebsch-Gordan recurrence in the native
lowering (pimi_iid_state), against

* brute force: the full d^N-dimensional rho^{otimes N}, linked to PI
  coordinates by pairings Tr(O rho) with products of collective units (a
  spanning set by Schur-Weyl duality), for qubits and qutrits;
* closed forms at large N: trace one, <sum_i n_i> = N rho_ee and
  <J+ J-> = N rho_ee + N(N-1)|rho_ge|^2;
* computational_product_state (the symmetric-sector formula of
  src/spin.jl:79-125) for pure basis states;
* the reference's validation (Hermitian, positive, trace one).
"""
import math

import numpy as np
import pytest

import bruteforce as bf
import pimi
from oracle import lowering as L


def _random_density(d, seed, rank=None):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((d, rank or d)) + 1j * rng.standard_normal((d, rank or d))
    rho = G @ G.conj().T
    return rho / np.trace(rho).real


@pytest.mark.parametrize("N,d,degree", [(3, 2, 3), (4, 2, 4), (3, 3, 3)])
def test_iid_state_against_full_tensor_power(N, d, degree):
    b = L.Basis(N, d)
    g = L.OneBodyGeometry(b)
    coll_pi = {(i, j): L.collective_operator(g, bf.unit(d, i, j)) for i in range(d) for j in range(d)}
    coll_f = {(i, j): bf.collective(bf.unit(d, i, j), N, d) for i in range(d) for j in range(d)}
    ident = L.collective_operator(g, np.eye(d) / N)
    Dh = d ** N
    Opi, Of = [], []
    for w in bf.collective_monomials(d, degree):
        a, A = ident.copy(), np.eye(Dh, dtype=complex)
        for ij in w:
            a = L.pi_product(b, a, coll_pi[ij])
            A = A @ coll_f[ij]
        Opi.append(a)
        Of.append(A)
    G = np.array(Opi).conj()
    assert np.linalg.matrix_rank(G, 1e-9) == b.n
    for seed, rank in ((1, None), (2, 1)):
        rho = _random_density(d, seed + 10 * N + d, rank)
        full = rho
        for _ in range(N - 1):
            full = np.kron(full, rho)
        want = np.linalg.lstsq(G, np.array([np.trace(O_.conj().T @ full) for O_ in Of]), rcond=None)[0]
        got = pimi.iid_state(pimi.PIBasis(N, d), rho).data
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), np.abs(got - want).max()


@pytest.mark.parametrize("N", [1, 7, 40, 120])
def test_iid_state_closed_forms(N):
    rho = np.array([[0.3, 0.2 - 0.1j], [0.2 + 0.1j, 0.7]])  # levels (g, e): rho_ee = 0.7
    sp = pimi.spin_matrices(2)
    comp = pimi.compile(pimi.PIModel(pimi.PIBasis(N, 2), []), host_only=True)
    x = pimi.iid_state(comp.basis, rho)
    idx, tv = comp.trace_functional()
    assert abs(np.dot(tv, x.data[idx]) - 1.0) <= 1e-12
    n = pimi.expect(comp.collective_operator(np.array([[0, 0], [0, 1]])), x).real
    assert abs(n - N * 0.7) <= 1e-12 * N
    jp = comp.collective_operator(sp.jp).data
    jm = comp.collective_operator(sp.jm).data
    jj = 0.0
    k = 0
    for s, m in enumerate(comp.basis.dims):  # (A B) coefficient blocks = A_nu B_nu / sqrt(f)
        o0, o1 = comp.basis.offsets[s], comp.basis.offsets[s + 1]
        blk = (jp[o0:o1].reshape(m, m, order="F") @ jm[o0:o1].reshape(m, m, order="F")) / tv[k]
        jj += np.vdot(blk.reshape(-1, order="F"), x.data[o0:o1]).real
        k += m
    want = N * 0.7 + N * (N - 1) * abs(rho[0, 1]) ** 2
    assert abs(jj - want) <= 1e-11 * want, (jj, want)


@pytest.mark.parametrize("N,d", [(5, 2), (30, 2), (4, 3)])
def test_iid_state_of_basis_states_is_the_computational_product_state(N, d):
    b = pimi.PIBasis(N, d)
    for level in range(1, d + 1):
        e = np.zeros((d, d))
        e[level - 1, level - 1] = 1.0
        np.testing.assert_allclose(pimi.iid_state(b, e).data, pimi.computational_product_state(b, level).data,
                                   rtol=0, atol=1e-13)


def test_iid_state_validation():
    b = pimi.PIBasis(3, 2)
    with pytest.raises(ValueError):
        pimi.iid_state(b, np.array([[0.5, 0.1], [0.3, 0.5]]))      # not Hermitian
    with pytest.raises(ValueError):
        pimi.iid_state(b, np.array([[1.2, 0.0], [0.0, -0.2]]))     # not positive
    with pytest.raises(ValueError):
        pimi.iid_state(b, np.array([[0.6, 0.0], [0.0, 0.6]]))      # trace 1.2
    with pytest.raises(ValueError):
        pimi.iid_state(b, np.eye(3) / 3)                            # wrong size
    # a restricted basis must keep every sector the state occupies
    with pytest.raises(ValueError):
        pimi.iid_state(pimi.PIBasis(3, 2, sectors=[(3, 0)]), np.eye(2) / 2)
    sym = pimi.iid_state(pimi.PIBasis(3, 2, sectors=[(3, 0)]), np.array([[0.0, 0.0], [0.0, 1.0]]))
    assert abs(sym.data.sum() - 1.0) <= 1e-14
