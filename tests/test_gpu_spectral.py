"""Arnoldi spectra (krylov_liouvillian_spectrum, src/krylov.jl:914-961) and
fixed-step RK4 evolution (evolve!, src/evolution.jl:75-118) on the device,
against the oracle: dense eigenvalues of the oracle's Liouvillian, the true
residuals of the returned Ritz pairs, and the oracle's RK4 restatement
(itself pinned to expm at fourth order).  The random start vector is numpy's
(Julia's RNG stream is not reproducible here), so iteration counts are not
compared."""
import numpy as np
import pytest

import pimi
from oracle import lowering as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu
SP2 = pimi.spin_matrices(2)


def _qubits(N_):
    model = pimi.qubit_ensemble_model(N_, hamiltonian=SP2.jz + SP2.jx, emission=0.1, dephasing=0.2, pumping=0.02,
                                      collective_emission=0.1)
    sp = L.spin_matrices(2)
    ref, _ = O.compile_plan(N_, 2, L.qubit_ensemble_terms(hamiltonian=sp["jz"] + sp["jx"], emission=0.1,
                                                          dephasing=0.2, pumping=0.02, collective_emission=0.1))
    return pimi.compile(model), ref


def _qutrits(N_):
    S = pimi.spin_matrices(3)
    model = pimi.PIModel(pimi.PIBasis(N_, 3), [pimi.LocalHamiltonian(S.jz + 0.7 * S.jx), pimi.LocalJump(S.jm, 0.1)])
    sp = L.spin_matrices(3)
    ref, _ = O.compile_plan(N_, 3, [L.Term("local_h", sp["jz"] + 0.7 * sp["jx"]), L.Term("local_jump", sp["jm"], 0.1)])
    return pimi.compile(model), ref


@pytest.mark.parametrize("which", ["LR", "LM", "SM"])
def test_full_krylov_space_gives_dense_eigenvalues(which):
    """krylovdim = n: the Ritz values are eigenvalues of L."""
    print(f"spectrum full space which={which}", flush=True)
    comp, ref = _qubits(6)
    n = len(comp.basis)
    info = pimi.krylov_liouvillian_spectrum(comp, nev=6, krylovdim=n, which=which, rng=3,
                                            require_convergence=False)
    dense = np.linalg.eigvals(O.dense_liouvillian(ref))
    scale = np.abs(dense).max()
    for lam in info["values"]:
        assert np.abs(dense - lam).min() <= 1e-9 * scale, lam
    if which == "LR":
        assert abs(info["values"][0]) <= 1e-9 * scale  # the steady state
    assert info["dimension"] == n and info["krylov_dimension"] == n


def test_ritz_pairs_residuals_are_true_residuals():
    print("spectrum Ritz residuals", flush=True)
    comp, ref = _qubits(10)
    info = pimi.krylov_liouvillian_spectrum(comp, nev=4, krylovdim=60, vectors=True, rng=7,
                                            require_convergence=False)
    V = info["vectors"]
    for c, lam in enumerate(info["values"]):
        v = V[:, c]
        true = np.linalg.norm(ref.apply(v) - lam * v) / np.linalg.norm(v)
        assert abs(true - info["residuals"][c]) <= 1e-8 * info["residual_scale"] + 1e-12, (true, info["residuals"][c])
    # converged values are eigenvalues of the dense operator
    dense = np.linalg.eigvals(O.dense_liouvillian(ref))
    for lam, ok in zip(info["values"], info["converged"]):
        if ok:
            assert np.abs(dense - lam).min() <= 1e-6 * info["residual_scale"]


def test_qutrit_spectrum_and_errors():
    print("spectrum qutrits", flush=True)
    comp, ref = _qutrits(3)
    n = len(comp.basis)
    info = pimi.krylov_liouvillian_spectrum(comp, nev=3, krylovdim=n, rng=1, require_convergence=False)
    dense = np.linalg.eigvals(O.dense_liouvillian(ref))
    for lam in info["values"]:
        assert np.abs(dense - lam).min() <= 1e-9 * np.abs(dense).max()
    with pytest.raises(ValueError):
        pimi.krylov_liouvillian_spectrum(comp, nev=0)
    with pytest.raises(ValueError):
        pimi.krylov_liouvillian_spectrum(comp, which="SR")
    with pytest.raises(pimi.PIMIError):
        pimi.krylov_liouvillian_spectrum(comp, initial_vector=np.zeros(n))
    with pytest.raises(ValueError):  # too small a space for the requested accuracy
        pimi.krylov_liouvillian_spectrum(comp, nev=4, krylovdim=5, rng=2)


@pytest.mark.parametrize("N_", [8, 20])
def test_rk4_matches_oracle_rk4(N_):
    print(f"rk4 N={N_}", flush=True)
    comp, ref = _qubits(N_)
    rho0 = pimi.computational_product_state(comp.basis, 2)
    out = pimi.time_evolve(comp, rho0, (0.0, 2.0), steps=80)
    want = O.rk4(ref, rho0.data, 0.0, 2.0, 80)
    err = np.linalg.norm(out.data - want) / np.linalg.norm(want)
    print(f"  rel err {err:.2e}", flush=True)
    assert err <= 1e-12
    # exact in-place use, as evolve!(rho, L, rho, tspan)
    st = pimi.PIState(comp.basis, rho0.data.copy())
    pimi.evolve(st, comp, st, (0.0, 2.0), steps=80)
    np.testing.assert_array_equal(st.data, out.data)


def test_rk4_tracks_the_exponential_and_rejects_bad_input():
    print("rk4 vs expm", flush=True)
    comp, ref = _qubits(5)
    rho0 = pimi.computational_product_state(comp.basis, 2)
    Ld = O.dense_liouvillian(ref)
    exact = O.expm(Ld * 1.0) @ rho0.data
    out = pimi.time_evolve(comp, rho0, (0.0, 1.0), steps=400)
    assert np.linalg.norm(out.data - exact) / np.linalg.norm(exact) <= 1e-9
    with pytest.raises(ValueError):
        pimi.time_evolve(comp, rho0, (0.0, 1.0), steps=0)
    other = pimi.computational_product_state(pimi.PIBasis(4, 2), 2)
    with pytest.raises(ValueError):
        pimi.time_evolve(comp, other, (0.0, 1.0))
