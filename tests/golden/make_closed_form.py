"""Closed-form fixtures for local-only qubit models (writes closed_form.json).

For a model with only one-body Hamiltonian and local jumps, the state stays a
product state: rho(t) = sigma(t)^{otimes N} with sigma(t) = exp(t L1) sigma(0)
for the 4x4 single-qubit Lindbladian L1.  Collective observables follow:
  <sum_i n_i>  = N sigma_ee,
  <J+ J->      = N sigma_ee + N (N-1) |sigma_ge|^2.
Levels are ordered (g, e); jz = diag(-1/2, 1/2), jm = |g><e|.  Nothing here
uses the PI machinery, so these numbers pin it from outside.
"""
import json
import os

import numpy as np
from scipy.linalg import expm, null_space

jz = np.diag([-0.5, 0.5]).astype(complex)
jm = np.array([[0, 1], [0, 0]], dtype=complex)
jp = jm.conj().T
jx = 0.5 * (jp + jm)
I2 = np.eye(2)


def lindblad1(H, jumps):
    L = -1j * (np.kron(I2, H) - np.kron(H.T, I2))
    for rate, A in jumps:
        AdA = A.conj().T @ A
        L += rate * (np.kron(A.conj(), A) - 0.5 * (np.kron(I2, AdA) + np.kron(AdA.T, I2)))
    return L


def observables(sigma, N):
    see = sigma[1, 1].real
    sge = sigma[0, 1]
    return N * see, N * see + N * (N - 1) * abs(sge) ** 2


def series(N, H, jumps, sigma0, ts):
    L1 = lindblad1(H, jumps)
    out = []
    for t in ts:
        s = (expm(t * L1) @ sigma0.reshape(-1, order="F")).reshape(2, 2, order="F")
        out.append(observables(s, N))
    return np.array(out)


def steady(N, H, jumps):
    L1 = lindblad1(H, jumps)
    v = null_space(L1)[:, 0]
    s = v.reshape(2, 2, order="F")
    s = s / np.trace(s)
    return observables(s, N)


def main():
    e = np.array([[0, 0], [0, 1]], dtype=complex)
    out = {}
    # driven_qubits(8; drive=1, detuning=1, emission=0.1, dephasing=0.2, pumping=0), rho0 = |e><e|^8
    H = 1.0 * jx + 1.0 * jz
    jumps = [(0.1, jm), (0.2, jz)]
    ts = [round(0.1 * k, 10) for k in range(41)]
    ser = series(8, H, jumps, e, ts)
    out["c1"] = {"N": 8, "t": ts, "n": ser[:, 0].tolist(), "JpJm": ser[:, 1].tolist(),
                 "steady_n": steady(8, H, jumps)[0], "steady_JpJm": steady(8, H, jumps)[1]}
    # qubit_ensemble_model(N; H=jz+jx, emission 0.1, dephasing 0.2, pumping 0.02), steady state
    H2 = jz + jx
    jumps2 = [(0.1, jm), (0.2, jz), (0.02, jp)]
    for N in (20, 120):
        n_ss, jj_ss = steady(N, H2, jumps2)
        out[f"local_steady_{N}"] = {"N": N, "steady_n": n_ss, "steady_JpJm": jj_ss}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "closed_form.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
