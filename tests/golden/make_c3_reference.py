"""FP64 oracle reference for the spin-1 precision workload (writes c3_reference.json).

Model: PIBasis(18, 3); LocalHamiltonian(1.0 Sz + 0.7 Sx), DirectPIHamiltonian
(Jz*Jz, rate 0.3/18), LocalJump(S-, rate 0.1); PI dimension 1,562,275.  The
oracle's sector-preconditioned GMRES (sector-weighted coordinates, backward
criterion, rtol 1e-12, 8 OpenMP threads) gives the stationary state; the
fixture holds <Jz>/N and <J+ J->/N^2 of it.  Takes tens of minutes on 8 cores.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import lowering as L  # noqa: E402
from oracle import oracle as O  # noqa: E402

N_, D = 18, 3


def terms(geom):
    """Jz^2 enters as its physical Schur blocks (Jz block squared per sector)."""
    s = L.spin_matrices(3)
    jz2 = []
    for q in range(len(geom.basis.dims)):
        K = L.collective_sparse_block(geom, s["jz"], q).dense()
        jz2.append(K @ K)
    return [L.Term("local_h", 1.0 * s["jz"] + 0.7 * s["jx"]),
            L.Term("direct_h", jz2, 0.3 / N_),
            L.Term("local_jump", s["jm"], 0.1)]


def product_coefficients(basis, a, b, tval):
    """Coefficient vector of (sum_i A_i)(sum_i B_i): per sector A_nu B_nu / sqrt(f)."""
    out = np.zeros(basis.n, dtype=complex)
    k = 0
    for s, m in enumerate(basis.dims):
        o0, o1 = basis.offsets[s], basis.offsets[s + 1]
        sf = tval[k]
        k += m
        out[o0:o1] = (a[o0:o1].reshape(m, m, order="F") @ b[o0:o1].reshape(m, m, order="F") / sf).reshape(
            -1, order="F")
    return out


def main():
    t0 = time.time()
    basis = L.Basis(N_, D)
    geom = L.OneBodyGeometry(basis)
    ref = O.OraclePlan(L.compile_model(basis, terms(geom), geom))
    s = L.spin_matrices(3)
    x, info = ref.steady_state(krylovdim=400, maxiter=2000, atol=0.0, rtol=1e-12, preconditioner="sector",
                               criterion="backward", weighting="sector", nthreads=8)
    jz = L.collective_operator(geom, s["jz"])
    jp = L.collective_operator(geom, s["jp"])
    jm = L.collective_operator(geom, s["jm"])
    jpjm = product_coefficients(ref.basis, jp, jm, ref.t_val)
    out = {"N": N_, "d": D, "n": int(ref.n), "converged": int(info["converged"]),
           "iterations": int(info["iterations"]),
           "relative_residual": info["normalized_residual"] / float(np.linalg.norm(x)),
           "jz_over_N": float(np.vdot(jz, x).real / N_), "jpjm_over_N2": float(np.vdot(jpjm, x).real / N_ ** 2),
           "trace": float(np.dot(ref.t_val, x[ref.t_idx]).real), "seconds": time.time() - t0}
    with open(os.path.join(HERE, "c3_reference.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(out)


if __name__ == "__main__":
    main()
