"""FP64 oracle reference for the N=120 qubit stationary state (writes c2_steady_reference.json).

Model: qubit_ensemble_model(120; hamiltonian = jz + jx, emission 0.1, dephasing
0.2, pumping 0.02, collective_emission 0.1), PI dimension 302,621.  The
oracle's inner-outer flexible GMRES (8 sector-preconditioned inner steps,
Krylov 300) in sector-weighted coordinates u = sqrt(f^nu) x, accepted on the
weighted relative residual |S L x|/(s |S x|) <= 1e-15 (8 OpenMP threads),
gives the stationary state, normalised to trace one.  At this N the
coefficient-space residual |L x|/(s |x|) does not resolve the populated
sectors (their coefficients scale as 1/sqrt(f^nu), f^nu up to 1e34):
unweighted solves at |L x|/(s |x|) ~ 1e-12 give <sum_i n_i> anywhere from 25.6
to 37.1, while two different weighted solvers (this one and plain
sector-preconditioned GMRES) agree on 22.3743025625 to 1e-12.  Both residuals
are recorded.  The state itself (4.8 MB) is not stored; the fixture
holds what pins it without it:
  - collective observables <sum_i n_i>, <Jz>, <Jx>, <J+ J->;
  - the population of every sector, p_nu = sqrt(f^nu) * trace(block_nu);
  - |x|_2 and sixteen projections <r_k, x> on seeded complex Gaussian vectors
    (numpy PCG64(20261204), r_k drawn in order k = 0..15, real then imaginary
    parts of each as one standard_normal((2, n)) call).
Takes a few minutes on 8 cores.
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

N_ = 120
SEED = 20261204
NPROJ = 16


def terms():
    sp = L.spin_matrices(2)
    return L.qubit_ensemble_terms(hamiltonian=sp["jz"] + sp["jx"], emission=0.1, dephasing=0.2, pumping=0.02,
                                  collective_emission=0.1)


def projection_vectors(n):
    rng = np.random.default_rng(SEED)
    out = []
    for _ in range(NPROJ):
        z = rng.standard_normal((2, n))
        out.append(z[0] + 1j * z[1])
    return out


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


def summarize(basis, geom, tval, x):
    """The fixture's quantities for a trace-one state x (also used by the tests)."""
    sp = L.spin_matrices(2)
    jp = L.collective_operator(geom, sp["jp"])
    jm = L.collective_operator(geom, sp["jm"])
    obs = {"n": L.collective_operator(geom, sp["jp"] @ sp["jm"]), "jz": L.collective_operator(geom, sp["jz"]),
           "jx": L.collective_operator(geom, sp["jx"]), "JpJm": product_coefficients(basis, jp, jm, tval)}
    vals = {k: float(np.vdot(v, x).real) for k, v in obs.items()}
    pops = []
    k = 0
    for s, m in enumerate(basis.dims):
        blk = x[basis.offsets[s]:basis.offsets[s + 1]].reshape(m, m, order="F")
        pops.append(float((tval[k] * np.trace(blk)).real))
        k += m
    proj = [complex(np.vdot(r, x)) for r in projection_vectors(basis.n)]
    return {"observables": vals, "populations": pops, "norm": float(np.linalg.norm(x)),
            "projections_re": [p.real for p in proj], "projections_im": [p.imag for p in proj]}


def main():
    t0 = time.time()
    basis = L.Basis(N_, 2)
    geom = L.OneBodyGeometry(basis)
    ref = O.OraclePlan(L.compile_model(basis, terms(), geom))
    x, info = ref.steady_state(krylovdim=400, maxiter=800, atol=0.0, rtol=1e-15, preconditioner="sector_inner",
                               inner_steps=8, criterion="weighted", weighting="sector", nthreads=8)
    tr = complex(np.dot(ref.t_val, x[ref.t_idx]))
    x = x / tr
    scale = info["operator_scale"]
    Lx = ref.apply(x, nthreads=8)
    sig = ref.sector_sigma()
    rel = float(np.linalg.norm(Lx) / (scale * np.linalg.norm(x)))
    relw = float(np.linalg.norm(sig * Lx) / (scale * np.linalg.norm(sig * x)))
    out = {"N": N_, "n": int(ref.n), "converged": int(info["converged"]), "iterations": int(info["iterations"]),
           "operator_scale": scale, "relative_residual": rel, "weighted_relative_residual": relw, "seed": SEED}
    out.update(summarize(basis, geom, ref.t_val, x))
    out["seconds"] = time.time() - t0
    with open(os.path.join(HERE, "c2_steady_reference.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print({k: out[k] for k in ("converged", "iterations", "relative_residual", "weighted_relative_residual",
                               "observables", "seconds")})


if __name__ == "__main__":
    main()
