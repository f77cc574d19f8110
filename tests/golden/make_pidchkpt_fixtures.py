"""Golden fixtures in the reference's portable checkpoint layout (PIDCHKPT v1,
src/checkpoints.jl:114-138), for exchanging states with a Julia-equipped run:
the oracle's C1 stationary state and its C1 state at t = 4.

C1: driven_qubits(8; drive=1, detuning=1, emission=0.1, dephasing=0.2,
pumping=0) (src/model_zoo.jl:214-236), PI dimension 165.
- c1_steady_state.pid: full (un-restarted) GMRES on the bordered operator
  (krylovdim 165, rtol 1e-14), trace one; no time.
- c1_expv_t4.pid: the adaptive Krylov exponential from |e><e|^8 over
  t in [0, 4] saved every 0.1 (Krylov 30, rtol 1e-14); time 4.0.

Both are written by the oracle on one thread (deterministic) through
pimi.save_checkpoint; run from the repository root:
    python tests/golden/make_pidchkpt_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "permutationalinvariantdynamics.jl_amd"))

import pimi  # noqa: E402
from oracle import lowering as L  # noqa: E402
from oracle import oracle as O  # noqa: E402

C1_PARAMS = dict(drive=1.0, detuning=1.0, emission=0.1, dephasing=0.2, pumping=0.0)


def c1_states():
    """(steady state, steady info, t = 4 state, expv times) from the oracle."""
    plan, geom = O.compile_plan(8, 2, L.driven_qubits_terms(**C1_PARAMS))
    x, info = plan.steady_state(krylovdim=plan.n, maxiter=plan.n, atol=0.0, rtol=1e-14, nthreads=1)
    basis = pimi.PIBasis(8, 2)
    rho0 = pimi.computational_product_state(basis, 2).data
    ts = np.asarray(pimi.api.saved_times((0.0, 4.0), 0.1), dtype=np.float64)  # Julia's 0:0.1:4
    sp = L.spin_matrices(2)
    n_op = L.collective_operator(geom, sp["jp"] @ sp["jm"])
    _, rho_t, _ = plan.solve_dynamics(rho0, ts, [n_op], krylovdim=30, atol=0.0, rtol=1e-14, nthreads=1)
    return x, info, rho_t, ts


def main():
    x, info, rho_t, ts = c1_states()
    basis = pimi.PIBasis(8, 2)
    meta = {"model": "driven_qubits(8; drive=1.0, detuning=1.0, emission=0.1, dephasing=0.2, pumping=0.0)",
            "writer": "oracle (tests/golden/make_pidchkpt_fixtures.py)"}
    pimi.save_checkpoint(os.path.join(HERE, "c1_steady_state.pid"), pimi.PIState(basis, x),
                         metadata={**meta, "solver": "GMRES krylovdim 165, rtol 1e-14",
                                   "normalized_residual": f"{info['normalized_residual']:.3e}"})
    pimi.save_checkpoint(os.path.join(HERE, "c1_expv_t4.pid"), pimi.PIState(basis, rho_t), time=float(ts[-1]),
                         metadata={**meta, "solver": "expv krylovdim 30, rtol 1e-14, saveat 0.1",
                                   "initial_state": "computational_product_state(basis, 2)"})


if __name__ == "__main__":
    main()
