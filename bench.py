"""Benchmark: steady-state solves/s of the PI Liouvillian on one MI355X.

Workload (one step = one stationary solve): the N=120 qubit ensemble,
H = sum_i (jz + jx), local emission 0.1, local dephasing 0.2 (D[jz]), local
pumping 0.02, collective emission 0.1; PI dimension C(123,3) = 302,621.
Complex FP64 throughout: flexible GMRES on the device (outer Krylov
dimension 200) whose right preconditioner is 8 steps of sector-preconditioned
GMRES on the sector-block-diagonal part of L, solved from rho0 = t/(t.t) to
relative residual |L rho|/(s |rho|) <= 1e-8.  One solve of plain
sector-preconditioned GMRES (Krylov dimension 1000) and one with FP32 inner
solves are timed beside it.
Inputs are synthesized on the host and resident in HBM before the timed
region; each solve's result is copied back (included in the timing).

Also reported: p50/p99 solve latency, the apply kernel against the HBM
roofline (algorithmic bytes per apply / event-timed average apply), and the
oracle's OpenMP CPU restatement timed on a bounded sample on the host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "permutationalinvariantdynamics.jl_amd"))
sys.path.insert(0, ROOT)

import ctypes as C  # noqa: E402

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
N_QUBITS = 120
RTOL = 1e-8
KRYLOV = 200          # outer (flexible) Krylov dimension of the headline solver
INNER_STEPS = 8
KRYLOV_PLAIN = 1000   # plain sector-preconditioned GMRES, no restart needed


def build_model(pimi):
    sp = pimi.spin_matrices(2)
    return pimi.qubit_ensemble_model(N_QUBITS, hamiltonian=sp.jz + sp.jx, emission=0.1, dephasing=0.2,
                                     pumping=0.02, collective_emission=0.1)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE runs (calibrated by a device copy), or None."""
    path = os.path.join(ROOT, "profiles", "k_apply_traffic_n120_fp64.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
    except OSError:
        return None
    return d["traffic_bytes"] if d.get("kernel") == kernel else None


def apply_roofline(pimi, comp, reps=200):
    N = pimi._native
    ws = comp.workspace(pimi.PIMI_C64, 1)
    ms = C.c_double()
    N.check(N.lib().pimi_apply_bench(ws, C.c_int64(1), C.c_int32(reps), C.byref(ms)))
    n = comp.info["n"]
    algorithmic = 2 * n * 16 + comp.info["table_bytes_c64"]
    achieved = algorithmic / (ms.value * 1e-3) / 1e9
    return ms.value, algorithmic, achieved


def cpu_baseline(steady_iters, budget_s):
    """Oracle (C/OpenMP restatement: output-sector-owned threaded apply, the
    same sector preconditioner, MGS2 GMRES) on a bounded sample of the same
    solve: runs of k1 and k2 Arnoldi steps fit t(k) = a k + b k^2 (apply and
    preconditioner per step, Gram-Schmidt growing with k), evaluated at the
    device solve's iteration count."""
    from oracle import lowering as L
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 64)
    sp = L.spin_matrices(2)
    ref, _ = O.compile_plan(N_QUBITS, 2, L.qubit_ensemble_terms(
        hamiltonian=sp["jz"] + sp["jx"], emission=0.1, dephasing=0.2, pumping=0.02, collective_emission=0.1))
    scale = ref.operator_scale(threads)
    ref.sector_preconditioner()

    def run(k):
        t0 = time.perf_counter()
        ref.steady_state(krylovdim=k, maxiter=k, atol=0.0, rtol=RTOL, operator_scale=scale, nthreads=threads,
                         preconditioner="sector", criterion="backward")
        return time.perf_counter() - t0

    k1 = 4
    t1 = run(k1)
    k2 = int(max(8, min(200, budget_s / max(t1 / k1, 1e-3) / 2)))
    t2 = run(k2)
    b = max((t2 / k2 - t1 / k1) / (k2 - k1), 0.0)
    a = max(t1 / k1 - b * k1, 0.0)
    est = a * steady_iters + b * steady_iters ** 2
    return {"value": 1.0 / est, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": (f"same preconditioned GMRES at N=120 on {threads} OpenMP threads: {k1} steps in {t1:.2f} s and "
                       f"{k2} steps in {t2:.2f} s, fit t(k)={a * 1e3:.1f} ms*k + {b * 1e3:.3f} ms*k^2, evaluated at "
                       f"the plain device solve's {steady_iters} steps = {est:.1f} s per solve")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--profile", action="store_true", help="apply + copy calibration only (counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import pimi
    comp = pimi.compile(build_model(pimi), backend="mi355x", device=0)
    if args.profile:
        # counter runs: 20 event-timed applies, then 20 device copies of 64 MiB
        # (the calibration kernel for the byte counters)
        N = pimi._native
        ms_apply, bytes_apply, achieved = apply_roofline(pimi, comp, reps=20)
        ms = C.c_double()
        N.check(N.lib().pimi_copy_bench(comp.workspace(pimi.PIMI_C64, 1), C.c_int64(64 << 20),
                                        C.c_int32(20), C.byref(ms)))
        print(json.dumps({"apply_ms": ms_apply, "apply_algorithmic_bytes": bytes_apply, "apply_GBps": achieved,
                          "copy_bytes": 64 << 20, "copy_ms": ms.value,
                          "copy_GBps": 2 * (64 << 20) / (ms.value * 1e-3) / 1e9}))
        return
    alg = pimi.GMRESAlgorithm(krylovdim=KRYLOV, maxiter=10 * KRYLOV, preconditioner="sector_inner",
                              inner_steps=INNER_STEPS)

    def solve(algorithm=alg, precision=pimi.PIMI_C64):
        return pimi.stationary_state(comp, algorithm=algorithm, atol=0.0, rtol=RTOL, return_info=True,
                                     criterion="backward", precision=precision)

    for _ in range(args.warmup):
        solve()
    lat = []
    infos = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        r = solve()
        lat.append(time.perf_counter() - s)
        infos.append(r.info)
        print(f"# solve {len(lat)}: {lat[-1]:.3f} s, iterations {r.info['iterations']}, "
              f"relative residual {r.info['relative_residual']:.3e}", file=sys.stderr, flush=True)
    total = time.perf_counter() - t0
    iters = int(np.median([i["iterations"] for i in infos]))
    side = {}
    for name, algorithm, prec in (
            ("plain_fp64", pimi.GMRESAlgorithm(krylovdim=KRYLOV_PLAIN, maxiter=KRYLOV_PLAIN,
                                               preconditioner="sector"), pimi.PIMI_C64),
            ("fp32_inner", alg, pimi.PIMI_C32)):
        s = time.perf_counter()
        r = solve(algorithm, prec)
        side[name] = {"ms": (time.perf_counter() - s) * 1e3, "iterations": int(r.info["iterations"]),
                      "relative_residual": r.info["relative_residual"]}
        print(f"# {name}: {side[name]}", file=sys.stderr, flush=True)
    ms_apply, bytes_apply, achieved = apply_roofline(pimi, comp)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(side["plain_fp64"]["iterations"], args.cpu_seconds)
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "steady-state solves/s & p99 latency at resid 1e-8; SpMV GB/s vs HBM peak",
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": "qubit ensemble N=120 (PI dim 302621): H=sum(jz+jx), local emission 0.1, "
                               "dephasing 0.2, pumping 0.02, collective emission 0.1; FP64 sector-"
                               "preconditioned flexible GMRES (8 inner steps) stationary state to |L rho|/(s|rho|) <= 1e-8",
                   "n": comp.info["n"], "krylovdim": KRYLOV, "inner_steps": INNER_STEPS, "rtol": RTOL},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "gmres_iterations": iters,
        "relative_residual": max(i["relative_residual"] for i in infos),
        "trace_error": max(i["trace_error"] for i in infos),
        "plain_fp64_sector_gmres": side["plain_fp64"],
        "fp32_inner": side["fp32_inner"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("pimi::k_apply<double>"),
                     "kernel": "k_apply<double>", "avg_ms": ms_apply, "algorithmic_bytes": bytes_apply},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
