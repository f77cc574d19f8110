"""Batched stationary solves and dynamics on one device (libpimi's batch engine).

Many independent models (different N, rates, fields) are solved together:
the engine admits queued jobs while its device-memory reservation holds
their Krylov state, advances all admitted jobs one step per tick with one
kernel launch per operation, and frees a job's memory as soon as it
finishes.  A stationary job is the solve of src/krylov.jl:797-891 with the
sector preconditioner and the scale-invariant stopping rule, run as GMRES
cycles inside FP64 iterative refinement (complex FP32 Arnoldi with
``precision=PIMI_C32``).  A dynamics job (FP64 engines) is solve_dynamics
with ExpvAlgorithm (src/highlevel.jl:559-733): its Arnoldi steps,
projections and observations share the ticks' launches with the other
jobs.  The reference solves one model at a time; this module has no
counterpart there.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .api import (CompiledPIModel, DynamicsStreamResult, ExpvAlgorithm, PIState, SteadyStateResult, _context,
                  _dynamics_inputs)


@dataclass
class BatchAlgorithm:
    """Per-job solver options: Krylov dimension of one GMRES cycle, residual
    reduction per cycle (inner_rtol), total Arnoldi steps and cycles, and the
    iteration coordinates (weighting="sector": u = sqrt(f^nu) x per sector,
    None: the PI coefficients)."""
    krylovdim: int = 200
    inner_rtol: float = 1e-4
    maxiter: int = 20000
    max_refinements: int = 50
    weighting: str | None = "sector"

    def __post_init__(self):
        if self.weighting not in (None, "sector"):
            raise ValueError("weighting must be None or \"sector\"")
        if not (1 <= self.krylovdim <= 4000):
            raise ValueError("krylovdim must lie in 1..4000")
        if not (0 < self.inner_rtol < 1):
            raise ValueError("inner_rtol must lie in (0, 1)")
        if self.maxiter < 1 or self.max_refinements < 1:
            raise ValueError("maxiter and max_refinements must be positive")


class BatchEngine:
    def __init__(self, precision=N.PIMI_C64, max_jobs=0, memory_bytes=0, device=0, timing=False):
        self.device = device
        self.precision = precision
        opts = N.BatchOpts(int(precision), int(max_jobs), int(memory_bytes), N.BATCH_TIMING if timing else 0, 0)
        h = C.c_void_p()
        N.check(N.lib().pimi_batch_create(_context(device), C.byref(opts), C.byref(h)))
        self._h = h
        self._jobs = {}   # job id -> compiled model (kept alive while the job exists)

    def close(self):
        if self._h:
            N.lib().pimi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def job_bytes(self, compiled: CompiledPIModel, krylovdim: int) -> int:
        b = C.c_int64()
        N.check(N.lib().pimi_batch_job_bytes(compiled._plan, C.c_int32(self.precision), C.c_int32(krylovdim),
                                             C.byref(b)))
        return b.value

    def submit(self, compiled: CompiledPIModel, algorithm: BatchAlgorithm | None = None, atol=0.0, rtol=1e-8,
               operator_scale=None) -> int:
        if compiled.device != self.device:
            raise ValueError("the compiled model lives on another device")
        a = algorithm or BatchAlgorithm()
        opts = N.GmresOpts(a.krylovdim, N.PRECOND_SECTOR, a.maxiter, float(atol), float(rtol),
                           float(operator_scale or 0.0), float(a.inner_rtol), int(a.max_refinements),
                           N.CRITERION_BACKWARD, 0, N.WEIGHT_SECTOR if a.weighting == "sector" else N.WEIGHT_NONE,
                           0)
        jid = C.c_int64()
        N.check(N.lib().pimi_batch_submit(self._h, compiled._plan, C.byref(opts), C.byref(jid)))
        self._jobs[jid.value] = (compiled, a)
        return jid.value

    def submit_dynamics(self, compiled: CompiledPIModel, rho0: PIState, tspan, saveat=None, algorithm=None,
                        observables=None) -> int:
        """Queue solve_dynamics(compiled, rho0, tspan; saveat, ExpvAlgorithm,
        observables) as a job of this (FP64) engine; fetch_dynamics returns it."""
        if compiled.device != self.device:
            raise ValueError("the compiled model lives on another device")
        algorithm = algorithm or ExpvAlgorithm()
        if not observables:
            raise ValueError("a batched dynamics job requires at least one observable")
        ts, names, obs, opts = _dynamics_inputs(compiled, rho0, tspan, saveat, algorithm, observables)
        rho = np.ascontiguousarray(rho0.data, dtype=np.complex128)
        tsc = np.ascontiguousarray(ts, dtype=np.float64)
        jid = C.c_int64()
        N.check(N.lib().pimi_batch_submit_expv(self._h, compiled._plan, N.ptr(rho), N.ptr(tsc), C.c_int64(len(tsc)),
                                               C.byref(opts), C.c_int32(len(names)), N.ptr(obs), C.byref(jid)))
        self._jobs[jid.value] = (compiled, ("dynamics", ts, names))
        return jid.value

    def fetch_dynamics(self, job: int, check=True) -> DynamicsStreamResult:
        """Result of a finished dynamics job; raises NotConverged (like
        solve_dynamics) unless check is False."""
        compiled, (_, ts, names) = self._jobs[job]
        out = np.zeros((len(ts), len(names)), dtype=np.complex128)
        final = np.zeros(len(compiled.basis), dtype=np.complex128)
        info = N.SolveInfo()
        st = N.lib().pimi_batch_fetch_expv(self._h, C.c_int64(job), N.ptr(out), N.ptr(final), C.byref(info))
        if st not in (N.PIMI_OK, N.PIMI_ENOTCONVERGED):
            N.check(st)
        del self._jobs[job]
        res = DynamicsStreamResult(ts, {name: out[:, k].copy() for k, name in enumerate(names)},
                                   PIState(compiled.basis, final), info.as_dict())
        if st == N.PIMI_ENOTCONVERGED and check:
            N.check(st)
        return res

    def is_dynamics(self, job: int) -> bool:
        meta = self._jobs[job][1]
        return isinstance(meta, tuple) and meta[0] == "dynamics"

    def run(self, max_ticks=1 << 30, stop_when_finished=False) -> int:
        fin = C.c_int64()
        N.check(N.lib().pimi_batch_run(self._h, C.c_int32(min(max_ticks, (1 << 31) - 1)),
                                       C.c_int32(1 if stop_when_finished else 0), C.byref(fin)))
        return fin.value

    def counts(self) -> dict:
        c = np.zeros(6, dtype=np.int64)
        N.check(N.lib().pimi_batch_counts(self._h, N.ptr(c)))
        return dict(zip(("queued", "active", "finished", "ticks", "bytes_in_use", "bytes_reserved"),
                        (int(v) for v in c)))

    def stats(self, reset=False) -> dict:
        """Per operation class: device ms (timing engines only), algorithmic
        bytes and flops accumulated over the ticks run so far."""
        out = np.zeros(24)
        N.check(N.lib().pimi_batch_stats(self._h, N.ptr(out), C.c_int32(1 if reset else 0)))
        return {name: {"ms": out[c], "bytes": out[8 + c], "flops": out[16 + c]}
                for c, name in enumerate(N.OP_CLASSES)}

    def next_finished(self):
        j = C.c_int64()
        N.check(N.lib().pimi_batch_next_finished(self._h, C.byref(j)))
        return None if j.value < 0 else j.value

    def fetch(self, job: int, check=True) -> SteadyStateResult:
        """Result of a finished job; raises NotConverged (like
        stationary_state) unless check is False."""
        compiled, alg = self._jobs[job]
        rho = np.zeros(len(compiled.basis), dtype=np.complex128)
        info = N.SolveInfo()
        st = N.lib().pimi_batch_fetch(self._h, C.c_int64(job), N.ptr(rho), C.byref(info))
        if st not in (N.PIMI_OK, N.PIMI_ENOTCONVERGED):
            N.check(st)
        del self._jobs[job]
        res = SteadyStateResult(PIState(compiled.basis, rho), info.as_dict(), alg)
        if st == N.PIMI_ENOTCONVERGED and check:
            N.check(st)
        return res

    def graph_counts(self):
        """(ticks launched as HIP graphs, graphs instantiated, graphs updated
        in place)."""
        out = np.zeros(3, dtype=np.int64)
        N.check(N.lib().pimi_batch_graph_counts(self._h, N.ptr(out)))
        return tuple(int(v) for v in out)

    def autotune(self, compiled_list, reps=5):
        """Pick the engine's banded-apply tile widths by timing launches at
        each width: one over a vector of every model (the width of sizes not
        tuned here, returned), then one per distinct model size (N), whose
        fastest width the engine's later launches use for that size.  The
        width of each model is left in `self.widths` (0: not banded)."""
        arr = (C.c_void_p * len(compiled_list))(*[c._plan.value for c in compiled_list])
        w = C.c_int32()
        per = np.zeros(len(compiled_list), dtype=np.int32)
        N.check(N.lib().pimi_batch_autotune_apply(self._h, arr, C.c_int32(len(compiled_list)), C.c_int32(reps),
                                                  C.byref(w), N.ptr(per)))
        self.widths = [int(v) for v in per]
        return w.value

    def apply(self, compiled_list, xs):
        """Y_k = L_k X_k for every model k in one launch (the engine's
        multi-model apply), computed in the engine precision; xs: one vector
        per model.  The multi-model form of the batch apply!
        (src/liouvillian.jl:3232-3251), across models instead of columns."""
        if len(xs) != len(compiled_list):
            raise ValueError("one input vector per model")
        for c, x in zip(compiled_list, xs):
            if np.shape(x) != (c.info["n"],):
                raise ValueError("input vector length must equal the model's PI dimension")
        x = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.complex128) for v in xs]))
        y = np.empty_like(x)
        arr = (C.c_void_p * len(compiled_list))(*[c._plan.value for c in compiled_list])
        N.check(N.lib().pimi_batch_apply(self._h, arr, C.c_int32(len(compiled_list)), N.ptr(x), N.ptr(y)))
        out, k = [], 0
        for c in compiled_list:
            n = c.info["n"]
            out.append(y[k:k + n])
            k += n
        return out

    def apply_bench(self, compiled_list, reps=20):
        """One launch applying L to a vector of every model, event-timed:
        (ms per launch, algorithmic bytes per launch)."""
        arr = (C.c_void_p * len(compiled_list))(*[c._plan.value for c in compiled_list])
        ms = C.c_double()
        nb = C.c_int64()
        N.check(N.lib().pimi_batch_apply_bench(self._h, arr, C.c_int32(len(compiled_list)), C.c_int32(reps),
                                               C.byref(ms), C.byref(nb)))
        return ms.value, nb.value


def batched_stationary_states(compiled_models, algorithm: BatchAlgorithm | None = None, atol=0.0, rtol=1e-8,
                              precision=N.PIMI_C64, engine: BatchEngine | None = None, check=True):
    """Stationary states of many compiled models, solved together; results in
    input order."""
    own = engine is None
    if own:
        # reserve what the jobs need (all resident at once), not the default
        # three quarters of free HBM: a large reservation costs time to map
        alg = algorithm or BatchAlgorithm()
        need = 0
        for c in compiled_models:
            b = C.c_int64()
            N.check(N.lib().pimi_batch_job_bytes(c._plan, C.c_int32(precision), C.c_int32(alg.krylovdim),
                                                 C.byref(b)))
            need += b.value + 256
        eng = BatchEngine(precision=precision, memory_bytes=need + (1 << 20),
                          device=compiled_models[0].device if compiled_models else 0)
    else:
        eng = engine
    try:
        ids = [eng.submit(c, algorithm, atol=atol, rtol=rtol) for c in compiled_models]
        out = {}
        while len(out) < len(ids):
            eng.run(stop_when_finished=True)
            while (j := eng.next_finished()) is not None:
                out[j] = eng.fetch(j, check=check)
            if eng.counts()["active"] == 0 and eng.counts()["queued"] == 0 and len(out) < len(ids):
                raise RuntimeError("batch engine stopped with unfinished jobs")
        return [out[j] for j in ids]
    finally:
        if own:
            eng.close()
