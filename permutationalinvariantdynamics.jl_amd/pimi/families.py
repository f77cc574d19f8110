"""Rate families and parameter scans (src/compiled_families.jl:1-300,
src/scans.jl:42-300,1457-1480).

compile_family(model; rate_indices) fixes a model's operators and selects
term rates to vary; specialize(family, rates) binds numbers to them.  The
expensive part of compiling, the Schur geometry (Clebsch-Gordan tables and
one-body transitions of the PI basis), depends on (N, d) only and is built
once per process in libpimi's lowering cache, so a specialization lowers only
the per-term blocks and uploads its stencil tables.  Rate indices here are
0-based positions in model.terms (the reference's are 1-based).

parameter_scan(plan) runs a family's scan points as one device batch: every
point is specialized on a host thread pool and all stationary solves advance
together in a BatchEngine (continuous batching by HBM footprint), instead of
the reference's one-point-at-a-time serial loop.
"""
from __future__ import annotations

import dataclasses
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .api import CompiledPIModel, PIModel, compile as pi_compile
from .batch import BatchAlgorithm, BatchEngine


def _real_finite(v, what):
    if isinstance(v, (bool, np.bool_)) or not isinstance(v, (int, float, np.integer, np.floating)):
        raise ValueError(f"{what} must be a real number")
    if not np.isfinite(float(v)):
        raise ValueError(f"{what} must be finite")
    return float(v)


@dataclass
class CompiledPIModelFamily:
    model: PIModel
    rate_indices: tuple
    default_rates: tuple
    estimates: dict = field(default_factory=dict)


def compile_family(model: PIModel, rate_indices=None) -> CompiledPIModelFamily:
    """Compile the fixed operator geometry of `model` once (host lowering) and
    select the term rates a specialization may change."""
    nterms = len(model.terms)
    if rate_indices is None:
        idx = list(range(nterms))
    else:
        idx = list(rate_indices)
        if not idx:
            raise ValueError("rate_indices cannot be empty")
        if not all(isinstance(i, (int, np.integer)) and not isinstance(i, (bool, np.bool_)) for i in idx):
            raise ValueError("rate_indices must contain integers")
        if not all(0 <= i < nterms for i in idx):
            raise IndexError(f"rate_indices {idx} out of range for {nterms} terms")
        if len(set(idx)) != len(idx):
            raise ValueError("rate_indices must be unique")
        idx = sorted(int(i) for i in idx)
    defaults = tuple(_real_finite(model.terms[i].rate, "every selected family rate in the prototype") for i in idx)
    t0 = time.perf_counter()
    proto = pi_compile(model, host_only=True)  # validates the model and builds the cached geometry
    est = dict(dimension=proto.info["n"], rate_indices=tuple(idx), geometry_reused=True,
               table_bytes_c64=proto.info["table_bytes_c64"], table_bytes_c32=proto.info["table_bytes_c32"],
               prototype_compile_seconds=time.perf_counter() - t0)
    return CompiledPIModelFamily(model, tuple(idx), defaults, est)


def _family_rates(family, rates):
    count = len(family.rate_indices)
    if np.isscalar(rates) and not isinstance(rates, (str, bytes)):
        if count != 1:
            raise ValueError("a scalar rate is valid only for a family with one selected term")
        values = (rates,)
    else:
        values = tuple(rates)
    if len(values) != count:
        raise ValueError(f"expected {count} family rates, got {len(values)}")
    return tuple(_real_finite(v, f"family rate {k}") for k, v in enumerate(values))


def _replace_rates(model, indices, rates):
    terms = list(model.terms)
    for i, r in zip(indices, rates):
        terms[i] = dataclasses.replace(terms[i], rate=r)
    return PIModel(model.basis, terms)


class SpecializedPIModel(CompiledPIModel):
    """A compiled model bound to one rate point of a family; it is a
    CompiledPIModel (apply, steady state, dynamics, spectra, scans)."""

    def __init__(self, family: CompiledPIModelFamily, rates, device=0):
        self.family = family
        self.rates = rates
        super().__init__(_replace_rates(family.model, family.rate_indices, rates), device=device)


def specialize(family: CompiledPIModelFamily, rates=None, device=0) -> SpecializedPIModel:
    """Bind numerical rates (in family.rate_indices order; a scalar when one
    rate is selected; None = the prototype's) without rebuilding geometry."""
    bound = family.default_rates if rates is None else _family_rates(family, rates)
    return SpecializedPIModel(family, bound, device=device)


@dataclass
class ParameterScanPoint:
    index: int
    parameter: object
    status: str
    output: object = None
    residual: float = None
    trace_error: float = None
    converged: bool = False
    iterations: int = None
    compile_seconds: float = 0.0
    solve_seconds: float = 0.0
    elapsed_seconds: float = 0.0
    warm_started: bool = False
    workspace_reused: bool = False
    diagnostics: object = None
    error_type: str = None
    message: str = None


@dataclass
class ParameterScanResult:
    task: str
    parameters: list
    points: list
    metadata: dict = field(default_factory=dict)

    def __len__(self):
        return len(self.points)

    def __getitem__(self, k):
        return self.points[k]

    def __iter__(self):
        return iter(self.points)


@dataclass
class ParameterScanPlan:
    """parameters: the scan points; rate_builder(parameter[, index]) maps one
    to family rates (default: the parameter is the rate tuple).  The task is
    the stationary state, solved to `rtol` by the batch engine's refined
    GMRES in `precision` (C32 Arnoldi in FP64 refinement by default)."""
    parameters: list
    family: CompiledPIModelFamily
    rate_builder: object = None
    task: str = "steady_state"
    rtol: float = 1e-8
    algorithm: BatchAlgorithm = None
    precision: int = N.PIMI_C32
    save_output: bool = True
    memory_bytes: int = 0
    compile_workers: int = 8

    def __post_init__(self):
        if self.task != "steady_state":
            raise ValueError("the device scan task is 'steady_state'")
        self.parameters = list(self.parameters)

    def rates_for(self, index):
        p = self.parameters[index]
        if self.rate_builder is None:
            return p
        try:
            return self.rate_builder(p, index)
        except TypeError:
            return self.rate_builder(p)


def parameter_scan(plan: ParameterScanPlan, indices=None, on_error="stop", device=0) -> ParameterScanResult:
    """Specialize and solve the selected points (all of them by default); one
    ParameterScanPoint per point in increasing index order.  on_error: 'stop'
    (raise at the first failure), 'record' or 'throw' (failures recorded, then
    'throw' raises after the scan)."""
    if on_error not in ("stop", "record", "throw"):
        raise ValueError("on_error must be 'stop', 'record', or 'throw'")
    sel = list(range(len(plan.parameters))) if indices is None else sorted(set(int(i) for i in indices))
    if any(not 0 <= i < len(plan.parameters) for i in sel):
        raise IndexError("scan index out of range")
    t_start = time.perf_counter()
    alg = plan.algorithm or BatchAlgorithm()
    points = {}
    comps = {}

    def build(i):
        t0 = time.perf_counter()
        comp = specialize(plan.family, plan.rates_for(i), device=device)
        comp.prepare()
        return i, comp, time.perf_counter() - t0

    def fail(i, e, compile_s=0.0):
        if on_error == "stop":
            raise e
        points[i] = ParameterScanPoint(i, plan.parameters[i], "failed", compile_seconds=compile_s,
                                       error_type=type(e).__name__, message=str(e))

    with ThreadPoolExecutor(max_workers=max(1, plan.compile_workers)) as pool:
        futs = [pool.submit(build, i) for i in sel]
        compile_s = {}
        for i, f in zip(sel, futs):
            try:
                _, comp, s = f.result()
                comps[i] = comp
                compile_s[i] = s
            except Exception as e:  # noqa: BLE001 -- recorded per point
                fail(i, e)
    eng = BatchEngine(precision=plan.precision, memory_bytes=plan.memory_bytes, device=device)
    try:
        jobs = {}
        t_sub = time.perf_counter()
        for i, comp in comps.items():
            jobs[eng.submit(comp, alg, rtol=plan.rtol)] = i
        eng.run()
        solve_s = time.perf_counter() - t_sub
        while (jid := eng.next_finished()) is not None:
            i = jobs[jid]
            r = eng.fetch(jid, check=False)
            ok = bool(r.info["converged"])
            pt = ParameterScanPoint(i, plan.parameters[i], "success" if ok else "failed",
                                    output=r.state if plan.save_output else None,
                                    residual=r.info["relative_residual"], trace_error=r.info["trace_error"],
                                    converged=ok, iterations=r.info["iterations"], compile_seconds=compile_s[i],
                                    solve_seconds=solve_s, elapsed_seconds=time.perf_counter() - t_start,
                                    diagnostics=r.info)
            if not ok:
                pt.error_type, pt.message = "NotConverged", "stationary solve did not converge"
                if on_error == "stop":
                    raise N.NotConverged(N.PIMI_ENOTCONVERGED, f"scan point {i}: stationary solve did not converge")
            points[i] = pt
    finally:
        eng.close()
        for comp in comps.values():
            comp.release_workspaces()
    out = ParameterScanResult(plan.task, plan.parameters, [points[i] for i in sel],
                              dict(seconds=time.perf_counter() - t_start, batched=True, points=len(sel)))
    if on_error == "throw" and any(p.status == "failed" for p in out.points):
        bad = next(p for p in out.points if p.status == "failed")
        raise RuntimeError(f"scan point {bad.index} failed: {bad.message}")
    return out
