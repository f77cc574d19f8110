"""Device-resident request loop: continuous batching of stationary-state and
dynamics requests on one device.  The reference has no serving layer; this
module drives the reference's two solver entry points (stationary_state,
src/highlevel.jl:321-357, and solve_dynamics, src/highlevel.jl:559-612) for a
stream of requests.

Stationary requests are continuously batched in one BatchEngine: a request
that arrives while others are being solved is admitted between two ticks as
soon as the engine's memory reservation holds its Krylov state (HBM-footprint
admission), and finished jobs free their memory at once.  Dynamics requests
are continuously batched the same way in a second, FP64 engine (its own
memory reservation and streams): their Arnoldi steps, projections and
observations are one launch per step kind over every dynamics request in
flight, and the loop ticks both engines in turn.  (`dynamics_mode="workers"`
runs each dynamics request on its own worker thread, workspace and HIP
stream instead.)  Model compilation and the stationary solver's host setup
run on a thread pool off the loop thread.

A request's result does not depend on the other requests in flight: the
engines' arithmetic per job depends only on that job, and a worker's
dynamics request runs on its own workspace with deterministic reductions.

With `snapshot_dir`, every finished request's state (the stationary state,
or a dynamics request's state at t_end) is written as a PIDCHKPT v1
checkpoint (src/checkpoints.jl:114-138) named by its request index, with the
request's parameters and solver figures as metadata, so served results can
be reloaded here or by the reference package.
"""
from __future__ import annotations

import os
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .api import ExpvAlgorithm, PIState, compile as pi_compile, computational_product_state, solve_dynamics
from .batch import BatchAlgorithm, BatchEngine
from .checkpoints import save_checkpoint


@dataclass
class ServedResult:
    index: int                 # position of the request in the submitted stream
    kind: str                  # "steady" or "dynamics"
    N: int
    arrival: float             # seconds after the stream started
    finished: float
    latency: float
    converged: bool
    state: PIState | None = None          # stationary state (trace one), or the dynamics state at t_end
    values: dict | None = None            # dynamics: observable series
    times: np.ndarray | None = None
    info: dict = field(default_factory=dict)
    snapshot: str | None = None           # PIDCHKPT file of the state (with snapshot_dir)


def dynamics_bytes(n, krylovdim):
    """Device bytes of one dynamics workspace (basis + work vectors)."""
    return (krylovdim + 10) * n * 16


class Server:
    def __init__(self, device=0, precision=N.PIMI_C32, steady_algorithm: BatchAlgorithm | None = None,
                 dynamics_algorithm: ExpvAlgorithm | None = None, rtol=1e-8, memory_bytes=0,
                 dynamics_workers=2, dynamics_memory_bytes=32 << 30, compile_workers=8, t_end=10.0, saveat=0.2,
                 snapshot_dir=None, dynamics_mode="engine"):
        if dynamics_mode not in ("engine", "workers"):
            raise ValueError("dynamics_mode must be 'engine' or 'workers'")
        self.device = device
        self.rtol = rtol
        self.steady_algorithm = steady_algorithm or BatchAlgorithm(krylovdim=200, inner_rtol=1e-4)
        self.dynamics_algorithm = dynamics_algorithm or ExpvAlgorithm(krylovdim=30, atol=0.0, rtol=rtol)
        self.dynamics_mode = dynamics_mode
        # the dynamics engine first: the stationary one reserves a share of what is left
        self.dyn_engine = BatchEngine(precision=N.PIMI_C64, memory_bytes=dynamics_memory_bytes, device=device) \
            if dynamics_mode == "engine" else None
        self.engine = BatchEngine(precision=precision, memory_bytes=memory_bytes, device=device)
        self.dynamics_workers = dynamics_workers
        self.dynamics_memory_bytes = dynamics_memory_bytes
        self.compile_workers = compile_workers
        self.t_end = t_end
        self.saveat = saveat
        self.snapshot_dir = snapshot_dir
        if snapshot_dir is not None:
            os.makedirs(snapshot_dir, exist_ok=True)

    def close(self):
        self.engine.close()
        if self.dyn_engine is not None:
            self.dyn_engine.close()

    def _dynamics_inputs(self, comp):
        rho0 = computational_product_state(comp.basis, 2)   # every particle in the upper level
        return rho0, {"n": np.array([[0, 0], [0, 1]], dtype=complex)}

    # -- request preparation (host, thread pool) --------------------------------
    def _prepare(self, req):
        comp = pi_compile(req.instance.model(), device=self.device)
        if req.kind == "steady":
            comp.prepare()
        return comp

    def _run_dynamics(self, comp, req):
        rho0, obs = self._dynamics_inputs(comp)
        return solve_dynamics(comp, rho0, (0.0, self.t_end), saveat=self.saveat, algorithm=self.dynamics_algorithm,
                              observables=obs)

    def _snapshot(self, result, req):
        """Write a finished request's state as a PIDCHKPT v1 file."""
        if self.snapshot_dir is None or result.state is None:
            return
        inst = req.instance
        md = {"request": result.index, "kind": result.kind, "N": inst.N, "omega": repr(inst.omega),
              "gamma": repr(inst.gamma), "gamma_phi": repr(inst.gamma_phi), "converged": int(result.converged)}
        for key in ("relative_residual", "iterations", "accepted", "applications"):
            if key in result.info:
                md[key] = repr(result.info[key])
        path = os.path.join(self.snapshot_dir, f"request_{result.index:06d}_{result.kind}.pid")
        result.snapshot = save_checkpoint(path, result.state, time=self.t_end if result.kind == "dynamics" else None,
                                          metadata=md)

    def serve(self, requests, realtime=True):
        """Run the stream; `requests` carry `arrival` (seconds), `kind` and
        `instance`.  With realtime, a request is not touched before its
        arrival time; otherwise all are available at once.  Returns the
        results in request order."""
        reqs = list(requests)
        order = sorted(range(len(reqs)), key=lambda i: reqs[i].arrival)
        results: list[ServedResult | None] = [None] * len(reqs)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        lock = threading.Lock()
        prepared_steady = queue.Queue()
        prepared_dyn = queue.Queue()   # dynamics requests for the dynamics engine
        dyn_queue = queue.Queue()      # dynamics requests for the workers
        dyn_budget = [self.dynamics_memory_bytes]
        budget_cv = threading.Condition()
        errors = []

        def prepare(i):
            req = reqs[i]
            try:
                comp = self._prepare(req)
            except Exception as e:  # reported by the loop
                errors.append(e)
                return
            if req.kind == "steady":
                prepared_steady.put((i, comp))
            elif self.dyn_engine is not None:
                prepared_dyn.put((i, comp))
            else:
                dyn_queue.put((i, comp))

        def dyn_worker():
            while True:
                item = dyn_queue.get()
                if item is None:
                    return
                i, comp = item
                need = dynamics_bytes(comp.info["n"], self.dynamics_algorithm.krylovdim)
                with budget_cv:   # admission by device footprint
                    while dyn_budget[0] < need and dyn_budget[0] < self.dynamics_memory_bytes:
                        budget_cv.wait()
                    dyn_budget[0] -= need
                try:
                    res = self._run_dynamics(comp, reqs[i])
                    t = now()
                    r = ServedResult(i, "dynamics", comp.basis.N, reqs[i].arrival, t, t - reqs[i].arrival,
                                     bool(res.info.get("converged", 1)), state=res.final_state, values=res.values,
                                     times=res.t, info=res.info)
                    self._snapshot(r, reqs[i])
                    with lock:
                        results[i] = r
                except Exception as e:
                    errors.append(e)
                finally:
                    comp.release_workspaces()
                    with budget_cv:
                        dyn_budget[0] += need
                        budget_cv.notify_all()

        pool = ThreadPoolExecutor(max_workers=self.compile_workers)
        nworkers = self.dynamics_workers if self.dyn_engine is None else 0
        workers = [threading.Thread(target=dyn_worker, daemon=True) for _ in range(nworkers)]
        for w in workers:
            w.start()

        def dispatch():   # hands each request to the preparation pool at its arrival time
            for i in order:
                if realtime:
                    delay = reqs[i].arrival - now()
                    if delay > 0:
                        time.sleep(delay)
                pool.submit(prepare, i)

        dispatcher = threading.Thread(target=dispatch, daemon=True)
        dispatcher.start()
        n_steady = sum(1 for r in reqs if r.kind == "steady")
        n_dyn = len(reqs) - n_steady if self.dyn_engine is not None else 0
        jobs = {}      # steady engine job id -> (request index, compiled model)
        eng, deng = self.engine, self.dyn_engine

        def dyn_loop():   # the dynamics engine's own loop (its ticks block only this thread)
            djobs, done = {}, 0
            try:
                while done < n_dyn and not errors:
                    while True:
                        try:
                            i, comp = prepared_dyn.get_nowait()
                        except queue.Empty:
                            break
                        rho0, obs = self._dynamics_inputs(comp)
                        jid = deng.submit_dynamics(comp, rho0, (0.0, self.t_end), saveat=self.saveat,
                                                   algorithm=self.dynamics_algorithm, observables=obs)
                        djobs[jid] = (i, comp)
                    c = deng.counts()
                    if c["active"] + c["queued"] > 0:
                        deng.run(max_ticks=1)
                    else:
                        time.sleep(2e-4)
                    while (jid := deng.next_finished()) is not None:
                        i, comp = djobs.pop(jid)
                        res = deng.fetch_dynamics(jid, check=False)
                        t = now()
                        r = ServedResult(i, "dynamics", comp.basis.N, reqs[i].arrival, t, t - reqs[i].arrival,
                                         bool(res.info.get("converged", 1)), state=res.final_state,
                                         values=res.values, times=res.t, info=res.info)
                        self._snapshot(r, reqs[i])
                        with lock:
                            results[i] = r
                        done += 1
            except Exception as e:
                errors.append(e)

        dyn_thread = threading.Thread(target=dyn_loop, daemon=True) if n_dyn else None
        if dyn_thread is not None:
            dyn_thread.start()
        done_steady = 0
        try:
            while done_steady < n_steady:
                if errors:
                    raise errors[0]
                # admit every prepared request that has arrived (between ticks)
                while True:
                    try:
                        i, comp = prepared_steady.get_nowait()
                    except queue.Empty:
                        break
                    jid = eng.submit(comp, self.steady_algorithm, rtol=self.rtol)
                    jobs[jid] = (i, comp)
                c = eng.counts()
                if c["active"] + c["queued"] > 0:
                    eng.run(max_ticks=1)
                else:
                    time.sleep(2e-4)
                while (jid := eng.next_finished()) is not None:
                    i, comp = jobs.pop(jid)
                    r = eng.fetch(jid, check=False)
                    t = now()
                    results[i] = ServedResult(i, "steady", comp.basis.N, reqs[i].arrival, t, t - reqs[i].arrival,
                                              bool(r.info["converged"]), state=r.state, info=r.info)
                    self._snapshot(results[i], reqs[i])
                    done_steady += 1
            if dyn_thread is not None:
                dyn_thread.join()
            dispatcher.join()
            pool.shutdown(wait=True)
            for _ in workers:
                dyn_queue.put(None)
            for w in workers:
                w.join()
            if errors:
                raise errors[0]
        finally:
            pool.shutdown(wait=False)
        return results


def summarize(results, wall=None):
    """p50/p99 latency (ms) per kind and overall throughput."""
    out = {}
    for kind in ("steady", "dynamics", "all"):
        rs = [r for r in results if r is not None and (kind == "all" or r.kind == kind)]
        if not rs:
            continue
        lat = np.array([r.latency for r in rs]) * 1e3
        span = wall if wall is not None else max(r.finished for r in rs) - min(r.arrival for r in rs)
        out[kind] = {"count": len(rs), "p50_ms": float(np.percentile(lat, 50)),
                     "p99_ms": float(np.percentile(lat, 99)), "per_s": len(rs) / span if span > 0 else None,
                     "converged": int(sum(r.converged for r in rs))}
    return out
