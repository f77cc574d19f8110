"""Batch engine: many stationary solves advanced together on the device.

Checks: results against the oracle's preconditioned GMRES (FP64 and FP32
Arnoldi engines), bit-identical results for a job alone and in a mixed batch,
admission under a small memory reservation, the closed form of the local-only
(batched-workload) model family, failure reporting, and the multi-model
apply probe.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import parity_log
import pimi
import pimi._native as N
from oracle import lowering as L
from oracle import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_closed_form as CF  # noqa: E402

pytestmark = pytest.mark.gpu
SP2 = pimi.spin_matrices(2)


def _c2_model(N_):
    return pimi.qubit_ensemble_model(N_, hamiltonian=SP2.jz + SP2.jx, emission=0.1, dephasing=0.2, pumping=0.02,
                                     collective_emission=0.1)


def _c2_terms():
    sp = L.spin_matrices(2)
    return L.qubit_ensemble_terms(hamiltonian=sp["jz"] + sp["jx"], emission=0.1, dephasing=0.2, pumping=0.02,
                                  collective_emission=0.1)


def _c4_model(N_, omega, gphi):
    return pimi.Models.driven_qubits(N_, drive=2 * omega, detuning=1.0, emission=0.1, dephasing=4 * gphi,
                                     pumping=0.0)


def _c4_closed_n(N_, omega, gphi):
    H = 2 * omega * CF.jx + 1.0 * CF.jz
    return CF.steady(N_, H, [(0.1, CF.jm), (4 * gphi, CF.jz)])[0]


@pytest.mark.parametrize("precision,weighting", [(N.PIMI_C64, "sector"), (N.PIMI_C32, "sector"),
                                                  (N.PIMI_C64, None)])
def test_batch_matches_oracle(precision, weighting):
    print(f"batch vs oracle precision={precision} weighting={weighting}", flush=True)
    sizes = (8, 13, 20)
    comps = [pimi.compile(_c2_model(n_)) for n_ in sizes]
    alg = pimi.BatchAlgorithm(krylovdim=150, inner_rtol=1e-4, weighting=weighting)
    res = pimi.batched_stationary_states(comps, alg, rtol=1e-11, precision=precision)
    for n_, r in zip(sizes, res):
        ref, _ = O.compile_plan(n_, 2, _c2_terms())
        x_ref, info_ref = ref.steady_state(krylovdim=400, maxiter=400, atol=0.0, rtol=1e-12,
                                           preconditioner="sector", criterion="backward", weighting=weighting)
        assert info_ref["converged"] and r.info["converged"]
        assert r.info["relative_residual"] <= 1e-11
        assert abs(r.info["trace_error"]) <= 1e-12
        err = np.linalg.norm(r.state.data - x_ref) / np.linalg.norm(x_ref)
        print(f"  N={n_} iterations {r.info['iterations']} cycles {r.info['refinements']} err {err:.2e}", flush=True)
        assert err <= 1e-8, (n_, err)


@pytest.mark.parametrize("precision", [N.PIMI_C64, N.PIMI_C32])
def test_batch_real_and_complex_factors_share_launches(precision):
    """Jobs whose preconditioner factor U is real (H = jz + jx: the GEMMs
    take U as a real operand) and complex (H = jz + jx + 0.4 jy) run in the
    same engine launches; every job reaches the oracle's solution."""
    print(f"batch real/complex factors precision={precision}", flush=True)
    sp = L.spin_matrices(2)
    cases = []
    for hy in (0.0, 0.4):
        for n_ in (13, 20):
            cases.append((n_, hy))
    comps = [pimi.compile(pimi.qubit_ensemble_model(n_, hamiltonian=SP2.jz + SP2.jx + hy * SP2.jy, emission=0.1,
                                                    dephasing=0.2, pumping=0.02, collective_emission=0.1))
             for n_, hy in cases]
    alg = pimi.BatchAlgorithm(krylovdim=150, inner_rtol=1e-4, weighting="sector")
    res = pimi.batched_stationary_states(comps, alg, rtol=1e-11, precision=precision)
    for (n_, hy), r in zip(cases, res):
        ref, _ = O.compile_plan(n_, 2, L.qubit_ensemble_terms(hamiltonian=sp["jz"] + sp["jx"] + hy * sp["jy"],
                                                              emission=0.1, dephasing=0.2, pumping=0.02,
                                                              collective_emission=0.1))
        x_ref, info_ref = ref.steady_state(krylovdim=400, maxiter=400, atol=0.0, rtol=1e-12,
                                           preconditioner="sector", criterion="backward", weighting="sector")
        assert info_ref["converged"] and r.info["converged"]
        assert r.info["relative_residual"] <= 1e-11
        err = np.linalg.norm(r.state.data - x_ref) / np.linalg.norm(x_ref)
        print(f"  N={n_} jy={hy} iterations {r.info['iterations']} err {err:.2e}", flush=True)
        assert err <= 1e-8, (n_, hy, err)


@pytest.mark.parametrize("precision", [N.PIMI_C64, N.PIMI_C32])
def test_batch_small_spaces(precision):
    """PI dimensions 4, 10 and 20 with krylovdim 150: each job's basis holds
    min(krylovdim, n) columns, as the reference's KrylovWorkspace."""
    print(f"batch small spaces precision={precision}", flush=True)
    sizes = (1, 2, 3)
    comps = [pimi.compile(_c2_model(n_)) for n_ in sizes]
    alg = pimi.BatchAlgorithm(krylovdim=150, inner_rtol=1e-4)
    res = pimi.batched_stationary_states(comps, alg, rtol=1e-11, precision=precision)
    for n_, r in zip(sizes, res):
        ref, _ = O.compile_plan(n_, 2, _c2_terms())
        x_ref, info_ref = ref.steady_state(krylovdim=400, maxiter=400, atol=0.0, rtol=1e-12,
                                           preconditioner="sector", criterion="backward")
        assert info_ref["converged"] and r.info["converged"]
        err = np.linalg.norm(r.state.data - x_ref) / np.linalg.norm(x_ref)
        assert err <= 1e-8, (n_, err)


@pytest.mark.parametrize("precision", [N.PIMI_C64, N.PIMI_C32])
def test_batch_alone_equals_mixed(precision):
    """A job's result does not depend on the other jobs in its launches."""
    print(f"batch alone vs mixed precision={precision}", flush=True)
    target = pimi.compile(_c4_model(40, 0.8, 0.01))
    others = [pimi.compile(_c2_model(17)), pimi.compile(_c4_model(20, 1.3, 0.05)),
              pimi.compile(_c4_model(61, 0.4, 0.002)), pimi.compile(_c2_model(9))]
    alg = pimi.BatchAlgorithm(krylovdim=120, inner_rtol=1e-4)
    alone = pimi.batched_stationary_states([target], alg, rtol=1e-9, precision=precision)[0]
    mixed = pimi.batched_stationary_states([others[0], others[1], target, others[2], others[3]], alg, rtol=1e-9,
                                           precision=precision)
    m = mixed[2]
    assert alone.info["iterations"] == m.info["iterations"]
    np.testing.assert_array_equal(alone.state.data, m.state.data)
    assert m.info["relative_residual"] <= 1e-9


def test_batch_admission_under_small_reservation():
    """With room for about two jobs at a time (admission is by footprint, so
    smaller jobs may share the room three at once) the queue drains and
    every result equals the unconstrained run bit for bit."""
    print("batch admission", flush=True)
    comps = [pimi.compile(_c4_model(n_, 0.5 + 0.1 * k, 0.01)) for k, n_ in enumerate((20, 30, 24, 36, 28, 32))]
    alg = pimi.BatchAlgorithm(krylovdim=100, inner_rtol=1e-4)
    free = pimi.batched_stationary_states(comps, alg, rtol=1e-9)
    probe = pimi.BatchEngine()
    biggest = max(probe.job_bytes(c, alg.krylovdim) for c in comps)
    probe.close()
    eng = pimi.BatchEngine(memory_bytes=2 * biggest + 4096)
    ids = [eng.submit(c, alg, rtol=1e-9) for c in comps]
    seen_active = seen_queued = 0
    out = {}
    while len(out) < len(ids):
        eng.run(max_ticks=1)
        c = eng.counts()
        seen_active = max(seen_active, c["active"])
        seen_queued = max(seen_queued, c["queued"])
        assert c["bytes_in_use"] <= c["bytes_reserved"]
        while (j := eng.next_finished()) is not None:
            out[j] = eng.fetch(j)
    eng.close()
    assert 1 <= seen_active < len(ids) and seen_queued > 0
    for j, r in zip(ids, free):
        np.testing.assert_array_equal(out[j].state.data, r.state.data)


@pytest.mark.parametrize("precision", [N.PIMI_C64, N.PIMI_C32])
def test_batch_local_models_against_closed_form(precision):
    """Local-only driven qubits (the batched workload's model family) have
    the product steady state sigma_ss^N: <sum_i n_i> = N sigma_ee (sector
    weighting, the engine's default)."""
    print(f"batch closed form precision={precision}", flush=True)
    rng = np.random.default_rng(20261203)
    params = [(n_, float(rng.uniform(0, 2)), float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-1)))))
              for n_ in (20, 40, 60, 80, 100, 120)]
    # weak drive: nearly pure steady states, where unweighted coordinates fail
    params += [(40, 0.045, 0.071), (80, 0.045, 0.071)]
    comps = [pimi.compile(_c4_model(*p)) for p in params]
    alg = pimi.BatchAlgorithm(krylovdim=300, inner_rtol=1e-4)
    res = pimi.batched_stationary_states(comps, alg, rtol=1e-12, precision=precision)
    for p, comp, r in zip(params, comps, res):
        n_op = comp.collective_operator(SP2.jp @ SP2.jm)
        val = pimi.expect(n_op, r.state).real
        ref = _c4_closed_n(*p)
        print(f"  N={p[0]} omega={p[1]:.3f} gphi={p[2]:.4f} iterations {r.info['iterations']} "
              f"cycles {r.info['refinements']} rel.res {r.info['relative_residual']:.2e} "
              f"<n> err {abs(val - ref) / ref:.2e}", flush=True)
        assert r.info["relative_residual"] <= 1e-12
        assert abs(val - ref) <= 1e-6 * ref, (p, val, ref)


def test_batch_512_seeded_instances_one_engine():
    """The batched form at its configured size: the 512 seeded driven-qubit
    models of the batched workload (N in {20..120}, Omega ~ U[0,2], gamma_phi
    log-uniform, PCG64 seed) in one FP32 engine with the tile width
    autotuned per N and graph-captured ticks, each to relative residual
    1e-8.  Every model converges within the bar, and <sum_i n_i> of every
    model is checked against its product-state closed form per qubit (the
    excited population): a relative residual of 1e-8 leaves weakly driven
    models, whose <sum_i n_i> is tiny, with a large relative but a small
    absolute error."""
    from pimi.workloads import batched_instances
    print("batch: 512 seeded instances in one engine", flush=True)
    inst = batched_instances(512)
    comps = [pimi.compile(i.model()) for i in inst]
    eng = pimi.BatchEngine(precision=N.PIMI_C32)
    try:
        eng.autotune(comps, reps=2)
        alg = pimi.BatchAlgorithm(krylovdim=300, inner_rtol=1e-4)
        res = pimi.batched_stationary_states(comps, alg, rtol=1e-8, engine=eng)
        ticks, inst_g, upd = eng.graph_counts()
    finally:
        eng.close()
    assert ticks > 0 and upd > 0
    worst, worst_rel, worst_res = 0.0, 0.0, 0.0
    for i, comp, r in zip(inst, comps, res):
        assert r.info["converged"]
        worst_res = max(worst_res, r.info["relative_residual"])
        val = pimi.expect(comp.collective_operator(SP2.jp @ SP2.jm), r.state).real
        ref = _c4_closed_n(i.N, i.omega, i.gamma_phi)
        worst = max(worst, abs(val - ref) / i.N)
        worst_rel = max(worst_rel, abs(val - ref) / ref)
    print(f"  largest relative error of <sum n_i>: {worst_rel:.2e}", flush=True)
    parity_log.record("batch512_relative_residual", worst_res, 1e-8)
    parity_log.record("batch512_excited_population_vs_closed_form", worst, 1e-4)
    assert worst_res <= 1e-8
    assert worst <= 1e-4


def test_batch_not_converged_reports():
    comp = pimi.compile(_c2_model(20))
    eng = pimi.BatchEngine()
    j = eng.submit(comp, pimi.BatchAlgorithm(krylovdim=5, maxiter=3, max_refinements=1), rtol=1e-12)
    eng.run()
    assert eng.next_finished() == j
    r = eng.fetch(j, check=False)
    assert r.info["converged"] == 0 and r.info["iterations"] <= 3
    j2 = eng.submit(comp, pimi.BatchAlgorithm(krylovdim=5, maxiter=3, max_refinements=1), rtol=1e-12)
    eng.run()
    with pytest.raises(pimi.NotConverged):
        eng.fetch(j2)
    eng.close()


def test_batch_rejects_bad_requests():
    comp = pimi.compile(_c2_model(8))
    eng = pimi.BatchEngine(memory_bytes=1 << 20)
    jid = C.c_int64()
    bad = N.GmresOpts(30, N.PRECOND_NONE, 100, 0.0, 1e-8, 0.0, 1e-4, 10, N.CRITERION_BACKWARD, 0)
    assert N.lib().pimi_batch_submit(eng._h, comp._plan, C.byref(bad), C.byref(jid)) == N.PIMI_EINVAL
    bad = N.GmresOpts(30, N.PRECOND_SECTOR, 100, 0.0, 1e-8, 0.0, 1e-4, 10, N.CRITERION_REFERENCE, 0)
    assert N.lib().pimi_batch_submit(eng._h, comp._plan, C.byref(bad), C.byref(jid)) == N.PIMI_EINVAL
    big = pimi.compile(_c2_model(60))
    with pytest.raises(pimi.PIMIError) as e:
        eng.submit(big, pimi.BatchAlgorithm(krylovdim=400))
    assert e.value.status == N.PIMI_ENOMEM
    eng.close()


def test_batch_apply_bench_counts_bytes():
    comps = [pimi.compile(_c4_model(n_, 0.7, 0.01)) for n_ in (20, 60, 100)]
    eng = pimi.BatchEngine(memory_bytes=1 << 24)
    ms, nbytes = eng.apply_bench(comps, reps=5)
    eng.close()
    assert ms > 0
    assert nbytes == sum(2 * c.info["n"] * 16 + c.info["table_bytes_c64"] for c in comps)


def test_batch_autotune_keeps_results():
    comps = [pimi.compile(_c4_model(n_, 0.7, 0.01)) for n_ in (16, 24)]
    alg = pimi.BatchAlgorithm(krylovdim=60, inner_rtol=1e-4)
    ref = pimi.batched_stationary_states(comps, alg, rtol=1e-9)
    eng = pimi.BatchEngine()
    assert eng.autotune(comps, reps=2) in (4, 8, 16)
    assert len(eng.widths) == len(comps) and all(w in (4, 8, 16) for w in eng.widths)
    out = pimi.batched_stationary_states(comps, alg, rtol=1e-9, engine=eng)
    eng.close()
    for a, b in zip(ref, out):
        np.testing.assert_array_equal(a.state.data, b.state.data)


def _engine_run(comps, dyn_comps, precision, graphs):
    """Stationary jobs of `comps` and dynamics jobs of `dyn_comps` (FP64
    engines only) in one engine, ticks captured as HIP graphs or launched
    eagerly (PIMI_NO_GRAPH=1, read when the engine is created)."""
    if graphs:
        os.environ.pop("PIMI_NO_GRAPH", None)
    else:
        os.environ["PIMI_NO_GRAPH"] = "1"
    try:
        eng = pimi.BatchEngine(precision=precision, memory_bytes=1 << 30)
    finally:
        os.environ.pop("PIMI_NO_GRAPH", None)
    try:
        ids = [eng.submit(c, pimi.BatchAlgorithm(krylovdim=60, inner_rtol=1e-4), rtol=1e-9) for c in comps]
        alg = pimi.ExpvAlgorithm(krylovdim=20, rtol=1e-10)
        dids = [eng.submit_dynamics(c, _excited(c), (0.0, 2.0), saveat=0.5, algorithm=alg,
                                    observables=_dyn_observables(c)) for c in dyn_comps]
        eng.run()
        out = [eng.fetch(j) for j in ids]
        dout = [eng.fetch_dynamics(j) for j in dids]
        return out, dout, eng.graph_counts()
    finally:
        eng.close()


@pytest.mark.parametrize("precision", [N.PIMI_C64, N.PIMI_C32])
def test_graph_ticks_equal_eager_ticks(precision):
    """The engine captures each tick's launches as a HIP graph and, when a
    tick of the same shape recurs (the Arnoldi ticks of a stable job set),
    updates that graph's launch parameters in place.  Results are bitwise
    those of the same ticks launched eagerly, for stationary jobs in both
    precisions and dynamics jobs beside them in FP64, and most graph ticks
    are in-place updates."""
    print(f"graph vs eager ticks, precision {precision}", flush=True)
    comps = [pimi.compile(_c4_model(n_, 0.7, 0.01)) for n_ in (16, 24, 32)]
    dyn = [pimi.compile(_c2_model(n_)) for n_ in (8, 14)] if precision == N.PIMI_C64 else []
    ref, dref, counts0 = _engine_run(comps, dyn, precision, graphs=False)
    out, dout, counts = _engine_run(comps, dyn, precision, graphs=True)
    assert counts0 == (0, 0, 0)
    ticks, inst, upd = counts
    assert ticks > 0 and inst + upd == ticks and upd >= ticks // 2, counts
    for a, b in zip(ref, out):
        assert a.info["converged"] and b.info["converged"]
        np.testing.assert_array_equal(a.state.data, b.state.data)
    for a, b in zip(dref, dout):
        np.testing.assert_array_equal(a.final_state.data, b.final_state.data)
        for k in a.values:
            np.testing.assert_array_equal(a.values[k], b.values[k])


def _excited(comp):
    return pimi.computational_product_state(comp.basis, 2)


def _dyn_observables(comp):
    return {"n": SP2.jp @ SP2.jm, "jx": SP2.jx}


def _relerr(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


def test_batched_dynamics_match_the_single_vector_path():
    """Dynamics jobs in the batch engine (the adaptive Krylov exponential of
    solve_dynamics, src/highlevel.jl:559-733, with its Arnoldi steps,
    projections and observations in the engine's ticks beside a stationary
    job) against pimi.solve_dynamics on the same models: observables at every
    saved time and final states agree to 1e-10 relative (both at rtol 1e-12;
    the engine's Arnoldi basis is orthogonalized by delayed
    reorthogonalization, the single path's by CGS2), and the job statistics
    are those of an adaptive run (accepted steps, factorizations)."""
    print("batched dynamics vs single path", flush=True)
    alg = pimi.ExpvAlgorithm(krylovdim=30, rtol=1e-12)
    comps = [pimi.compile(pimi.Models.driven_qubits(8, drive=1.0, detuning=1.0, emission=0.1, dephasing=0.2,
                                                    pumping=0.0)),
             pimi.compile(_c2_model(20))]
    eng = pimi.BatchEngine(precision=N.PIMI_C64, memory_bytes=1 << 30)
    try:
        steady = eng.submit(pimi.compile(_c2_model(13)), pimi.BatchAlgorithm(krylovdim=150), rtol=1e-10)
        jobs = [eng.submit_dynamics(c, _excited(c), (0.0, 4.0), saveat=0.1, algorithm=alg,
                                    observables=_dyn_observables(c)) for c in comps]
        eng.run()
        res = [eng.fetch_dynamics(j) for j in jobs]
        assert eng.fetch(steady).info["converged"]
    finally:
        eng.close()
    for c, r in zip(comps, res):
        ref = pimi.solve_dynamics(c, _excited(c), (0.0, 4.0), saveat=0.1, algorithm=alg,
                                  observables=_dyn_observables(c))
        assert np.array_equal(r.t, ref.t)
        err = max(_relerr(r.values[k], ref.values[k]) for k in ref.values)
        serr = _relerr(r.final_state.data, ref.final_state.data)
        parity_log.record(f"batched_dynamics_n{c.basis.N}_observables", err, 1e-10)
        parity_log.record(f"batched_dynamics_n{c.basis.N}_state", serr, 1e-10)
        assert err <= 1e-10 and serr <= 1e-10, (err, serr)
        assert r.info["converged"] and r.info["accepted"] >= 40 and r.info["factorizations"] >= r.info["accepted"]


def test_batched_dynamics_c1_closed_form_and_alone_equals_mixed():
    """The C1 dynamics (driven_qubits(8) from the excited state, t in [0, 4])
    as a batch-engine job against the closed form of the single-qubit
    Lindbladian (<sum n_i>(t), 41 times), and bit for bit the same alone and
    in a batch with other dynamics and stationary jobs."""
    print("batched dynamics C1 closed form and determinism", flush=True)
    model = pimi.Models.driven_qubits(8, drive=1.0, detuning=1.0, emission=0.1, dephasing=0.2, pumping=0.0)
    comp = pimi.compile(model)
    alg = pimi.ExpvAlgorithm(krylovdim=30, rtol=1e-12)
    obs = {"n": SP2.jp @ SP2.jm}

    def run(extra):
        eng = pimi.BatchEngine(precision=N.PIMI_C64, memory_bytes=1 << 30)
        try:
            others = []
            if extra:
                c2 = pimi.compile(_c2_model(17))
                others.append(eng.submit_dynamics(c2, _excited(c2), (0.0, 2.0), saveat=0.5, algorithm=alg,
                                                  observables=_dyn_observables(c2)))
                eng.submit(pimi.compile(_c2_model(9)), pimi.BatchAlgorithm(krylovdim=100), rtol=1e-10)
            j = eng.submit_dynamics(comp, _excited(comp), (0.0, 4.0), saveat=0.1, algorithm=alg, observables=obs)
            eng.run()
            return eng.fetch_dynamics(j)
        finally:
            eng.close()

    alone, mixed = run(False), run(True)
    assert np.array_equal(alone.values["n"], mixed.values["n"])
    assert np.array_equal(alone.final_state.data, mixed.final_state.data)
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "closed_form.json")) as fh:
        g = json.load(fh)["c1"]
    np.testing.assert_allclose(alone.t, g["t"], rtol=0, atol=0)
    gold = np.asarray(g["n"])
    err = _relerr(alone.values["n"].real, gold)
    parity_log.record("batched_dynamics_c1_closed_form", err, 1e-8)
    assert err <= 1e-8, err


def test_batched_dynamics_rejects_bad_requests():
    comp = pimi.compile(_c2_model(6))
    eng32 = pimi.BatchEngine(precision=N.PIMI_C32, memory_bytes=1 << 26)
    try:
        with pytest.raises(pimi.PIMIError):
            eng32.submit_dynamics(comp, _excited(comp), (0.0, 1.0), saveat=0.5, observables=_dyn_observables(comp))
    finally:
        eng32.close()
    eng = pimi.BatchEngine(precision=N.PIMI_C64, memory_bytes=1 << 26)
    try:
        with pytest.raises(ValueError):
            eng.submit_dynamics(comp, _excited(comp), (0.0, 1.0), saveat=0.5)
        with pytest.raises(ValueError):
            eng.submit_dynamics(comp, _excited(comp), (0.0, 1.0), saveat=0.5, observables=_dyn_observables(comp),
                                algorithm=pimi.ExpvAlgorithm(safety=1.5))
        # the native checks behind the Python ones: non-monotone output times
        opts = N.ExpvOpts(30, 0, 0.0, 1e-10, 0.9, 1000)
        rho = np.ascontiguousarray(_excited(comp).data)
        ts = np.array([0.0, 1.0, 0.5])
        obs = np.ascontiguousarray(comp.collective_operator(SP2.jz).data[None, :])
        jid = C.c_int64()
        st = N.lib().pimi_batch_submit_expv(eng._h, comp._plan, N.ptr(rho), N.ptr(ts), C.c_int64(3), C.byref(opts),
                                            C.c_int32(1), N.ptr(obs), C.byref(jid))
        assert st == N.PIMI_EINVAL
    finally:
        eng.close()


@pytest.mark.parametrize("N_", [1, 2, 3])
def test_batched_dynamics_small_spaces_and_special_times(N_):
    """Dynamics jobs where the Krylov space is the whole space (exact
    breakdowns: n = 4, 9, 16 below the Krylov dimension), a repeated save
    time (a zero-length interval) and a zero initial state (the zero-state
    projection), against the single-vector path."""
    print(f"batched dynamics small spaces N={N_}", flush=True)
    c = pimi.compile(_c2_model(N_))
    alg = pimi.ExpvAlgorithm(krylovdim=30, rtol=1e-12)
    obs = _dyn_observables(c)
    times = [0.0, 0.5, 0.5, 1.25, 3.0]
    zero = pimi.PIState(c.basis, np.zeros(len(c.basis), dtype=np.complex128))
    eng = pimi.BatchEngine(precision=N.PIMI_C64, memory_bytes=1 << 26)
    try:
        j1 = eng.submit_dynamics(c, _excited(c), (0.0, 3.0), saveat=times, algorithm=alg, observables=obs)
        j2 = eng.submit_dynamics(c, zero, (0.0, 3.0), saveat=times, algorithm=alg, observables=obs)
        eng.run()
        r1, r2 = eng.fetch_dynamics(j1), eng.fetch_dynamics(j2)
    finally:
        eng.close()
    ref = pimi.solve_dynamics(c, _excited(c), (0.0, 3.0), saveat=times, algorithm=alg, observables=obs)
    assert np.array_equal(r1.t, ref.t)
    err = max(_relerr(r1.values[k], ref.values[k]) for k in ref.values)
    assert err <= 1e-12, err
    assert _relerr(r1.final_state.data, ref.final_state.data) <= 1e-12
    assert np.array_equal(r1.values["n"][1], r1.values["n"][2])  # the repeated time sees the same state
    assert all(np.all(v == 0) for v in r2.values.values()) and np.all(r2.final_state.data == 0)
    assert r2.info["converged"]


def test_batched_dynamics_max_steps_reports_failure():
    """A dynamics job whose attempts run out (max_steps) ends as not converged
    with the reference's message; the other jobs of the batch are unaffected."""
    c = pimi.compile(_c2_model(10))
    obs = _dyn_observables(c)
    eng = pimi.BatchEngine(precision=N.PIMI_C64, memory_bytes=1 << 28)
    try:
        bad = eng.submit_dynamics(c, _excited(c), (0.0, 50.0), saveat=25.0, observables=obs,
                                  algorithm=pimi.ExpvAlgorithm(krylovdim=4, rtol=1e-12, max_steps=3))
        good = eng.submit_dynamics(c, _excited(c), (0.0, 1.0), saveat=0.5, observables=obs,
                                   algorithm=pimi.ExpvAlgorithm(krylovdim=30, rtol=1e-10))
        eng.run()
        rb = eng.fetch_dynamics(bad, check=False)
        assert not rb.info["converged"] and rb.info["iterations"] == 3
        with pytest.raises(pimi.NotConverged):
            eng2 = pimi.BatchEngine(precision=N.PIMI_C64, memory_bytes=1 << 28)
            try:
                j = eng2.submit_dynamics(c, _excited(c), (0.0, 50.0), saveat=25.0, observables=obs,
                                         algorithm=pimi.ExpvAlgorithm(krylovdim=4, rtol=1e-12, max_steps=3))
                eng2.run()
                eng2.fetch_dynamics(j)
            finally:
                eng2.close()
        rg = eng.fetch_dynamics(good)
        assert rg.info["converged"]
    finally:
        eng.close()
