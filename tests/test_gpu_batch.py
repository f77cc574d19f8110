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
    out = pimi.batched_stationary_states(comps, alg, rtol=1e-9, engine=eng)
    eng.close()
    for a, b in zip(ref, out):
        np.testing.assert_array_equal(a.state.data, b.state.data)
