"""Serving loop: a Poisson stream of mixed stationary and dynamics requests.

Each request's result is the same alone and in the mix (to the bit), the
stationary states meet the relative residual and the closed form of the
local-only model family, and the dynamics follow the closed form
<sum_i n_i>(t) = N sigma_ee(t).
"""
import os
import sys

import numpy as np
import pytest

import pimi
from pimi.serving import Server, summarize
from pimi.workloads import Request, served_requests

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_closed_form as CF  # noqa: E402

pytestmark = pytest.mark.gpu


def _closed(inst, ts=None):
    H = 2 * inst.omega * CF.jx + 1.0 * CF.jz
    jumps = [(inst.gamma, CF.jm)] + ([(4 * inst.gamma_phi, CF.jz)] if inst.gamma_phi > 0 else [])
    if ts is None:
        return CF.steady(inst.N, H, jumps)[0]
    e = np.array([[0, 0], [0, 1]], dtype=complex)
    return CF.series(inst.N, H, jumps, e, ts)[:, 0]


def test_serving_alone_equals_mix_and_closed_forms():
    reqs = served_requests(rate=500.0, count=10, nmin=16, nmax=40)
    print("serving kinds", [r.kind for r in reqs], "N", [r.instance.N for r in reqs], flush=True)
    # rtol 1e-11: the observable error of a stationary state is the residual
    # times the conditioning (up to ~1e3 on weakly driven members), so the
    # closed-form comparison to 1e-6 needs a residual well below 1e-9
    srv = Server(precision=pimi.PIMI_C32, rtol=1e-11, memory_bytes=4 << 30, dynamics_workers=2,
                 dynamics_memory_bytes=2 << 30, compile_workers=4, t_end=2.0, saveat=0.2)
    try:
        mix = srv.serve(reqs, realtime=True)
        print("summary", summarize(mix), flush=True)
        for r, q in zip(mix, reqs):
            assert r is not None and r.converged, (r, q)
            alone = srv.serve([Request(0.0, q.kind, q.instance)], realtime=False)[0]
            if q.kind == "steady":
                assert r.info["relative_residual"] <= 1e-11
                np.testing.assert_array_equal(alone.state.data, r.state.data)
                comp = pimi.compile(q.instance.model())
                val = pimi.expect(comp.collective_operator(np.array([[0, 0], [0, 1]], dtype=complex)), r.state).real
                ref = _closed(q.instance)
                assert abs(val - ref) <= 1e-6 * max(ref, 1e-3), (q.instance, val, ref)
            else:
                np.testing.assert_array_equal(alone.values["n"], r.values["n"])
                ref = _closed(q.instance, list(r.times))
                np.testing.assert_allclose(r.values["n"].real, ref, rtol=1e-6, atol=1e-9)
    finally:
        srv.close()


def test_serving_large_n_fp32_against_closed_form():
    """The top of the served size range, where FP32 range hazards live
    (t.t = 2^N overflows FP32 from N = 128, f^nu reaches 8.8e45 at N = 160):
    FP32-Arnoldi stationary jobs at N = 128 and 160 and an N = 160 dynamics
    request to t = 10 (save every 0.2), against the product-state closed forms."""
    from pimi.workloads import Instance
    import parity_log
    reqs = [Request(0.0, "steady", Instance(128, 0.7, 0.1, 0.02)),
            Request(0.0, "steady", Instance(160, 1.3, 0.15, 0.05)),
            Request(0.0, "dynamics", Instance(160, 0.9, 0.12, 0.03))]
    srv = Server(precision=pimi.PIMI_C32, rtol=1e-11, memory_bytes=8 << 30, dynamics_workers=1,
                 dynamics_memory_bytes=4 << 30, compile_workers=3, t_end=10.0, saveat=0.2)
    try:
        out = srv.serve(reqs, realtime=False)
        for r, q in zip(out, reqs):
            assert r is not None and r.converged, (r, q)
            if q.kind == "steady":
                assert r.info["relative_residual"] <= 1e-11
                comp = pimi.compile(q.instance.model())
                val = pimi.expect(comp.collective_operator(np.array([[0, 0], [0, 1]], dtype=complex)),
                                  r.state).real
                ref = _closed(q.instance)
                err = abs(val - ref) / ref
                parity_log.record(f"served_fp32_steady_N{q.instance.N}_n", err, 1e-6)
                assert err <= 1e-6, (q.instance, val, ref)
            else:
                assert len(r.times) == 51
                ref = _closed(q.instance, list(r.times))
                err = np.abs(r.values["n"].real - ref).max() / np.abs(ref).max()
                parity_log.record(f"served_dynamics_N{q.instance.N}_n", err, 1e-6)
                assert err <= 1e-6
    finally:
        srv.close()


def test_serving_snapshots_reload_bit_for_bit(tmp_path):
    """With snapshot_dir every finished request leaves a PIDCHKPT v1 file whose
    state reloads bit for bit, with the request's parameters as metadata."""
    from pimi.workloads import Instance
    reqs = [Request(0.0, "steady", Instance(12, 0.6, 0.1, 0.02)), Request(0.0, "dynamics", Instance(9, 0.8, 0.1, 0.0))]
    srv = Server(precision=pimi.PIMI_C64, rtol=1e-10, memory_bytes=1 << 30, dynamics_workers=1,
                 dynamics_memory_bytes=1 << 30, compile_workers=2, t_end=1.0, saveat=0.5,
                 snapshot_dir=str(tmp_path))
    try:
        out = srv.serve(reqs, realtime=False)
    finally:
        srv.close()
    for r, q in zip(out, reqs):
        assert r.snapshot and os.path.exists(r.snapshot)
        cp = pimi.load_checkpoint(r.snapshot)
        np.testing.assert_array_equal(cp.state.data, r.state.data)
        assert cp.metadata["kind"] == q.kind and int(cp.metadata["N"]) == q.instance.N
        assert float(cp.metadata["omega"]) == q.instance.omega
        assert (cp.time is None) == (q.kind == "steady")


def test_serving_dynamics_engine_matches_workers():
    """Dynamics requests batched in the FP64 dynamics engine (the default)
    against the same requests on worker streams (pimi_expv_series): the
    observable series agree to 1e-10 relative, and every request's series
    follows the closed form."""
    from pimi.workloads import Instance
    import parity_log
    reqs = [Request(0.0, "dynamics", Instance(n_, om, 0.1, gp))
            for n_, om, gp in ((12, 0.7, 0.02), (30, 1.1, 0.0), (57, 0.4, 0.05))]
    out = {}
    for mode in ("engine", "workers"):
        srv = Server(precision=pimi.PIMI_C64, rtol=1e-11, memory_bytes=1 << 30, dynamics_workers=2,
                     dynamics_memory_bytes=2 << 30, compile_workers=3, t_end=3.0, saveat=0.25, dynamics_mode=mode)
        try:
            out[mode] = srv.serve(reqs, realtime=False)
        finally:
            srv.close()
    worst = 0.0
    for a, b, q in zip(out["engine"], out["workers"], reqs):
        assert a.converged and b.converged
        np.testing.assert_array_equal(a.times, b.times)
        worst = max(worst, float(np.abs(a.values["n"] - b.values["n"]).max() / np.abs(b.values["n"]).max()))
        ref = _closed(q.instance, list(a.times))
        np.testing.assert_allclose(a.values["n"].real, ref, rtol=1e-8, atol=1e-10)
    parity_log.record("served_dynamics_engine_vs_workers", worst, 1e-10)
    assert worst <= 1e-10, worst
