"""Precision form: spin-1 ensemble N=18 (PI dimension 1,562,275, blocks up to
442).  Complex FP32 Arnoldi/apply inside FP64 refinement (inner residual
reduction 1e-5 per cycle, outer relative residual 1e-9), against the FP64
oracle's stationary state (tests/golden/c3_reference.json, written by
make_c3_reference.py); the error of each observable is reported and bounded.
The plain FP64 device path is checked the same way.
"""
import json
import os
import time

import numpy as np
import pytest

import pimi

pytestmark = pytest.mark.gpu
S3 = pimi.spin_matrices(3)
N_ = 18


def _model():
    B = pimi.PIBasis(N_, 3)
    Jz = pimi.collective_operator(B, S3.jz)
    return pimi.PIModel(B, [pimi.LocalHamiltonian(1.0 * S3.jz + 0.7 * S3.jx),
                            pimi.DirectPIHamiltonian(Jz * Jz, rate=0.3 / N_), pimi.LocalJump(S3.jm, rate=0.1)])


def _observables(comp, rho):
    b = comp.basis
    jz = comp.collective_operator(S3.jz).data
    jp = comp.collective_operator(S3.jp).data
    jm = comp.collective_operator(S3.jm).data
    idx, tv = comp.trace_functional()
    jj = np.zeros_like(jp)
    k = 0
    for s, m in enumerate(b.dims):   # (A B) coefficient blocks = A_nu B_nu / sqrt(f)
        o0, o1 = b.offsets[s], b.offsets[s + 1]
        jj[o0:o1] = ((jp[o0:o1].reshape(m, m, order="F") @ jm[o0:o1].reshape(m, m, order="F")) / tv[k]).reshape(
            -1, order="F")
        k += m
    return np.vdot(jz, rho).real / N_, np.vdot(jj, rho).real / N_ ** 2


def test_spin1_precision_form_against_fp64_oracle():
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "c3_reference.json")))
    comp = pimi.compile(_model()).prepare()
    assert comp.info["n"] == gold["n"] == 1562275
    runs = {}
    t = time.time()
    r = pimi.batched_stationary_states([comp], pimi.BatchAlgorithm(krylovdim=200, inner_rtol=1e-5), rtol=1e-9,
                                       precision=pimi.PIMI_C32)[0]
    runs["fp32_arnoldi_fp64_refinement"] = (r, time.time() - t)
    t = time.time()
    r2 = pimi.stationary_state(comp, algorithm=pimi.GMRESAlgorithm(200, 2000, "sector_inner", inner_steps=8),
                               atol=0.0, rtol=1e-9, return_info=True, criterion="backward", weighting="sector",
                               precision=pimi.PIMI_C32)
    runs["fgmres_fp32_inner"] = (r2, time.time() - t)
    t = time.time()
    r3 = pimi.stationary_state(comp, algorithm=pimi.GMRESAlgorithm(400, 2000, "sector"), atol=0.0, rtol=1e-9,
                               return_info=True, criterion="backward", weighting="sector")
    runs["plain_fp64"] = (r3, time.time() - t)
    for name, (res, secs) in runs.items():
        jz, jj = _observables(comp, res.state.data)
        e1 = abs(jz - gold["jz_over_N"]) / abs(gold["jz_over_N"])
        e2 = abs(jj - gold["jpjm_over_N2"]) / abs(gold["jpjm_over_N2"])
        print(f"  {name}: {secs:.2f} s, iterations {res.info['iterations']}, relative residual "
              f"{res.info['relative_residual']:.2e}, <Jz>/N err {e1:.2e}, <J+J->/N^2 err {e2:.2e}", flush=True)
        assert res.info["relative_residual"] <= 1e-9
        assert e1 <= 1e-7 and e2 <= 1e-7, (name, e1, e2)
