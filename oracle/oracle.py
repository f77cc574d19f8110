"""Oracle front end: exact lowering (lowering.py) + C numerics (pi_oracle.c).

TEST INFRASTRUCTURE ONLY -- used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker, never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import lowering as low

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libpi_oracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _LIB.oracle_steady_state.restype = C.c_int
        _LIB.oracle_expv.restype = C.c_int
        _LIB.oracle_solve_dynamics.restype = C.c_int
        _LIB.oracle_operator_scale.restype = C.c_double
    return _LIB


class _CSCSet(C.Structure):
    _fields_ = [("cp_base", C.c_void_p), ("colptr", C.c_void_p), ("rows", C.c_void_p), ("vals", C.c_void_p)]


class _Plan(C.Structure):
    _fields_ = [
        ("ns", C.c_int32), ("dims", C.c_void_p), ("offs", C.c_void_p), ("n", C.c_int64),
        ("has_h", C.c_int32), ("has_q", C.c_int32), ("H", _CSCSet), ("Q", _CSCSet),
        ("ncg", C.c_int32), ("cg_scale", C.c_void_p), ("K", _CSCSet),
        ("nob", C.c_int32), ("ob_scale", C.c_void_p), ("ob_br_base", C.c_void_p),
        ("br_out", C.c_void_p), ("br_in", C.c_void_p), ("br_r", C.c_void_p),
        ("br_use_support", C.c_void_p), ("br_sup_base", C.c_void_p),
        ("sup_rows", C.c_void_p), ("sup_cols", C.c_void_p), ("sup_vals", C.c_void_p),
        ("br_dense_base", C.c_void_p), ("dense", C.c_void_p), ("max_dim", C.c_int64),
    ]


class _SteadyInfo(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("restarts", C.c_int64), ("applications", C.c_int64),
                ("converged", C.c_int32), ("residual", C.c_double), ("normalized_residual", C.c_double),
                ("trace_error", C.c_double), ("operator_scale", C.c_double), ("linear_residual", C.c_double),
                ("weighted_relative_residual", C.c_double)]


class _ExpvInfo(C.Structure):
    _fields_ = [("converged", C.c_int32), ("accepted", C.c_int64), ("rejected", C.c_int64),
                ("applications", C.c_int64), ("factorizations", C.c_int64), ("attempts", C.c_int64),
                ("reached", C.c_double), ("estimated_error", C.c_double)]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _cscset(blocks):
    cp_base, colptr, rows, vals = [], [], [], []
    for B in blocks:
        cp_base.append(len(colptr))
        base = len(rows)
        colptr.extend(base + p for p in B.colptr)
        rows.extend(B.rows)
        vals.extend(B.vals)
    arrs = (np.array(cp_base, dtype=np.int64), np.array(colptr, dtype=np.int64),
            np.array(rows, dtype=np.int32), np.array(vals, dtype=np.complex128))
    return arrs


class OraclePlan:
    """Flattened FusedStaticPIKernel handed to the C oracle."""

    def __init__(self, plan: low.FusedPlan):
        b = plan.basis
        self.basis = b
        self.plan = plan
        keep = []
        ns = len(b.sectors)
        dims = np.array(b.dims, dtype=np.int64)
        offs = np.array(b.offsets[:-1], dtype=np.int64)
        keep += [dims, offs]
        P = _Plan()
        P.ns = ns
        P.dims = _ptr(dims)
        P.offs = _ptr(offs)
        P.n = b.n
        P.max_dim = int(max(b.dims))

        def setcsc(field, blocks):
            arrs = _cscset(blocks)
            keep.extend(arrs)
            setattr(P, field, _CSCSet(_ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]), _ptr(arrs[3])))

        P.has_h = plan.H is not None
        P.has_q = plan.Q is not None
        empty = [low.CSC(m, m, [0] * (m + 1), [], []) for m in b.dims]
        setcsc("H", plan.H if plan.H is not None else empty)
        setcsc("Q", plan.Q if plan.Q is not None else empty)
        P.ncg = len(plan.collective_gains)
        cg_scale = np.array([g[0] for g in plan.collective_gains] or [0], dtype=np.complex128)
        keep.append(cg_scale)
        P.cg_scale = _ptr(cg_scale)
        setcsc("K", [blk for g in plan.collective_gains for blk in g[1]] or empty)
        P.nob = len(plan.onebody_gains)
        ob_scale = np.array([g[0] for g in plan.onebody_gains] or [0], dtype=np.complex128)
        base = [0]
        br_out, br_in, br_r, use, sup_base, sup_rows, sup_cols, sup_vals, dbase, dense = (
            [], [], [], [], [0], [], [], [], [], [])
        dense_len = 0
        for _, branches in plan.onebody_gains:
            for (li, ni, r, Cm, u, sup) in branches:
                br_out.append(li)
                br_in.append(ni)
                br_r.append(r)
                use.append(1 if u else 0)
                for (row, col, v) in (sup if u else ()):
                    sup_rows.append(row)
                    sup_cols.append(col)
                    sup_vals.append(v)
                sup_base.append(len(sup_rows))
                dbase.append(dense_len)
                if not u:
                    dense.append(np.asarray(Cm, dtype=np.complex128).reshape(-1, order="F"))
                    dense_len += Cm.size
            base.append(len(br_out))
        arrs = dict(
            ob_scale=ob_scale, ob_br_base=np.array(base, dtype=np.int64),
            br_out=np.array(br_out or [0], dtype=np.int32), br_in=np.array(br_in or [0], dtype=np.int32),
            br_r=np.array(br_r or [0], dtype=np.float64), br_use_support=np.array(use or [0], dtype=np.int32),
            br_sup_base=np.array(sup_base, dtype=np.int64), sup_rows=np.array(sup_rows or [0], dtype=np.int32),
            sup_cols=np.array(sup_cols or [0], dtype=np.int32),
            sup_vals=np.array(sup_vals or [0], dtype=np.complex128),
            br_dense_base=np.array(dbase or [0], dtype=np.int64),
            dense=(np.concatenate(dense) if dense else np.zeros(1, dtype=np.complex128)))
        for k, v in arrs.items():
            keep.append(v)
            setattr(P, k, _ptr(v))
        self._keep = keep
        self._P = P
        t_idx, t_val = low.trace_functional(b)
        self.t_idx = t_idx
        self.t_val = t_val

    @property
    def n(self):
        return self.basis.n

    def apply_diag(self, x, nthreads=1):
        """y = D x, D the sector-diagonal part of L (the inner solver's operator)."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty_like(x)
        lib().oracle_apply_diag(C.byref(self._P), _ptr(x), _ptr(y), C.c_int(nthreads))
        return y

    def apply(self, x, nthreads=1):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty_like(x)
        if nthreads > 1:
            lib().oracle_apply_threaded(C.byref(self._P), _ptr(x), _ptr(y), C.c_int(nthreads))
        else:
            lib().oracle_apply(C.byref(self._P), _ptr(x), _ptr(y))
        return y

    def operator_scale(self, nthreads=1):
        return lib().oracle_operator_scale(C.byref(self._P), C.c_int(nthreads))

    def sector_preconditioner(self):
        """(U, inv_den) in PI layout from LAPACK eigh of each fused
        Hamiltonian block (independent of the product's eigensolver)."""
        if getattr(self, "_pc", None) is None:
            b = self.basis
            U = np.zeros(b.n, dtype=np.complex128)
            inv = np.zeros(b.n, dtype=np.complex128)
            for s, m in enumerate(b.dims):
                Hs = self.plan.H[s].dense() if self.plan.H is not None else np.zeros((m, m), complex)
                Qs = self.plan.Q[s].dense() if self.plan.Q is not None else np.zeros((m, m), complex)
                Hs = 0.5 * (Hs + Hs.conj().T)
                d, Us = np.linalg.eigh(Hs)
                q = np.real(np.einsum("ka,kc,ca->a", Us.conj(), Qs, Us))
                den = -1j * (d[:, None] - d[None, :]) - 0.5 * (q[:, None] + q[None, :])
                scale = max(np.max(np.abs(d) + np.abs(q)), 1e-300)
                iv = np.where(np.abs(den) > 1e-13 * scale, 1.0 / np.where(den == 0, 1, den), 0)
                o0, o1 = b.offsets[s], b.offsets[s + 1]
                U[o0:o1] = Us.reshape(-1, order="F")
                inv[o0:o1] = iv.reshape(-1, order="F")
            self._pc = (U, inv)
        return self._pc

    def sector_sigma(self):
        """Per-coordinate sqrt(f^nu): the trace weight of the coordinate's sector."""
        b = self.basis
        sig = np.empty(self.n)
        k = 0
        for s, m in enumerate(b.dims):
            sig[b.offsets[s]:b.offsets[s + 1]] = self.t_val[k]
            k += m
        return sig

    def steady_state(self, krylovdim=30, maxiter=500, atol=0.0, rtol=1e-10, operator_scale=0.0, nthreads=1,
                     preconditioner=None, criterion="reference", weighting=None, inner_steps=0,
                     initial_state=None):
        """weighting="sector": iterate on u = S x with S = sqrt(f^nu) per sector;
        criterion="weighted" (with it) accepts on |S L x|/(s |S x|).
        preconditioner="sector_inner": flexible GMRES whose directions are
        `inner_steps` (default 8) sector-preconditioned GMRES steps on the
        sector-diagonal part of L."""
        if preconditioner == "sector_inner":
            preconditioner, inner_steps = "sector", (inner_steps or 8)
        x = np.zeros(self.n, dtype=np.complex128)
        sig = None
        if weighting == "sector":
            if criterion not in ("backward", "weighted"):
                raise ValueError("sector weighting needs the backward or the weighted criterion")
            sig = np.ascontiguousarray(self.sector_sigma())
        elif weighting is not None:
            raise ValueError("unknown weighting")
        x0 = None if initial_state is None else np.ascontiguousarray(initial_state, dtype=np.complex128)
        info = _SteadyInfo()
        pu = pi = None
        if preconditioner == "sector":
            U, inv = self.sector_preconditioner()
            pu, pi = _ptr(U), _ptr(inv)
        elif preconditioner is not None:
            raise ValueError("unknown preconditioner")
        lib().oracle_steady_state(C.byref(self._P), C.c_int(nthreads), _ptr(self.t_idx), _ptr(self.t_val),
                                  C.c_int64(len(self.t_idx)), C.c_int(krylovdim), C.c_int64(maxiter),
                                  C.c_double(atol), C.c_double(rtol), C.c_double(operator_scale), _ptr(x),
                                  C.byref(info), pu, pi, C.c_int({"backward": 1, "weighted": 2}.get(criterion, 0)),
                                  None if sig is None else _ptr(sig), C.c_int(int(inner_steps)),
                                  None if x0 is None else _ptr(x0))
        return x, {k: getattr(info, k) for k, _ in _SteadyInfo._fields_}

    def expv(self, b, t, krylovdim=30, atol=1e-10, rtol=1e-8, safety=0.9, max_steps=10000, nthreads=1):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        y = np.empty_like(b)
        info = _ExpvInfo()
        lib().oracle_expv(C.byref(self._P), C.c_int(nthreads), _ptr(y), _ptr(b), C.c_double(t),
                          C.c_int(krylovdim), C.c_double(atol), C.c_double(rtol), C.c_double(safety),
                          C.c_int64(max_steps), C.byref(info))
        return y, {k: getattr(info, k) for k, _ in _ExpvInfo._fields_}

    def solve_dynamics(self, rho0, ts, observables, krylovdim=30, atol=1e-10, rtol=1e-8, safety=0.9,
                       max_steps=10000, nthreads=1):
        rho = np.array(rho0, dtype=np.complex128, copy=True)
        ts = np.ascontiguousarray(ts, dtype=np.float64)
        obs = np.ascontiguousarray(np.array(observables, dtype=np.complex128).reshape(len(observables), -1))
        out = np.zeros((len(ts), len(observables)), dtype=np.complex128)
        info = _ExpvInfo()
        lib().oracle_solve_dynamics(C.byref(self._P), C.c_int(nthreads), _ptr(rho), _ptr(ts),
                                    C.c_int64(len(ts)), _ptr(obs), C.c_int32(len(observables)), _ptr(out),
                                    C.c_int(krylovdim), C.c_double(atol), C.c_double(rtol),
                                    C.c_double(safety), C.c_int64(max_steps), C.byref(info))
        return out, rho, {k: getattr(info, k) for k, _ in _ExpvInfo._fields_}


def expm(A):
    A = np.asfortranarray(A, dtype=np.complex128)
    q = A.shape[0]
    E = np.zeros((q, q), dtype=np.complex128, order="F")
    lib().oracle_expm(C.c_int(q), _ptr(A), _ptr(E))
    return E


def compile_plan(N, d, terms):
    basis = low.Basis(N, d)
    geom = low.OneBodyGeometry(basis)
    plan = low.compile_model(basis, terms, geom)
    return OraclePlan(plan), geom


def dense_liouvillian(plan: OraclePlan):
    """L as a dense n x n matrix, column k = L e_k from the oracle's apply
    (small n only)."""
    n = plan.n
    L = np.zeros((n, n), dtype=np.complex128)
    e = np.zeros(n, dtype=np.complex128)
    for k in range(n):
        e[k] = 1.0
        L[:, k] = plan.apply(e)
        e[k] = 0.0
    return L


def rk4(plan: OraclePlan, x, t0, t1, steps):
    """Fixed-step RK4, restating evolve! (src/evolution.jl:75-118): three
    scratch vectors, k2 accumulates k1 + 2k2 + 2k3 + k4 and x += (h/6) k2."""
    h = (t1 - t0) / steps
    x = np.array(x, dtype=np.complex128)
    for _ in range(steps):
        k1 = plan.apply(x)
        k2 = k1.copy()
        tmp = x + (h / 2) * k1
        k1 = plan.apply(tmp)
        k2 = k2 + 2 * k1
        tmp = x + (h / 2) * k1
        k1 = plan.apply(tmp)
        k2 = k2 + 2 * k1
        tmp = x + h * k1
        k1 = plan.apply(tmp)
        k2 = k2 + k1
        x = x + (h / 6) * k2
    return x
