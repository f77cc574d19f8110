"""Arnoldi spectra and fixed-step RK4 evolution on the device.

krylov_liouvillian_spectrum follows src/krylov.jl:914-961: the Arnoldi
factorization (the large basis, the applies and the orthogonalization) runs
in libpimi (pimi_arnoldi); the small projected eigenproblem, the Ritz
ordering (src/krylov.jl:893-898) and the residual estimates are host work on
the (k+1) x k Hessenberg matrix, as in the reference.

evolve / time_evolve follow src/evolution.jl:75-130: fixed-step RK4 with the
reference's three-scratch stage arithmetic, every stage on the device
(pimi_evolve_rk4).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .api import CompiledPIModel, PIState

_EPS = np.finfo(np.float64).eps
_TINY = np.finfo(np.float64).tiny


def _ritz_order(values, which):
    """src/krylov.jl:893-898 (stable sorts, as Julia's sortperm)."""
    if which == "LR":
        return np.argsort(-values.real, kind="stable")
    if which == "LM":
        return np.argsort(-np.abs(values), kind="stable")
    if which == "SM":
        return np.argsort(np.abs(values), kind="stable")
    raise ValueError("which must be 'LR', 'LM', or 'SM'")


def krylov_liouvillian_spectrum(compiled: CompiledPIModel, nev=6, krylovdim=None, which="LR", initial_vector=None,
                                atol=1e-10, rtol=1e-8, vectors=False, rng=None, require_convergence=True):
    """Selected Liouvillian eigenvalues by single-vector Arnoldi; returns a
    dict with the reference's info fields (values, residuals, converged,
    iterations, krylov_dimension, dimension, which, residual_scale,
    residual_tolerance, and vectors when requested).  The random start
    vector comes from numpy (seeded by `rng`), not Julia's RNG stream."""
    n = len(compiled.basis)
    if not 0 < nev <= n:
        raise ValueError("nev must lie between 1 and the operator dimension")
    if which not in ("LR", "LM", "SM"):
        raise ValueError("which must be 'LR', 'LM', or 'SM'")
    kd = max(20, 2 * nev + 4) if krylovdim is None else int(krylovdim)
    m = min(n, max(kd, nev + 1))
    if initial_vector is None:
        g = rng if isinstance(rng, np.random.Generator) else np.random.default_rng(rng)
        q = (g.standard_normal(n) + 1j * g.standard_normal(n)) / np.sqrt(2)
    else:
        q = np.asarray(initial_vector, dtype=np.complex128)
        if q.shape != (n,):
            raise ValueError("initial_vector has wrong length")
    q = np.ascontiguousarray(q)
    ws = compiled.workspace(N.PIMI_C64, krylovdim=max(m, 30))
    H = np.zeros((m + 1) * m, dtype=np.complex128)
    k = C.c_int32()
    N.check(N.lib().pimi_arnoldi(ws, N.ptr(q), C.c_int32(m), C.c_double(_EPS), N.ptr(H), C.byref(k)))
    k = k.value
    Hm = H.reshape(m, m + 1).T  # (m+1) x m, column-major on the wire
    vals, Y = np.linalg.eig(Hm[:k, :k])
    order = _ritz_order(vals, which)[:min(nev, k)]
    vals, Y = vals[order], Y[:, order]
    residuals = np.abs(Hm[k, k - 1] * Y[-1, :])
    scale = max(max((np.linalg.norm(Hm[:k + 1, j]) for j in range(k)), default=0.0), _TINY)
    tolerance = atol + rtol * scale
    converged = residuals <= tolerance
    if require_convergence and not converged.all():
        raise ValueError("Arnoldi spectrum did not converge; maximum requested Ritz residual="
                         f"{residuals.max()}; increase krylovdim")
    info = dict(values=vals, residuals=residuals, converged=converged, iterations=k, krylov_dimension=m,
                dimension=n, which=which, residual_scale=scale, residual_tolerance=tolerance,
                workspace_reused=False)
    if vectors:
        Yf = np.ascontiguousarray(Y.T)  # column c contiguous
        out = np.zeros((Y.shape[1], n), dtype=np.complex128)
        N.check(N.lib().pimi_arnoldi_vectors(ws, C.c_int32(k), N.ptr(Yf), C.c_int32(Y.shape[1]), N.ptr(out)))
        info["vectors"] = out.T
    return info


def evolve(dest: PIState, compiled: CompiledPIModel, src: PIState, tspan, steps=256):
    """evolve!(dest, L, src, tspan; steps): RK4 in `steps` equal steps; dest
    may be src (exact in-place use).  Returns dest."""
    if dest.basis != src.basis or src.basis != compiled.basis:
        raise ValueError("source and destination use incompatible PI bases")
    if int(steps) <= 0:
        raise ValueError("steps must be positive")
    t0, t1 = (float(v) for v in tspan)
    x0 = np.ascontiguousarray(src.data, dtype=np.complex128)
    out = np.empty_like(x0)
    ws = compiled.workspace(N.PIMI_C64)
    N.check(N.lib().pimi_evolve_rk4(ws, N.ptr(x0), C.c_double(t0), C.c_double(t1), C.c_int64(int(steps)),
                                    N.ptr(out)))
    dest.data[:] = out
    return dest


def time_evolve(compiled: CompiledPIModel, rho: PIState, tspan, steps=256) -> PIState:
    """The PI state obtained by propagating rho over tspan (src/evolution.jl:128-130)."""
    out = PIState(rho.basis, np.array(rho.data, dtype=np.complex128))
    return evolve(out, compiled, rho, tspan, steps=steps)
