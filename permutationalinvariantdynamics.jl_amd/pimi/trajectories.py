"""Conditional (no-jump) trajectory cohorts on the device: the batched
trajectory primitives of src/batched_trajectories.jl:266-598
(batched_conditional_action!, batched_conditional_rk4!,
batched_channel_intensities!, batched_apply_jumps!).  A cohort is an n x ncols
matrix of PI states that share one time step.  Jump sampling
(batched_trajectory_rngs, batched_sample_jumps; src/batched_trajectories.jl:609-669)
is host code drawing from one private stream per trajectory.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .api import CompiledPIModel


def _cols(compiled, X):
    X = np.asarray(X, dtype=np.complex128)
    if X.ndim == 1:
        X = X[:, None]
    if X.shape[0] != len(compiled.basis):
        raise ValueError("batched trajectory matrix has the wrong leading dimension")
    return np.asfortranarray(X)


def trajectory_channels(compiled: CompiledPIModel) -> int:
    c = C.c_int32()
    N.check(N.lib().pimi_trajectory_channels(compiled._plan, C.byref(c)))
    return c.value


def batched_conditional_action(compiled: CompiledPIModel, X):
    """(Y, intensities): the normalized conditional drift of every column and
    its total jump intensity."""
    X = _cols(compiled, X)
    Y = np.empty_like(X)
    lam = np.zeros(X.shape[1])
    N.check(N.lib().pimi_conditional_action(compiled.workspace(N.PIMI_C64), N.ptr(X), N.ptr(Y),
                                            C.c_int64(X.shape[1]), N.ptr(lam), C.c_int32(0)))
    return Y, lam


def batched_conditional_rk4(compiled: CompiledPIModel, X, step):
    """One fixed RK4 step of every column, each then trace-normalized (a new array)."""
    X = np.array(_cols(compiled, X), order="F", copy=True)
    N.check(N.lib().pimi_conditional_rk4(compiled.workspace(N.PIMI_C64), N.ptr(X), C.c_int64(X.shape[1]),
                                         C.c_double(step), C.c_int32(0)))
    return X


def batched_channel_intensities(compiled: CompiledPIModel, X):
    """rates[k, c]: intensity of jump channel k+1 (model term order) in column c."""
    X = _cols(compiled, X)
    nch = trajectory_channels(compiled)
    rates = np.zeros((nch, X.shape[1]), order="F")
    N.check(N.lib().pimi_channel_intensities(compiled.workspace(N.PIMI_C64), N.ptr(X), C.c_int64(X.shape[1]),
                                             N.ptr(rates), C.c_int32(0)))
    return rates


def batched_apply_jumps(compiled: CompiledPIModel, X, channels):
    """Apply the selected jump (1-based channel, 0 = none) to every column (a new array)."""
    X = np.array(_cols(compiled, X), order="F", copy=True)
    ch = np.ascontiguousarray(channels, dtype=np.int32)
    if ch.shape != (X.shape[1],):
        raise ValueError("channels must contain one entry per batch column")
    N.check(N.lib().pimi_apply_jumps(compiled.workspace(N.PIMI_C64), N.ptr(X), C.c_int64(X.shape[1]), N.ptr(ch),
                                     C.c_int32(0)))
    return X


def batched_trajectory_rngs(seed: int, count: int):
    """One independent random stream per trajectory index
    (batched_trajectory_rngs, src/batched_trajectories.jl:609-618): the
    master seed expands into `count` child streams (numpy SeedSequence spawn,
    PCG64), one per trajectory, kept with its trajectory when columns are
    regrouped.  The reference expands a MersenneTwister(seed) into per-index
    MersenneTwister streams; the streams differ, the contract (one private
    stream per trajectory, a fixed draw pattern) is the same."""
    if isinstance(count, bool) or int(count) != count:
        raise ValueError("trajectory RNG count must be an integer, not Bool")
    if count <= 0:
        raise ValueError("trajectory RNG count must be positive")
    return [np.random.Generator(np.random.PCG64(s)) for s in np.random.SeedSequence(int(seed)).spawn(int(count))]


def _select_jump_channel(rates, u):
    """First channel whose cumulative rate exceeds u (1-based); a trailing
    zero channel is never chosen (src/trajectories.jl:796-809)."""
    cumulative = 0.0
    last_positive = 0
    for index, rate in enumerate(rates, start=1):
        if rate > 0:
            last_positive = index
        cumulative += rate
        if u < cumulative:
            return index
    if last_positive == 0:
        raise RuntimeError("cannot select a jump channel from zero total intensity")
    return last_positive


def batched_sample_jumps(channels, rates, step, rngs):
    """One fixed-step jump decision per column from channel intensities
    rates[k, c] (batched_sample_jumps!, src/batched_trajectories.jl:632-669):
    column c jumps with probability 1 - exp(-lambda_c h), lambda_c the sum of
    its rates, using one draw from rngs[c]; a jump selects channel k with
    probability rates[k, c]/lambda_c from a second draw.  channels[c] is set
    to the 1-based channel, or 0 for no jump.  step is used as given (never
    capped).  Host code: the draws are sequential per stream, as in the
    reference; the rates come from batched_channel_intensities."""
    rates = np.asarray(rates)
    if rates.ndim != 2:
        raise ValueError("rates must be a (channels x columns) matrix")
    columns = rates.shape[1]
    if len(channels) != columns:
        raise ValueError("channels must contain one entry per rate column")
    if isinstance(channels, np.ndarray) and channels.dtype == np.bool_:
        raise ValueError("jump-decision storage must contain integer indices, not Bool")
    if len(rngs) < columns:
        raise ValueError("one RNG stream is required per rate column")
    if not np.issubdtype(rates.dtype, np.floating):
        raise ValueError("batched jump intensities must use a real floating scalar type")
    if isinstance(step, bool):
        raise ValueError("batched trajectory step must be a real number, not Bool")
    h = float(step)
    if not (np.isfinite(h) and h > 0):
        raise ValueError("batched trajectory step must be positive")
    for c in range(columns):
        col = rates[:, c]
        if not np.all(np.isfinite(col)):
            raise ValueError(f"jump intensity is nonfinite in batch column {c + 1}")
        if np.any(col < 0):
            raise ValueError(f"jump intensity is negative in batch column {c + 1}")
        lam = float(np.sum(col))
        if not np.isfinite(lam):
            raise ValueError(f"total jump intensity overflows in batch column {c + 1}")
        probability = -np.expm1(-lam * h)
        if lam > 0 and rngs[c].random() < probability:
            channels[c] = _select_jump_channel(col, rngs[c].random() * lam)
        else:
            channels[c] = 0
    return channels
