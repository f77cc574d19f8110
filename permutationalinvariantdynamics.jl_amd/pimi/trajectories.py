"""Conditional (no-jump) trajectory cohorts on the device: the batched
trajectory primitives of src/batched_trajectories.jl:266-598
(batched_conditional_action!, batched_conditional_rk4!,
batched_channel_intensities!, batched_apply_jumps!).  A cohort is an n x ncols
matrix of PI states that share one time step; jump sampling (the reference's
per-trajectory MersenneTwister streams) stays with the caller.
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
