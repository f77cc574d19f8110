"""Versioned PI state checkpoints in the reference's portable binary layout
(PIDCHKPT v1, src/checkpoints.jl:1-160), so that states move between this
backend and the reference package.

Layout (little-endian): the 8 magic bytes "PIDCHKPT"; UInt16 schema version
(1); UInt8 scalar code (16/32/64 = Float16/32/64); Int64 N; Int32 d; Int32
sector count, then each retained sector's d parts as Int64; UInt8 time flag
and, when 1, the time in the scalar type; Int64 coordinate count, then (real,
imaginary) pairs in the scalar type; Int32 metadata count, then (key, value)
string pairs sorted by key, each string an UInt32 byte length and UTF-8
bytes.  BigFloat checkpoints (code 128) carry arbitrary-precision decimal
strings; numpy has no such type, so they are rejected on save and load.
"""
from __future__ import annotations

import io
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from .api import PIBasis, PIState

PI_CHECKPOINT_VERSION = 1
_MAGIC = b"PIDCHKPT"
_CODES = {16: np.float16, 32: np.float32, 64: np.float64}
_COMPLEX = {np.dtype(np.float16): None, np.dtype(np.float32): np.complex64, np.dtype(np.float64): np.complex128}


@dataclass
class PIStateCheckpoint:
    schema_version: int
    state: PIState
    time: object = None
    metadata: dict = field(default_factory=dict)


def _real_type(state: PIState):
    return {np.dtype(np.complex64): np.float32, np.dtype(np.complex128): np.float64}[state.data.dtype]


def make_checkpoint(state: PIState, time=None, metadata=None) -> PIStateCheckpoint:
    """PIStateCheckpoint(state; time, metadata): a detached copy; metadata keys
    and values are stored as strings; the time must be exactly representable
    in the state's precision."""
    R = _real_type(state)
    stored = None
    if time is not None:
        if isinstance(time, (complex, np.complexfloating)) or not np.isreal(time):
            raise ValueError("checkpoint time must be real")
        stored = R(time)
        if not np.isfinite(stored):
            raise ValueError(f"checkpoint time is not finite in state precision {np.dtype(R).name}")
        if float(stored) != float(time):
            raise ValueError(f"checkpoint time cannot be represented exactly in state precision {np.dtype(R).name}")
    md = {str(k): str(v) for k, v in (metadata or {}).items()}
    return PIStateCheckpoint(PI_CHECKPOINT_VERSION, PIState(state.basis, np.array(state.data)), stored, md)


def checkpoint_state(checkpoint: PIStateCheckpoint) -> PIState:
    """A detached PIState (src/checkpoints.jl:43)."""
    return PIState(checkpoint.state.basis, np.array(checkpoint.state.data))


def _write_string(f, s):
    b = s.encode("utf-8")
    if len(b) > 0xFFFFFFFF:
        raise ValueError("checkpoint string is too long")
    f.write(struct.pack("<I", len(b)))
    f.write(b)


def save_checkpoint(path, checkpoint_or_state, time=None, metadata=None, format="auto"):
    """save_checkpoint(path, checkpoint) or save_checkpoint(path, state; time,
    metadata).  Only the portable binary format is available (JLD2/HDF5 need
    the reference's optional packages).  Returns the path."""
    fmt = format
    if fmt == "auto":
        ext = os.path.splitext(str(path))[1].lower()
        fmt = "jld2" if ext in (".jld2", ".jld") else "hdf5" if ext in (".h5", ".hdf5") else "pid"
    if fmt != "pid":
        raise ValueError(f"checkpoint format :{fmt} requires the corresponding optional JLD2 or HDF5 package")
    cp = checkpoint_or_state if isinstance(checkpoint_or_state, PIStateCheckpoint) else \
        make_checkpoint(checkpoint_or_state, time=time, metadata=metadata)
    R = _real_type(cp.state)
    code = {np.float32: 32, np.float64: 64}[R]
    b = cp.state.basis
    buf = io.BytesIO()
    buf.write(_MAGIC)
    buf.write(struct.pack("<HB", cp.schema_version, code))
    buf.write(struct.pack("<qii", b.N, b.d, len(b.sectors)))
    for p in b.sectors:
        buf.write(struct.pack(f"<{b.d}q", *p))
    if cp.time is None:
        buf.write(b"\x00")
    else:
        buf.write(b"\x01")
        buf.write(np.asarray(cp.time, dtype=R).astype(np.dtype(R).newbyteorder("<")).tobytes())
    data = np.asarray(cp.state.data)
    buf.write(struct.pack("<q", data.size))
    pairs = np.empty(2 * data.size, dtype=np.dtype(R).newbyteorder("<"))
    pairs[0::2] = data.real
    pairs[1::2] = data.imag
    buf.write(pairs.tobytes())
    buf.write(struct.pack("<i", len(cp.metadata)))
    for k in sorted(cp.metadata):
        _write_string(buf, k)
        _write_string(buf, cp.metadata[k])
    with open(path, "wb") as f:
        f.write(buf.getvalue())
    return str(path)


def load_checkpoint(path, format="auto") -> PIStateCheckpoint:
    """Read a portable PI checkpoint; every field is validated (magic, version,
    scalar code, basis metadata, time flag, coordinate count, string lengths,
    duplicate keys, trailing data) and violations raise ValueError."""
    fmt = format
    if fmt == "auto":
        ext = os.path.splitext(str(path))[1].lower()
        fmt = "jld2" if ext in (".jld2", ".jld") else "hdf5" if ext in (".h5", ".hdf5") else "pid"
    if fmt != "pid":
        raise ValueError(f"checkpoint format :{fmt} requires the corresponding optional JLD2 or HDF5 package")
    with open(path, "rb") as f:
        raw = f.read()
    pos = 0

    def take(nbytes):
        nonlocal pos
        if pos + nbytes > len(raw):
            raise ValueError("truncated PI checkpoint")
        out = raw[pos:pos + nbytes]
        pos += nbytes
        return out

    def unpack(fmt_):
        return struct.unpack("<" + fmt_, take(struct.calcsize("<" + fmt_)))

    if len(raw) < len(_MAGIC) or raw[:len(_MAGIC)] != _MAGIC:
        raise ValueError("file is not a PI checkpoint")
    take(len(_MAGIC))
    version, code = unpack("HB")
    if version != PI_CHECKPOINT_VERSION:
        raise ValueError(f"unsupported PI checkpoint schema version {version}")
    if code == 128:
        raise ValueError("BigFloat checkpoints need arbitrary-precision scalars, which this package does not provide")
    if code not in _CODES:
        raise ValueError(f"unknown PI checkpoint scalar code {code}")
    R = np.dtype(_CODES[code]).newbyteorder("<")
    if _COMPLEX[np.dtype(_CODES[code])] is None:
        raise ValueError("Float16 checkpoints are not supported by this package's states")
    N_, d, ns = unpack("qii")
    if not (N_ >= 0 and d >= 1 and ns >= 1):
        raise ValueError("invalid PI checkpoint basis metadata")
    sectors = [unpack(f"{d}q") for _ in range(ns)]
    basis = PIBasis(N_, d, sectors=sectors)
    if [tuple(p) for p in sectors] != basis.sectors:
        raise ValueError("PI checkpoint sectors are not in basis order")
    (flag,) = unpack("B")
    if flag not in (0, 1):
        raise ValueError("invalid PI checkpoint time flag")
    time = None
    if flag == 1:
        time = np.frombuffer(take(R.itemsize), dtype=R)[0].astype(_CODES[code])
    (count,) = unpack("q")
    if count != len(basis):
        raise ValueError("PI checkpoint coefficient length does not match its basis")
    pairs = np.frombuffer(take(2 * count * R.itemsize), dtype=R)
    # complex(re, im) exactly, as the reference builds it: re + 1j*im would
    # turn (2, inf) into (nan, inf) and lose the sign of a -0.0 real part
    data = np.empty(count, dtype=_COMPLEX[np.dtype(_CODES[code])])
    data.real = pairs[0::2]
    data.imag = pairs[1::2]
    (mcount,) = unpack("i")
    if mcount < 0:
        raise ValueError("invalid PI checkpoint metadata count")
    md = {}
    for _ in range(mcount):
        (kl,) = unpack("I")
        if kl > len(raw) - pos:
            raise ValueError("checkpoint string length exceeds the remaining file payload")
        key = take(kl).decode("utf-8")
        if key in md:
            raise ValueError(f"duplicate PI checkpoint metadata key {key}")
        (vl,) = unpack("I")
        if vl > len(raw) - pos:
            raise ValueError("checkpoint string length exceeds the remaining file payload")
        md[key] = take(vl).decode("utf-8")
    if pos != len(raw):
        raise ValueError("trailing data in PI checkpoint")
    return PIStateCheckpoint(version, PIState(basis, data), time, md)
