"""PI state checkpoints in the reference's portable layout (PIDCHKPT v1,
src/checkpoints.jl:1-160); cases follow test/test_research_utilities.jl:125-181.
The byte-layout test spells out the file field by field from the reference's
writer, so a reordered or resized field fails it."""
import struct

import numpy as np
import pytest

import pimi


def test_restricted_float32_roundtrip(tmp_path):
    b = pimi.PIBasis(3, 2, sectors=[(3, 0)])
    rho = pimi.computational_product_state(b, 2, T=np.float32)
    cp = pimi.make_checkpoint(rho, time=np.float32(0.25), metadata={"model": "test", "seed": 4})
    assert cp.schema_version == pimi.PI_CHECKPOINT_VERSION == 1
    path = str(tmp_path / "state.pid")
    assert pimi.save_checkpoint(path, cp) == path
    got = pimi.load_checkpoint(path)
    assert got.schema_version == cp.schema_version
    assert (got.state.basis.N, got.state.basis.d, got.state.basis.sectors) == (3, 2, [(3, 0)])
    assert got.state.data.dtype == np.complex64
    np.testing.assert_array_equal(got.state.data, rho.data)
    assert got.time == np.float32(0.25) and got.time.dtype == np.float32
    assert got.metadata == {"model": "test", "seed": "4"}
    detached = pimi.checkpoint_state(got)
    detached.data[0] += 1
    assert not np.array_equal(detached.data, got.state.data)


def test_byte_layout(tmp_path):
    b = pimi.PIBasis(2, 2)  # sectors (2,0) m=3, (1,1) m=1 -> n = 10
    data = np.arange(10) * (0.5 - 0.25j)
    path = str(tmp_path / "s.pid")
    pimi.save_checkpoint(path, pimi.PIState(b, data), time=0.125, metadata={"z": "last", "a": "first"})
    raw = open(path, "rb").read()
    want = b"PIDCHKPT" + struct.pack("<H", 1) + struct.pack("<B", 64)
    want += struct.pack("<q", 2) + struct.pack("<i", 2) + struct.pack("<i", 2)
    want += struct.pack("<qq", 2, 0) + struct.pack("<qq", 1, 1)
    want += b"\x01" + struct.pack("<d", 0.125) + struct.pack("<q", 10)
    for v in data:
        want += struct.pack("<dd", v.real, v.imag)
    want += struct.pack("<i", 2)
    for k, v in (("a", "first"), ("z", "last")):  # sorted by key
        want += struct.pack("<I", len(k)) + k.encode() + struct.pack("<I", len(v)) + v.encode()
    assert raw == want


def test_complete_qutrit_basis_roundtrip(tmp_path):
    b = pimi.PIBasis(4, 3)
    rng = np.random.default_rng(1)
    st = pimi.PIState(b, rng.standard_normal(len(b)) + 1j * rng.standard_normal(len(b)))
    path = str(tmp_path / "q.pid")
    pimi.save_checkpoint(path, st)
    got = pimi.load_checkpoint(path)
    assert got.time is None and got.metadata == {}
    assert got.state.basis == b
    np.testing.assert_array_equal(got.state.data, st.data)


def test_invalid_files_and_values(tmp_path):
    bad = tmp_path / "corrupt.pid"
    bad.write_bytes(b"not a checkpoint")
    with pytest.raises(ValueError):
        pimi.load_checkpoint(str(bad))
    good = str(tmp_path / "g.pid")
    b = pimi.PIBasis(2, 2)
    pimi.save_checkpoint(good, pimi.PIState(b, np.ones(len(b))), metadata={"k": "v"})
    raw = open(good, "rb").read()
    for name, blob in (("trailing", raw + b"\x00"), ("truncated", raw[:-3]),
                       ("version", raw[:8] + struct.pack("<H", 2) + raw[10:]),
                       ("code", raw[:10] + struct.pack("<B", 8) + raw[11:]),
                       ("bigfloat", raw[:10] + struct.pack("<B", 128) + raw[11:])):
        p = tmp_path / f"{name}.pid"
        p.write_bytes(blob)
        with pytest.raises(ValueError):
            pimi.load_checkpoint(str(p))
    # a duplicated metadata key
    dup = raw[:-(4 + 4 + 1 + 1)] + struct.pack("<i", 2)
    dup += (struct.pack("<I", 1) + b"k" + struct.pack("<I", 1) + b"v") * 2
    p = tmp_path / "dup.pid"
    p.write_bytes(dup)
    with pytest.raises(ValueError):
        pimi.load_checkpoint(str(p))
    rho32 = pimi.computational_product_state(b, 1, T=np.float32)
    with pytest.raises(ValueError):  # 0.1 is not a float32
        pimi.make_checkpoint(rho32, time=0.1)
    with pytest.raises(ValueError):
        pimi.make_checkpoint(rho32, time=np.inf)
    with pytest.raises(ValueError):
        pimi.save_checkpoint(str(tmp_path / "x.jld2"), rho32)


def test_restricted_basis_cannot_compile():
    b = pimi.PIBasis(3, 2, sectors=[(3, 0)])
    model = pimi.PIModel(b, [pimi.LocalJump(pimi.spin_matrices(2).jm, 0.1)])
    with pytest.raises(ValueError):
        pimi.compile(model, host_only=True)
    with pytest.raises(ValueError):
        pimi.PIBasis(3, 2, sectors=[(2, 2)])


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_special_values_roundtrip_byte_for_byte(tmp_path, dtype):
    """Coefficients are rebuilt as complex(re, im) exactly: -0.0 real parts,
    infinite and NaN parts survive a save/load/save cycle bit for bit (a
    re + 1j*im reconstruction turns (2, inf) into (nan, inf))."""
    b = pimi.PIBasis(2, 2)  # n = 10
    R = np.float64 if dtype == np.complex128 else np.float32
    re = np.array([-0.0, 2.0, np.inf, -np.inf, np.nan, 0.0, 1.5, -0.0, 3.0, -2.0], dtype=R)
    im = np.array([1.0, np.inf, -0.0, 2.0, 0.5, np.nan, -np.inf, -0.0, 0.0, 7.0], dtype=R)
    data = np.empty(10, dtype=dtype)
    data.real, data.imag = re, im
    p1, p2 = str(tmp_path / "a.pid"), str(tmp_path / "b.pid")
    pimi.save_checkpoint(p1, pimi.PIState(b, data))
    got = pimi.load_checkpoint(p1)
    assert got.state.data.dtype == dtype
    np.testing.assert_array_equal(got.state.data.real.view(np.uint64 if R == np.float64 else np.uint32),
                                  re.view(np.uint64 if R == np.float64 else np.uint32))
    np.testing.assert_array_equal(got.state.data.imag.view(np.uint64 if R == np.float64 else np.uint32),
                                  im.view(np.uint64 if R == np.float64 else np.uint32))
    pimi.save_checkpoint(p2, got)
    assert open(p1, "rb").read() == open(p2, "rb").read()


def _golden(name):
    import os
    return os.path.join(os.path.dirname(__file__), "golden", name)


def test_golden_pidchkpt_fixtures_reload_and_pin():
    """The golden states exchanged in the reference's portable layout
    (tests/golden/make_pidchkpt_fixtures.py): both load with every field
    validated, carry the C1 basis (N=8, d=2, all five sectors, n = 165) and
    their time; the oracle, rerun on one thread, reproduces them to 1e-14;
    and <sum_i n_i> of each matches the closed form of the local-only C1
    model (product states: steady state and t = 4)."""
    import json
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_pidchkpt_fixtures as G
    from oracle import lowering as L
    from oracle import oracle as O
    steady = pimi.load_checkpoint(_golden("c1_steady_state.pid"))
    dyn = pimi.load_checkpoint(_golden("c1_expv_t4.pid"))
    for cp in (steady, dyn):
        b = cp.state.basis
        assert (b.N, b.d, len(b.sectors), cp.state.data.shape) == (8, 2, 5, (165,))
        assert cp.state.data.dtype == np.complex128 and cp.schema_version == 1
    assert steady.time is None and dyn.time == 4.0
    x, _, rho_t, _ = G.c1_states()
    assert np.abs(steady.state.data - x).max() <= 1e-14 * np.abs(x).max()
    assert np.abs(dyn.state.data - rho_t).max() <= 1e-14 * np.abs(rho_t).max()
    gold = json.load(open(_golden("closed_form.json")))["c1"]
    _, geom = O.compile_plan(8, 2, L.driven_qubits_terms(**G.C1_PARAMS))
    sp = L.spin_matrices(2)
    n_op = L.collective_operator(geom, sp["jp"] @ sp["jm"])
    assert abs(np.vdot(n_op, steady.state.data).real - gold["steady_n"]) <= 1e-10 * abs(gold["steady_n"])
    k4 = int(np.argmin(np.abs(np.asarray(gold["t"]) - 4.0)))
    assert abs(np.vdot(n_op, dyn.state.data).real - gold["n"][k4]) <= 1e-10 * abs(gold["n"][k4])
