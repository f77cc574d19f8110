"""The banded apply tables (the qubit plans' device format) evaluated on the
CPU from their host copies (pimi_plan_banded_export) against the oracle's
apply, chains and lane shifts as the kernel defines them: every entry kind -- left, right, complex and real sandwiches, the
imaginary left/right entries, the merged left+right diagonal and the paired
real sandwiches -- must reproduce L x to rounding.  This checks the table
construction without a GPU; the device kernel is checked against the same
oracle by the GPU tests."""
import ctypes as C

import numpy as np
import pytest

import pimi
import pimi._native as N
from oracle import lowering as L
from oracle import oracle as O


def _tables(comp):
    counts = np.zeros(4, dtype=np.int64)
    N.check(N.lib().pimi_plan_banded_export(comp._plan, N.ptr(counts), None, None, None, None))
    ng, ngr, ne, nc = (int(v) for v in counts)
    cbase = np.zeros(max(ng, 1), dtype=np.int32)
    chains = np.zeros((max(ngr, 1), 7), dtype=np.int64)
    ents = np.zeros((max(ne, 1), 7), dtype=np.int32)
    coef = np.zeros(max(nc, 1), dtype=np.complex128)
    N.check(N.lib().pimi_plan_banded_export(comp._plan, N.ptr(counts), N.ptr(cbase), N.ptr(chains), N.ptr(ents),
                                            N.ptr(coef)))
    return cbase[:ng], chains[:ngr], ents[:ne], coef[:nc]


_SHIFT = 2              # largest row shift within a chain (kDiaShift)
_ROWS = 64 - _SHIFT     # output rows of a banded tile (kDiaRows)


def _coefficient(kind, qa, qb, qa2, qb2, coef, m):
    reals = coef.view(np.float64)  # packed real vectors: element i at real index 2*slot + i
    row = lambda q: coef[q:q + m][:, None]                     # noqa: E731
    col = lambda q: coef[q:q + m][None, :]                     # noqa: E731
    rrow = lambda q: reals[2 * q:2 * q + m][:, None]           # noqa: E731
    rcol = lambda q: reals[2 * q:2 * q + m][None, :]           # noqa: E731
    return {0: lambda: row(qa), 1: lambda: col(qb), 2: lambda: row(qa) * col(qb),
            3: lambda: rrow(qa) * rcol(qb), 4: lambda: 1j * rrow(qa), 5: lambda: 1j * rcol(qb),
            6: lambda: row(qa) + col(qb), 7: lambda: rrow(qa) * rcol(qb) + rrow(qa2) * rcol(qb2)}[int(kind)]()


def _banded_apply(comp, x):
    """y = L x from the tables, entry by entry: an entry of chain (dA, c0)
    with row shift sh and strip column dB reads X_in[a + dA + sh, b + c0 + dB]."""
    cbase, chains, ents, coef = _tables(comp)
    b = comp.basis
    y = np.zeros_like(x)
    kinds = set()
    for s, m in enumerate(b.dims):
        o = b.offsets[s]
        Y = np.zeros((m, m), dtype=complex)
        a = np.arange(m)[:, None]
        bb = np.arange(m)[None, :]
        for g in range(cbase[s], cbase[s + 1]):
            in_off, mn, dA, c0, span, e0, e1 = chains[g]
            Xn = x[in_off:in_off + mn * mn].reshape(mn, mn, order="F")
            for kind, dB, qa, qb, qa2, qb2, sh in ents[e0:e1]:
                kinds.add(int(kind))
                r, c = a + dA + sh, bb + c0 + dB
                ok = (r >= 0) & (r < mn) & (c >= 0) & (c < mn)
                Xs = np.where(ok, Xn[np.clip(r, 0, mn - 1), np.clip(c, 0, mn - 1)], 0)
                Y += _coefficient(kind, qa, qb, qa2, qb2, coef, m) * Xs
        y[o:o + m * m] = Y.reshape(-1, order="F")
    return y, kinds


def _lane_apply(comp, x, TB):
    """The same as the kernel computes it: tiles of _ROWS rows and TB
    columns, 64 lanes holding strip rows a0 + l + dA (clamped into the input
    block; columns outside the block read zero), shifted up one lane per row
    of sh with a zero entering at lane 63, lanes l < _ROWS stored."""
    cbase, chains, ents, coef = _tables(comp)
    b = comp.basis
    y = np.full_like(x, np.nan)
    lanes = np.arange(64)
    for s, m in enumerate(b.dims):
        o = b.offsets[s]
        Y = np.zeros((m, m), dtype=complex)
        full = {}
        for g in range(cbase[s], cbase[s + 1]):
            for e in range(chains[g][5], chains[g][6]):
                kind, dB, qa, qb, qa2, qb2, sh = ents[e]
                full[e] = _coefficient(kind, qa, qb, qa2, qb2, coef, m)
        for b0 in range(0, m, TB):
            cols = np.arange(b0, b0 + TB + 2)
            for a0 in range(0, m, _ROWS):
                acc = np.zeros((64, TB), dtype=complex)
                arow = a0 + lanes
                crow = np.minimum(arow, m - 1)
                for g in range(cbase[s], cbase[s + 1]):
                    in_off, mn, dA, c0, span, e0, e1 = chains[g]
                    Xn = x[in_off:in_off + mn * mn].reshape(mn, mn, order="F")
                    r = np.clip(arow + dA, 0, mn - 1)
                    c = cols + c0
                    nk = TB if span == 0 else TB + 2
                    st = np.zeros((64, TB + 2), dtype=complex)
                    okc = (c >= 0) & (c < mn) & (np.arange(TB + 2) < nk)
                    st[:, okc] = Xn[np.ix_(r, c[okc])]
                    shifted = 0
                    for e in range(e0, e1):
                        kind, dB, qa, qb, qa2, qb2, sh = ents[e]
                        while shifted < sh:
                            st = np.vstack([st[1:], np.zeros((1, TB + 2))])
                            shifted += 1
                        cf = full[e]
                        colw = np.minimum(np.arange(b0, b0 + TB), m - 1)
                        coeff = np.broadcast_to(cf, (m, m))[np.ix_(crow, colw)]
                        acc += coeff * st[:, dB:dB + TB]
                keep = (lanes < _ROWS) & (arow < m)
                nb = min(TB, m - b0)
                Y[arow[keep][:, None], np.arange(b0, b0 + nb)[None, :]] = acc[keep][:, :nb]
        y[o:o + m * m] = Y.reshape(-1, order="F")
    return y


def _c2_model(N_):
    sp = pimi.spin_matrices(2)
    return pimi.qubit_ensemble_model(N_, hamiltonian=sp.jz + sp.jx, emission=0.1, dephasing=0.2, pumping=0.02,
                                     collective_emission=0.1)


@pytest.mark.parametrize("N_", [1, 2, 8, 13])
def test_banded_tables_reproduce_the_oracle_apply(N_):
    comp = pimi.compile(_c2_model(N_), host_only=True)
    sp = L.spin_matrices(2)
    ref, _ = O.compile_plan(N_, 2, L.qubit_ensemble_terms(hamiltonian=sp["jz"] + sp["jx"], emission=0.1,
                                                          dephasing=0.2, pumping=0.02, collective_emission=0.1))
    rng = np.random.default_rng(N_)
    x = rng.standard_normal(ref.n) + 1j * rng.standard_normal(ref.n)
    y, kinds = _banded_apply(comp, x)
    want = ref.apply(x)
    assert np.abs(y - want).max() <= 1e-13 * np.abs(want).max()
    for TB in (4, 8):
        assert np.abs(_lane_apply(comp, x, TB) - want).max() <= 1e-13 * np.abs(want).max()
    if N_ >= 8:
        # qubit plans use the specialised kinds: imaginary left/right, merged diagonal, paired sandwiches
        assert {4, 5, 6, 7} <= kinds, kinds


def test_banded_tables_general_hamiltonian():
    """A Hamiltonian with a jy part (complex off-diagonals) keeps the general
    left/right kinds where the coefficients are not purely imaginary."""
    sp = pimi.spin_matrices(2)
    model = pimi.qubit_ensemble_model(9, hamiltonian=0.8 * sp.jz + 0.5 * sp.jx + 0.3 * sp.jy, emission=0.1,
                                      dephasing=0.05)
    comp = pimi.compile(model, host_only=True)
    lsp = L.spin_matrices(2)
    ref, _ = O.compile_plan(9, 2, L.qubit_ensemble_terms(hamiltonian=0.8 * lsp["jz"] + 0.5 * lsp["jx"] +
                                                         0.3 * lsp["jy"], emission=0.1, dephasing=0.05))
    x = np.random.default_rng(3).standard_normal(ref.n) + 0j
    y, kinds = _banded_apply(comp, x)
    assert np.abs(y - ref.apply(x)).max() <= 1e-13 * np.abs(ref.apply(x)).max()
    assert np.abs(_lane_apply(comp, x, 8) - ref.apply(x)).max() <= 1e-13 * np.abs(ref.apply(x)).max()


def test_banded_chain_structure():
    """Chain invariants the kernel relies on: row shifts within a chain in
    0.._SHIFT and non-decreasing (the strip is only ever shifted up), strip
    columns within the chain's span, span at most 2; and one chain per input
    sector for the N=120 qubit model (its own block and the two neighbours:
    one strip load each instead of one per row shift)."""
    comp = pimi.compile(_c2_model(120), host_only=True)
    cbase, chains, ents, coef = _tables(comp)
    assert len(cbase) == len(comp.basis.dims) + 1
    for in_off, mn, dA, c0, span, e0, e1 in chains:
        assert 0 <= span <= 2 and e1 > e0
        sh = ents[e0:e1, 6]
        assert sh.min() >= 0 and sh.max() <= _SHIFT and np.all(np.diff(sh) >= 0)
        assert np.all((ents[e0:e1, 1] >= 0) & (ents[e0:e1, 1] <= span))
    per_sector = np.diff(cbase)
    sources = [len({int(c[0]) for c in chains[cbase[s]:cbase[s + 1]]}) for s in range(len(per_sector))]
    assert np.array_equal(per_sector, sources)
    assert per_sector.max() == 3 and len(chains) <= 3 * len(per_sector)


def _programs(comp, compile_=0):
    ln, npr = C.c_int64(), C.c_int32()
    pid = np.zeros(len(comp.basis.dims), dtype=np.int32)
    N.check(N.lib().pimi_plan_banded_programs(comp._plan, 0, None, C.c_int64(0), C.byref(ln), C.byref(npr), None))
    buf = C.create_string_buffer(ln.value + 1)
    N.check(N.lib().pimi_plan_banded_programs(comp._plan, compile_, buf, C.c_int64(ln.value + 1), C.byref(ln),
                                              C.byref(npr), N.ptr(pid)))
    return buf.value.decode(), npr.value, pid


@pytest.mark.parametrize("N_", [1, 2, 120])
def test_banded_programs_cover_every_sector_and_compile(N_):
    """The compiled banded programs: one per distinct sector structure (the
    qubit model at N=120 has three: the first sector, the last, and every
    other), every sector assigned one, each program's entries the sector's
    own (kinds, strip columns, row shifts in table order), and the source
    compiles for gfx950 (hiprtc, no device)."""
    comp = pimi.compile(_c2_model(N_), host_only=True)
    src, count, pid = _programs(comp, compile_=1)
    assert pid.min() >= 0 and pid.max() == count - 1
    if N_ == 120:
        assert count == 3 and pid[0] == 0 and pid[-1] == 2 and np.all(pid[1:-1] == 1)
    cbase, chains, ents, _ = _tables(comp)
    for s in range(len(pid)):
        body = src.split(f"void prog_{pid[s]}(")[1].split("\n}\n")[0]
        calls = [ln.strip() for ln in body.splitlines() if "L.template entry<" in ln]
        want = [(int(k), int(d)) for c in chains[cbase[s]:cbase[s + 1]] for k, d in ents[c[5]:c[6], :2]]
        got = [tuple(int(v) for v in ln.split("entry<")[1].split(">")[0].split(",")) for ln in calls]
        assert got == want
        shifts = body.count("dia_shift<R, TB>(st);")
        assert shifts == sum(int(ents[c[6] - 1, 6]) for c in chains[cbase[s]:cbase[s + 1]])
