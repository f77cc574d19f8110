"""The banded apply tables (the qubit plans' device format) evaluated on the
CPU from their host copies (pimi_plan_banded_export) against the oracle's
apply: every entry kind -- left, right, complex and real sandwiches, the
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
    gbase = np.zeros(max(ng, 1), dtype=np.int32)
    groups = np.zeros((max(ngr, 1), 7), dtype=np.int64)
    ents = np.zeros((max(ne, 1), 6), dtype=np.int32)
    coef = np.zeros(max(nc, 1), dtype=np.complex128)
    N.check(N.lib().pimi_plan_banded_export(comp._plan, N.ptr(counts), N.ptr(gbase), N.ptr(groups), N.ptr(ents),
                                            N.ptr(coef)))
    return gbase[:ng], groups[:ngr], ents[:ne], coef[:nc]


def _banded_apply(comp, x):
    """y = L x from the tables, entry by entry, as the kernel defines them."""
    gbase, groups, ents, coef = _tables(comp)
    b = comp.basis
    y = np.zeros_like(x)
    reals = coef.view(np.float64)  # packed real vectors: element i at real index 2*slot + i
    kinds = set()
    for s, m in enumerate(b.dims):
        o = b.offsets[s]
        Y = np.zeros((m, m), dtype=complex)
        a = np.arange(m)[:, None]
        bb = np.arange(m)[None, :]
        for g in range(gbase[s], gbase[s + 1]):
            in_off, mn, dA, c0, span, e0, e1 = groups[g]
            Xn = x[in_off:in_off + mn * mn].reshape(mn, mn, order="F")
            for kind, dB, qa, qb, qa2, qb2 in ents[e0:e1]:
                kinds.add(int(kind))
                r, c = a + dA, bb + c0 + dB
                ok = (r >= 0) & (r < mn) & (c >= 0) & (c < mn)
                Xs = np.where(ok, Xn[np.clip(r, 0, mn - 1), np.clip(c, 0, mn - 1)], 0)
                row = lambda q: coef[q:q + m][:, None]                     # noqa: E731
                col = lambda q: coef[q:q + m][None, :]                     # noqa: E731
                rrow = lambda q: reals[2 * q:2 * q + m][:, None]           # noqa: E731
                rcol = lambda q: reals[2 * q:2 * q + m][None, :]           # noqa: E731
                coeff = {0: lambda: row(qa), 1: lambda: col(qb), 2: lambda: row(qa) * col(qb),
                         3: lambda: rrow(qa) * rcol(qb), 4: lambda: 1j * rrow(qa), 5: lambda: 1j * rcol(qb),
                         6: lambda: row(qa) + col(qb), 7: lambda: rrow(qa) * rcol(qb) + rrow(qa2) * rcol(qb2)}[int(kind)]()
                Y += coeff * Xs
        y[o:o + m * m] = Y.reshape(-1, order="F")
    return y, kinds


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
