"""Oracle: exact-arithmetic CPU restatement of the reference's PI lowering.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
module; only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` use it, and only as the checker.

It restates, with Python integers and ``fractions.Fraction`` in the role of
Julia's ``BigInt``/``Rational{BigInt}``, the setup half of the reference's
hot path:

* partitions / GT patterns / content        -- src/partitions.jl:29-45,
  src/gtpatterns.jl:1-106
* one-box Clebsch--Gordan coefficient       -- src/cgc.jl:6-39
* checked square root of an exact ratio     -- src/partitions.jl:445-484
* one-box branching weight                  -- src/partitions.jl:519-541
* OneBodyGeometry (transitions, packed
  contractions, branch scales, connections) -- src/geometry.jl:405-432,465-490,519-562
* collective sparse Schur blocks           -- src/geometry.jl:820-865
* collective_operator                       -- src/geometry.jl:1282-1308
* compile_term for the built-in one-body
  terms and the static fusion              -- src/liouvillian.jl:270-326,1290-1371,
                                               1543-1604,1700-1738
* trace functional                          -- src/source_protocol.jl:15-34

Float rounding follows the reference: a CG coefficient or branch scale is
``sqrt(RN(q))`` of the exact rational ``q`` (``float(Fraction)`` is correctly
rounded), and ``sqrt(f)`` of a sector multiplicity is ``sqrt(RN(f))``.
"""
from __future__ import annotations

import math
from fractions import Fraction
from functools import lru_cache

import numpy as np


# ---------------------------------------------------------------------------
# Combinatorics (src/partitions.jl, src/gtpatterns.jl)
# ---------------------------------------------------------------------------

def partitions(N: int, d: int):
    """Partitions of N into exactly d parts, descending lexicographic order
    (src/partitions.jl:29-45)."""
    out = []

    def rec(prefix, left, cap):
        if len(prefix) == d:
            if left == 0:
                out.append(tuple(prefix))
            return
        for x in range(min(cap, left), -1, -1):
            rec(prefix + [x], left - x, x)

    rec([], N, N)
    return out


def removable_corners(p):
    """One-based rows from which a box can be removed (src/partitions.jl:54)."""
    D = len(p)
    return [i + 1 for i in range(D) if p[i] > (0 if i == D - 1 else p[i + 1])]


def remove_corner(p, row):
    return tuple(p[i] - (1 if i + 1 == row else 0) for i in range(len(p)))


def _gidx(i, k):
    # one-based (i, k) -> zero-based index into the row-by-row entry tuple
    return k * (k - 1) // 2 + i - 1


def gt_entry(g, i, k):
    return g[_gidx(i, k)]


def _interlacing_rows(upper):
    k = len(upper) - 1
    out = []

    def rec(v, i):
        if i > k:
            out.append(list(v))
            return
        for x in range(upper[i - 1], upper[i] - 1, -1):
            v.append(x)
            rec(v, i + 1)
            v.pop()

    rec([], 1)
    return out


@lru_cache(maxsize=None)
def gt_patterns(p):
    """All GT patterns with top row p, as entry tuples stored row by row
    (row 1 first), sorted ascending (src/gtpatterns.jl:92-106)."""
    D = len(p)
    out = []

    def descend(rows):
        if len(rows) == D:
            ordered = list(reversed(rows))
            out.append(tuple(x for r in ordered for x in r))
            return
        for row in _interlacing_rows(rows[-1]):
            rows.append(row)
            descend(rows)
            rows.pop()

    descend([list(p)])
    out.sort()
    return tuple(out)


def content(g, D):
    """Weight content: successive row-sum differences (src/gtpatterns.jl:47-65)."""
    res = []
    for K in range(1, D + 1):
        v = sum(g[_gidx(i, K)] for i in range(1, K + 1))
        v -= sum(g[_gidx(i, K - 1)] for i in range(1, K))
        res.append(v)
    return tuple(res)


def shape_weight(g, D):
    return sum(g[_gidx(i, D)] for i in range(1, D + 1))


def symmetric_group_dimension(p):
    """Hook-length formula (src/partitions.jl:119-127)."""
    n = sum(p)
    if n == 0:
        return 1
    den = 1
    D = len(p)
    for i in range(D):
        for j in range(1, p[i] + 1):
            below = sum(1 for k in range(i + 1, D) if p[k] >= j)
            den *= p[i] - j + below + 1
    return math.factorial(n) // den


def unitary_group_dimension(p):
    """Weyl dimension (src/partitions.jl:595-602)."""
    D = len(p)
    q = Fraction(1)
    for i in range(D - 1):
        for j in range(i + 1, D):
            q *= Fraction(p[i] - p[j] + j - i, j - i)
    assert q.denominator == 1
    return q.numerator


def sqrt_exact_ratio(q: Fraction) -> float:
    """sqrt of an exact nonnegative rational, rounded as
    src/partitions.jl:445-484 does: RN(sqrt(RN(q / 4^k))) * 2^k, which equals
    sqrt(RN(q)) whenever q is in the normal double range."""
    if q == 0:
        return 0.0
    if q < 0:
        raise ValueError("negative exact ratio")
    num, den = q.numerator, q.denominator
    e = num.bit_length() - den.bit_length()
    below = (num < (den << e)) if e >= 0 else ((num << (-e)) < den)
    if below:
        e -= 1
    r = e // 2  # fld
    shift = 2 * r
    scaled = Fraction(num, den << shift) if shift >= 0 else Fraction(num << (-shift), den)
    return math.ldexp(math.sqrt(float(scaled)), r)


def sqrt_exact_integer(f: int) -> float:
    """src/partitions.jl:486-513: sqrt(RN(f)) for representable f."""
    return math.sqrt(float(f))


# ---------------------------------------------------------------------------
# One-box Clebsch--Gordan coefficient (src/cgc.jl:6-39)
# ---------------------------------------------------------------------------

def triangular_shift(mu, lam, j, D):
    if not (0 <= j < D):
        return None
    tau = []
    for k in range(1, D + 1):
        dif = [gt_entry(lam, i, k) - gt_entry(mu, i, k) for i in range(1, k + 1)]
        if k <= j:
            if any(x != 0 for x in dif):
                return None
        else:
            if not (dif.count(1) == 1 and all(x in (0, 1) for x in dif)):
                return None
            tau.append(dif.index(1) + 1)
    return tau


def _p(g, i, k):
    return gt_entry(g, i, k) + k - i


def cgc_exact(mu, j, lam, D):
    """Return (sign, q) with the coefficient = sign*sqrt(q), or None for an
    exact zero (src/cgc.jl:20-39)."""
    if shape_weight(lam, D) != shape_weight(mu, D) + 1:
        return None
    tau = triangular_shift(mu, lam, j, D)
    if tau is None:
        return None
    tj = tau[0]
    q = Fraction(1)
    for k in range(1, j + 1):
        q *= _p(mu, tj, j + 1) - _p(mu, k, j)
    for k in range(1, j + 2):
        if k == tj:
            continue
        q /= _p(mu, tj, j + 1) - _p(mu, k, j + 1)
    sign = 1
    for l in range(j + 2, D + 1):
        tprev = tau[l - j - 2]
        tl = tau[l - j - 1]
        sign *= -1 if tprev < tl else 1
        for k in range(1, l + 1):
            if k == tl:
                continue
            q *= Fraction(_p(mu, tprev, l - 1) - _p(mu, k, l) + 1,
                          _p(mu, tl, l) - _p(mu, k, l))
        for k in range(1, l):
            if k == tprev:
                continue
            q *= Fraction(_p(mu, tl, l) - _p(mu, k, l - 1),
                          _p(mu, tprev, l - 1) - _p(mu, k, l - 1) + 1)
    if q < 0 and abs(float(q)) < 1e-14:
        q = -q
    if q < 0:
        raise ValueError("negative squared CG coefficient")
    return sign, q


def cgc(mu, j, lam, D) -> float:
    r = cgc_exact(mu, j, lam, D)
    if r is None:
        return 0.0
    sign, q = r
    return float(sign) * sqrt_exact_ratio(q)


def one_box_branch_weight(lam, mu):
    """Exact |lambda| f^mu / f^lambda (src/partitions.jl:519-541)."""
    D = len(lam)
    row = None
    for i in range(D):
        if lam[i] == mu[i] + 1:
            row = i
            break
    assert row is not None
    shifted = [lam[i] + D - (i + 1) for i in range(D)]
    pivot = shifted[row]
    result = Fraction(pivot)
    for i in range(row):
        diff = shifted[i] - pivot
        result *= Fraction(diff + 1, diff)
    for i in range(row + 1, D):
        diff = pivot - shifted[i]
        result *= Fraction(diff - 1, diff)
    return result


# ---------------------------------------------------------------------------
# Basis and one-body geometry (src/basis.jl:11-35, src/geometry.jl:450-562)
# ---------------------------------------------------------------------------

class Basis:
    """PIBasis restated: sectors in descending lexicographic order, GT
    patterns ascending, column-major blocks (src/basis.jl:11-35)."""

    def __init__(self, N, d):
        self.N = N
        self.d = d
        self.sectors = partitions(N, d)
        self.patterns = [gt_patterns(p) for p in self.sectors]
        self.dims = [len(g) for g in self.patterns]
        self.offsets = [0]
        for m in self.dims:
            self.offsets.append(self.offsets[-1] + m * m)
        self.n = self.offsets[-1]
        self.index = {p: i for i, p in enumerate(self.sectors)}
        self.removed = [[remove_corner(p, r) for r in removable_corners(p)]
                        for p in self.sectors]

    def multiplicity(self, s):
        return symmetric_group_dimension(self.sectors[s])


def _onebody_transitions(parent_patterns, child_patterns, D):
    """src/geometry.jl:465-490: per child pattern, the list of
    (parent index, one-based local label, CG value) in the reference's order."""
    parent_by_weight = {}
    for pi_, parent in enumerate(parent_patterns):
        parent_by_weight.setdefault(content(parent, D), []).append(pi_)
    transitions = []
    for child in child_patterns:
        cw = content(child, D)
        lst = []
        for label in range(D):
            pw = tuple(cw[l] + (1 if l == label else 0) for l in range(D))
            for pi_ in parent_by_weight.get(pw, []):
                v = cgc(child, label, parent_patterns[pi_], D)
                if v != 0.0:
                    lst.append((pi_, label, v))
        transitions.append(lst)
    return transitions


class OneBodyGeometry:
    """src/geometry.jl:519-562.  contractions[(li, mu, ni)] maps a cell
    (a, c) (zero-based) to its ordered list of (i, j, value) products;
    scales[(li, mu, ni)] = sqrt(w(l,mu) w(n,mu)); connections[(li, ni)] lists
    the common children in removable-corner order of the left sector."""

    def __init__(self, basis: Basis, diagonal_only=False):
        b = basis
        D = b.d
        self.basis = b
        trans = {}
        for s, children in enumerate(b.removed):
            for mu in children:
                trans[(s, mu)] = _onebody_transitions(b.patterns[s], gt_patterns(mu), D)
        self.contractions = {}
        self.scales = {}
        self.connections = {}
        ns = len(b.sectors)
        for li in range(ns):
            for ni in range(ns):
                if diagonal_only and li != ni:
                    continue
                common = [mu for mu in b.removed[li] if mu in b.removed[ni]]
                if common:
                    self.connections[(li, ni)] = common
                for mu in common:
                    left = trans[(li, mu)]
                    right = trans[(ni, mu)]
                    cells = {}
                    for child in range(len(left)):
                        for (a, i, x) in left[child]:
                            for (c, j, y) in right[child]:
                                v = x * y
                                if v == 0.0:
                                    continue
                                cells.setdefault((a, c), []).append((i, j, v))
                    self.contractions[(li, mu, ni)] = cells
                    w = one_box_branch_weight(b.sectors[li], mu) * \
                        one_box_branch_weight(b.sectors[ni], mu)
                    self.scales[(li, mu, ni)] = sqrt_exact_ratio(w)

    def conn(self, li, ni):
        return self.connections.get((li, ni), [])


def _contract(terms, X):
    """src/geometry.jl:820-823: left fold of conj(z)*X[i,j] from zero."""
    acc = 0j
    for (i, j, z) in terms:
        acc = acc + z * complex(X[i, j])
    return acc


def collective_sparse_block(geom: OneBodyGeometry, X, s):
    """src/geometry.jl:840-865: exact-nonzero entries of the collective Schur
    block of sum_i X_i in sector s, column-major order -> CSC."""
    b = geom.basis
    n = b.dims[s]
    conns = geom.conn(s, s)
    tables = [(geom.scales[(s, mu, s)], geom.contractions[(s, mu, s)]) for mu in conns]
    colptr = [0]
    rows = []
    vals = []
    for col in range(n):
        for row in range(n):
            value = 0j
            for scale, cells in tables:
                value = value + scale * _contract(cells.get((row, col), ()), X)
            if value == 0:
                continue
            rows.append(row)
            vals.append(value)
        colptr.append(len(rows))
    return CSC(n, n, colptr, rows, vals)


class CSC:
    def __init__(self, m, n, colptr, rows, vals):
        self.m, self.n = m, n
        self.colptr = list(colptr)
        self.rows = list(rows)
        self.vals = list(vals)

    def dense(self):
        A = np.zeros((self.m, self.n), dtype=complex)
        for c in range(self.n):
            for p in range(self.colptr[c], self.colptr[c + 1]):
                A[self.rows[p], c] += self.vals[p]
        return A

    @staticmethod
    def from_dense(A):
        m, n = A.shape
        colptr = [0]
        rows, vals = [], []
        for c in range(n):
            for r in range(m):
                v = complex(A[r, c])
                if v != 0:
                    rows.append(r)
                    vals.append(v)
            colptr.append(len(rows))
        return CSC(m, n, colptr, rows, vals)


def sparse_gram(K: CSC) -> CSC:
    """K'K with exact zeros dropped (src/liouvillian.jl:1318-1341)."""
    A = K.dense()
    return CSC.from_dense(A.conj().T @ A)


def collective_block_dense(geom, X, s):
    return collective_sparse_block(geom, X, s).dense()


# ---------------------------------------------------------------------------
# Terms, compile and static fusion (src/liouvillian.jl)
# ---------------------------------------------------------------------------

class Term:
    """kind in {'local_h', 'collective_h', 'local_jump', 'collective_jump',
    'direct_h'}.  ``op`` is a d x d complex matrix, or for 'direct_h' a list
    of physical Schur blocks (one dense matrix per sector)."""

    def __init__(self, kind, op, rate=1.0, hbar=1.0):
        self.kind = kind
        self.op = op
        self.rate = rate
        self.hbar = hbar


class FusedPlan:
    """FusedStaticPIKernel restated (src/liouvillian.jl:539-545): fused H and
    loss CSC blocks per sector, collective gains [(scale, [K per sector])],
    one-body gains [(scale, branches)] where each branch is
    (out_sector, in_sector, r, dense C, use_support, support list)."""

    def __init__(self, basis):
        self.basis = basis
        self.H = None
        self.Q = None
        self.collective_gains = []
        self.onebody_gains = []


def _static_contraction(C):
    """src/liouvillian.jl:270-292: support kept when 2 nnz^2 <= m_l m_n (m_l+m_n)."""
    ml, mn = C.shape
    sup = []
    for col in range(mn):
        for row in range(ml):
            v = complex(C[row, col])
            if v != 0:
                sup.append((row, col, v))
    nnz = len(sup)
    use = 2 * nnz * nnz <= ml * mn * (ml + mn)
    return use, sup


def _fuse_sparse(blocks_scales, basis):
    """src/liouvillian.jl:1543-1604: column-by-column accumulation in kernel
    order, exact zeros removed."""
    out = []
    for s in range(len(basis.sectors)):
        m = basis.dims[s]
        colptr = [0]
        rows, vals = [], []
        for col in range(m):
            acc = {}
            for scale, blocks in blocks_scales:
                B = blocks[s]
                for p in range(B.colptr[col], B.colptr[col + 1]):
                    r = B.rows[p]
                    acc[r] = acc.get(r, 0j) + scale * B.vals[p]
            for r in sorted(acc):
                if acc[r] != 0:
                    rows.append(r)
                    vals.append(acc[r])
            colptr.append(len(rows))
        out.append(CSC(m, m, colptr, rows, vals))
    return out


def compile_model(basis: Basis, terms, geom: OneBodyGeometry | None = None) -> FusedPlan:
    """compile(model; backend=:matrixfree) restated for the built-in one-body
    terms: compile_term per term, then _fuse_fixed_kernels
    (src/liouvillian.jl:1290-1371,1700-1738)."""
    if geom is None:
        geom = OneBodyGeometry(basis)
    ns = len(basis.sectors)
    hams = []       # (scale, [CSC per sector])
    losses = []     # (scale, [CSC per sector])
    plan = FusedPlan(basis)
    for t in terms:
        if t.kind in ("local_h", "collective_h"):
            blocks = [collective_sparse_block(geom, t.op, s) for s in range(ns)]
            hams.append((complex(t.rate / t.hbar), blocks))
        elif t.kind == "direct_h":
            blocks = [CSC.from_dense(np.asarray(Bk)) for Bk in t.op]
            hams.append((complex(t.rate / t.hbar), blocks))
        elif t.kind == "collective_jump":
            K = [collective_sparse_block(geom, t.op, s) for s in range(ns)]
            Q = [sparse_gram(k) for k in K]
            losses.append((complex(t.rate), Q))
            plan.collective_gains.append((complex(t.rate), K))
        elif t.kind == "local_jump":
            L = np.asarray(t.op, dtype=complex)
            Q = [collective_sparse_block(geom, L.conj().T @ L, s) for s in range(ns)]
            losses.append((complex(t.rate), Q))
            branches = []
            for li in range(ns):
                for ni in range(ns):
                    for mu in geom.conn(li, ni):
                        key = (li, mu, ni)
                        cells = geom.contractions[key]
                        C = np.zeros((basis.dims[li], basis.dims[ni]), dtype=complex)
                        for (a, c), terms_ in cells.items():
                            C[a, c] = _contract(terms_, L)
                        use, sup = _static_contraction(C)
                        branches.append((li, ni, geom.scales[key], C, use, sup))
            plan.onebody_gains.append((complex(t.rate), branches))
        else:
            raise ValueError(t.kind)
    plan.H = _fuse_sparse(hams, basis) if hams else None
    plan.Q = _fuse_sparse(losses, basis) if losses else None
    return plan


def trace_functional(basis: Basis):
    """src/source_protocol.jl:15-34: (indices, values) with sqrt(f^nu) on the
    diagonal of every block."""
    idx, vals = [], []
    for s, p in enumerate(basis.sectors):
        m = basis.dims[s]
        sc = sqrt_exact_integer(symmetric_group_dimension(p))
        for a in range(m):
            idx.append(basis.offsets[s] + a + a * m)
            vals.append(sc)
    return np.array(idx, dtype=np.int64), np.array(vals, dtype=float)


def collective_operator(geom: OneBodyGeometry, X):
    """src/geometry.jl:1282-1308: coefficient vector of sum_i X_i."""
    b = geom.basis
    out = np.zeros(b.n, dtype=complex)
    for s, p in enumerate(b.sectors):
        m = b.dims[s]
        K = collective_block_dense(geom, X, s)
        sc = sqrt_exact_integer(symmetric_group_dimension(p))
        out[b.offsets[s]:b.offsets[s + 1]] = (sc * K).reshape(-1, order="F")
    return out


def pi_product(basis: Basis, a, b):
    """src/operators.jl:120-135: coefficient blocks C_A C_B / sqrt(f)."""
    out = np.zeros(basis.n, dtype=complex)
    for s, p in enumerate(basis.sectors):
        m = basis.dims[s]
        o0, o1 = basis.offsets[s], basis.offsets[s + 1]
        A = a[o0:o1].reshape(m, m, order="F")
        B = b[o0:o1].reshape(m, m, order="F")
        sc = sqrt_exact_integer(symmetric_group_dimension(p))
        out[o0:o1] = ((A @ B) / sc).reshape(-1, order="F")
    return out


def physical_blocks(basis: Basis, vec):
    """coefficient_block / sqrt(f) per sector (src/basis.jl:163-177)."""
    out = []
    for s, p in enumerate(basis.sectors):
        m = basis.dims[s]
        o0, o1 = basis.offsets[s], basis.offsets[s + 1]
        sc = sqrt_exact_integer(symmetric_group_dimension(p))
        out.append(vec[o0:o1].reshape(m, m, order="F") / sc)
    return out


def symmetric_occupation_index(counts):
    """One-based pattern index of an occupation state in the fully symmetric
    sector (src/spin.jl:109-125)."""
    D = len(counts)
    remaining = sum(counts)
    zb = 0
    for level in range(1, D):
        trailing = D - level
        c = counts[level - 1]
        zb += math.comb(remaining + trailing, trailing) - math.comb(remaining - c + trailing, trailing)
        remaining -= c
    return zb + 1


def computational_product_state(basis: Basis, level: int):
    """|level><level|^N (one-based level) -- a single unit coordinate in the
    fully symmetric sector (src/spin.jl:79-86 via src/spin.jl:109-125)."""
    counts = [0] * basis.d
    counts[level - 1] = basis.N
    a = symmetric_occupation_index(counts) - 1
    m = basis.dims[0]
    v = np.zeros(basis.n, dtype=complex)
    v[basis.offsets[0] + a + a * m] = 1.0
    return v


def spin_matrices(d: int):
    """src/spin.jl:13-36: ascending |-j>..|j> ordering."""
    j = (d - 1) / 2
    jp = np.zeros((d, d), dtype=complex)
    for col in range(1, d):
        m = -j + (col - 1)
        jp[col, col - 1] = math.sqrt((j - m) * (j + m + 1))
    jm = jp.conj().T.copy()
    jx = (jp + jm) * 0.5
    jy = (jp - jm) / 2j
    jz = np.diag([complex(-j + k) for k in range(d)])
    return dict(j=j, jx=jx, jy=jy, jz=jz, jp=jp, jm=jm)


def qubit_ensemble_terms(hamiltonian=None, emission=0.0, dephasing=0.0, pumping=0.0,
                         collective_emission=0.0, collective_dephasing=0.0,
                         collective_pumping=0.0):
    """src/spin.jl:496-513 term order: H, local jm, jz, jp, collective jm, jz, jp;
    zero rates omitted."""
    sp = spin_matrices(2)
    terms = []
    if hamiltonian is not None:
        terms.append(Term("collective_h", np.asarray(hamiltonian, dtype=complex)))
    for rate, op, kind in ((emission, sp["jm"], "local_jump"), (dephasing, sp["jz"], "local_jump"),
                           (pumping, sp["jp"], "local_jump"),
                           (collective_emission, sp["jm"], "collective_jump"),
                           (collective_dephasing, sp["jz"], "collective_jump"),
                           (collective_pumping, sp["jp"], "collective_jump")):
        if rate != 0:
            terms.append(Term(kind, op, rate))
    return terms


def driven_qubits_terms(drive=0.7, detuning=0.0, emission=0.12, pumping=0.02, dephasing=0.0):
    """Models.driven_qubits (src/model_zoo.jl:214-236): H = drive/2 (2 jx) +
    detuning/2 (2 jz)."""
    sp = spin_matrices(2)
    H = (drive / 2) * (2.0 * sp["jx"]) + (detuning / 2) * (2.0 * sp["jz"])
    return qubit_ensemble_terms(hamiltonian=H, emission=emission, pumping=pumping,
                                dephasing=dephasing)
