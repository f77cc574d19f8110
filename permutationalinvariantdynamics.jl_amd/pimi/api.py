"""Host-side mirror of the reference API for the MI355X backend.

Names, argument meanings and error behaviour follow the reference package
(PermutationalInvariantDynamics.jl); every numeric call goes through libpimi.

  spin_matrices                 src/spin.jl:13-36
  PIBasis                       src/basis.jl:11-35
  LocalHamiltonian, ...         src/terms.jl:118-160
  qubit_ensemble_model          src/spin.jl:496-518
  Models.driven_qubits          src/model_zoo.jl:214-236
  compile(model; backend)       src/liouvillian.jl:4040-4098
  GMRESAlgorithm, ExpvAlgorithm src/highlevel.jl:32-40,81-124
  stationary_state              src/highlevel.jl:321-357 (krylov route, src/krylov.jl:797-891)
  solve_dynamics                src/highlevel.jl:559-733 (expv route)
  computational_product_state   src/spin.jl:79-86
  collective_operator           src/geometry.jl:1282-1308
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from fractions import Fraction
from types import SimpleNamespace

import numpy as np

from . import _native as N


# ---------------------------------------------------------------------------
# local operators, basis, terms, models
# ---------------------------------------------------------------------------

def spin_matrices(d: int = 2):
    """Spin-j matrices, j=(d-1)/2, ascending |-j>,...,|j> ordering."""
    if d < 1:
        raise ValueError("d must be positive")
    j = (d - 1) / 2
    jp = np.zeros((d, d), dtype=complex)
    for col in range(1, d):
        m = -j + (col - 1)
        jp[col, col - 1] = math.sqrt((j - m) * (j + m + 1))
    jm = jp.conj().T.copy()
    jx = (jp + jm) * 0.5
    jy = (jp - jm) / 2j
    jz = np.diag([complex(-j + k) for k in range(d)])
    return SimpleNamespace(j=j, jx=jx, jy=jy, jz=jz, jp=jp, jm=jm)


def _partitions(N, d):
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


def _weyl_dimension(p):
    q = Fraction(1)
    for i in range(len(p) - 1):
        for j in range(i + 1, len(p)):
            q *= Fraction(p[i] - p[j] + j - i, j - i)
    return int(q)


class PIBasis:
    """PI operator basis for N identical d-level systems (src/basis.jl:11-35):
    sectors in descending lexicographic order, block dimension = Weyl
    dimension.  `sectors` retains a subset of the complete list (kept in that
    order); the device backend compiles complete bases only."""

    def __init__(self, N: int, d: int, sectors=None):
        if N < 0:
            raise ValueError("N must be nonnegative")
        if d < 1:
            raise ValueError("d must be positive")
        self.N, self.d = int(N), int(d)
        full = _partitions(self.N, self.d)
        if sectors is None:
            self.sectors = full
        else:
            want = [tuple(int(v) for v in p) for p in sectors]
            if not want:
                raise ValueError("a PI basis needs at least one sector")
            if len(set(want)) != len(want):
                raise ValueError("sectors must be unique")
            bad = [p for p in want if p not in full]
            if bad:
                raise ValueError(f"{bad[0]} is not a partition of N={self.N} into {self.d} parts")
            self.sectors = [p for p in full if p in want]
        self.complete = len(self.sectors) == len(full)
        self.dims = [_weyl_dimension(p) for p in self.sectors]
        self.offsets = np.concatenate([[0], np.cumsum([m * m for m in self.dims])]).astype(np.int64)

    def __len__(self):
        return int(self.offsets[-1])

    def __eq__(self, other):
        return isinstance(other, PIBasis) and (self.N, self.d) == (other.N, other.d) and \
            self.sectors == other.sectors

    def __hash__(self):
        return hash((self.N, self.d, tuple(self.sectors)))

    def __repr__(self):
        return f"PIBasis(N={self.N}, d={self.d}, sectors={len(self.sectors)}, dimension={len(self)})"


@dataclass
class LocalHamiltonian:
    operator: np.ndarray
    rate: float = 1.0
    hbar: float = 1.0
    kind = N.TERM_LOCAL_H


@dataclass
class CollectiveHamiltonian:
    operator: np.ndarray
    rate: float = 1.0
    hbar: float = 1.0
    kind = N.TERM_COLLECTIVE_H


@dataclass
class LocalJump:
    operator: np.ndarray
    rate: float = 1.0
    kind = N.TERM_LOCAL_JUMP


@dataclass
class CollectiveJump:
    operator: np.ndarray
    rate: float = 1.0
    kind = N.TERM_COLLECTIVE_JUMP


class CollectiveOperator:
    """Symbolic sum_i X_i; its PI coefficient vector comes from libpimi."""

    def __init__(self, basis: PIBasis, X):
        self.basis = basis
        self.local = np.asarray(X, dtype=complex)

    def __mul__(self, other):
        if isinstance(other, CollectiveOperator):
            if other.basis != self.basis:
                raise ValueError("operators live on different PI bases")
            return CollectiveProduct(self.basis, self.local, other.local)
        return NotImplemented


class CollectiveProduct:
    """(sum_i X_i)(sum_i Y_i) as a PI operator."""

    def __init__(self, basis, X, Y):
        self.basis = basis
        self.X, self.Y = X, Y


@dataclass
class DirectPIHamiltonian:
    operator: CollectiveProduct
    rate: float = 1.0
    hbar: float = 1.0
    kind = N.TERM_DIRECT_PRODUCT_H


class PIModel:
    def __init__(self, basis: PIBasis, terms):
        self.basis = basis
        self.terms = tuple(terms)
        d = basis.d
        for t in self.terms:
            op = t.operator
            if isinstance(op, CollectiveProduct):
                if op.basis != basis:
                    raise ValueError("direct PI operator uses a different basis")
                continue
            if np.asarray(op).shape != (d, d):
                raise ValueError(f"term operator must be {d}x{d}")


def qubit_ensemble_model(basis_or_N, hamiltonian=None, emission=0, dephasing=0, pumping=0,
                         collective_emission=0, collective_dephasing=0, collective_pumping=0):
    """src/spin.jl:496-518: H then jumps jm, jz, jp (local) and jm, jz, jp
    (collective); exact zero rates are omitted."""
    b = basis_or_N if isinstance(basis_or_N, PIBasis) else PIBasis(basis_or_N, 2)
    if b.d != 2:
        raise ValueError("qubit_ensemble_model requires d == 2")
    sp = spin_matrices(2)
    terms = []
    if hamiltonian is not None:
        if isinstance(hamiltonian, CollectiveProduct):
            terms.append(DirectPIHamiltonian(hamiltonian))
        else:
            terms.append(CollectiveHamiltonian(np.asarray(hamiltonian, dtype=complex)))
    for rate, op, cls in ((emission, sp.jm, LocalJump), (dephasing, sp.jz, LocalJump), (pumping, sp.jp, LocalJump),
                          (collective_emission, sp.jm, CollectiveJump),
                          (collective_dephasing, sp.jz, CollectiveJump),
                          (collective_pumping, sp.jp, CollectiveJump)):
        if rate != 0:
            terms.append(cls(op, float(rate)))
    return PIModel(b, terms)


class Models:
    @staticmethod
    def driven_qubits(N, drive=0.7, detuning=0.0, emission=0.12, pumping=0.02, dephasing=0.0):
        """src/model_zoo.jl:214-236: H = drive*jx + detuning*jz per particle."""
        if emission < 0 or pumping < 0 or dephasing < 0:
            raise ValueError("rates must be nonnegative")
        sp = spin_matrices(2)
        H = (drive / 2) * (2.0 * sp.jx) + (detuning / 2) * (2.0 * sp.jz)
        return qubit_ensemble_model(N, hamiltonian=H, emission=emission, pumping=pumping, dephasing=dephasing)


# ---------------------------------------------------------------------------
# states and operators
# ---------------------------------------------------------------------------

class PIState:
    def __init__(self, basis: PIBasis, data):
        data = np.asarray(data)
        data = data.astype(np.complex64 if data.dtype in (np.complex64, np.float32) else np.complex128, copy=False)
        if data.shape != (len(basis),):
            raise ValueError("coefficient vector has wrong length")
        self.basis = basis
        self.data = data


class PIOperator(PIState):
    pass


def _symmetric_occupation_index(counts):
    D = len(counts)
    remaining = sum(counts)
    zb = 0
    for level in range(1, D):
        trailing = D - level
        c = counts[level - 1]
        zb += math.comb(remaining + trailing, trailing) - math.comb(remaining - c + trailing, trailing)
        remaining -= c
    return zb + 1


def computational_product_state(basis: PIBasis, level: int, T=np.float64) -> PIState:
    """|level><level|^N (one-based level) in the fully symmetric sector
    (src/spin.jl:79-86); T is the real scalar type (float64 or float32)."""
    if not (1 <= level <= basis.d):
        raise ValueError(f"level must be a one-based local-state index in 1:{basis.d}")
    sym = tuple([basis.N] + [0] * (basis.d - 1))
    if sym not in basis.sectors:
        raise ValueError("the basis does not retain the fully symmetric sector")
    s = basis.sectors.index(sym)
    counts = [0] * basis.d
    counts[level - 1] = basis.N
    a = _symmetric_occupation_index(counts) - 1
    m = basis.dims[s]
    v = np.zeros(len(basis), dtype=np.complex64 if np.dtype(T) == np.float32 else np.complex128)
    v[basis.offsets[s] + a + a * m] = 1.0
    return PIState(basis, v)


_IID_PLANS = {}


def iid_state(basis: PIBasis, rho, atol=0.0, rtol=None) -> PIState:
    """PI representation of rho^{otimes N} for a one-particle density matrix
    (iid_state, src/operators.jl:197-250): validated as the reference does
    (Hermitian within atol + rtol |rho|, positive within atol + rtol |A| with
    A = (rho + rho^H)/2, trace one within atol + rtol, the N-th power of the
    trace too), then every Schur block is built on the host by the one-box
    Clebsch-Gordan recurrence D_lambda = sum_{i,o} A[i,o] U_i D_mu U_o^T
    (pimi_iid_state); the coefficient block is sqrt(f^lambda) D_lambda.  A
    restricted basis must retain every sector that carries weight."""
    rho = np.asarray(rho, dtype=np.complex128)
    d, Np = basis.d, basis.N
    if rho.shape != (d, d):
        raise ValueError(f"rho must have size ({d}, {d})")
    rtol = math.sqrt(np.finfo(np.float64).eps) if rtol is None else float(rtol)
    atol = float(atol)
    if atol < 0 or rtol < 0:
        raise ValueError("atol and rtol must be nonnegative")
    herr = float(np.abs(rho - rho.conj().T).max())
    if not herr <= atol + rtol * float(np.abs(rho).max()):
        raise ValueError(f"rho must be Hermitian: error={herr}")
    A = (rho + rho.conj().T) / 2
    tol = atol + rtol * float(np.abs(A).max())
    lo = float(np.linalg.eigvalsh(A).min())
    if lo < -(tol + 32 * d * np.finfo(np.float64).eps * max(float(np.abs(A).max()), 1.0)):
        raise ValueError(f"rho must be positive within tolerance {tol}")
    tr = float(np.trace(A).real)
    if not abs(tr - 1.0) <= atol + rtol:
        raise ValueError(f"rho must have trace one: trace={tr}")
    tt = tr ** Np
    if not (math.isfinite(tt) and abs(tt - 1.0) <= atol + rtol):
        raise ValueError(f"the accepted one-particle trace error is amplified at N={Np}: the tensor-power trace "
                         f"would be {tt}")
    full = PIBasis(Np, d)
    if Np == 0:
        out = np.ones(1, dtype=np.complex128)
    else:
        key = (Np, d)
        if key not in _IID_PLANS:
            _IID_PLANS[key] = CompiledPIModel(PIModel(full, []), host_only=True)
        plan = _IID_PLANS[key]
        Ac = np.ascontiguousarray(A.reshape(-1, order="F"))
        out = np.zeros(len(full), dtype=np.complex128)
        N.check(N.lib().pimi_iid_state(plan._plan, N.ptr(Ac), N.ptr(out)))
    if basis.complete:
        return PIState(basis, out)
    data = np.zeros(len(basis), dtype=np.complex128)
    for s, p in enumerate(basis.sectors):
        fs = full.sectors.index(p)
        data[basis.offsets[s]:basis.offsets[s + 1]] = out[full.offsets[fs]:full.offsets[fs + 1]]
    if Np == 0:
        tr_b = float(data.sum().real)
    else:
        idx, tv = _IID_PLANS[(Np, d)].trace_functional()
        keep = np.zeros(len(full), dtype=bool)
        for p in basis.sectors:
            fs = full.sectors.index(p)
            keep[full.offsets[fs]:full.offsets[fs + 1]] = True
        tr_b = float(sum(tv[k] * out[i_].real for k, i_ in enumerate(idx) if keep[i_]))
    if not abs(tr_b - 1.0) <= atol + rtol:
        raise ValueError(f"basis omits Schur-sector support of rho^N: represented trace={tr_b}")
    return PIState(basis, data)


def expect(op, state):
    return complex(np.vdot(op.data, state.data))


# ---------------------------------------------------------------------------
# compile(model; backend=:mi355x)
# ---------------------------------------------------------------------------

@dataclass
class GMRESAlgorithm:
    """GMRESAlgorithm(krylovdim=30, maxiter=500, preconditioner=nothing).
    preconditioner="sector" selects the device sector preconditioner (right
    preconditioning; see pimi_preconditioner in include/pimi.h)."""
    krylovdim: int = 30
    maxiter: int = 500
    preconditioner: object = None
    inner_steps: int = 8

    def __post_init__(self):
        if self.krylovdim < 1 or self.maxiter < 1:
            raise ValueError("krylovdim and maxiter must be positive")
        if self.preconditioner not in (None, "sector", "sector_inner"):
            raise ValueError("preconditioner must be nothing, \"sector\" or \"sector_inner\" on the mi355x backend")


@dataclass
class ExpvAlgorithm:
    krylovdim: int = 30
    atol: float = 1e-10
    rtol: float = 1e-8
    max_steps: int = 10_000
    safety: float = 0.9

    def __post_init__(self):
        if self.atol < 0 or self.rtol < 0:
            raise ValueError("atol and rtol must be finite nonnegative real numbers")
        if not (0 < self.safety < 1):
            raise ValueError("safety must lie strictly between zero and one")


_CONTEXTS = {}


def _context(device):
    if device not in _CONTEXTS:
        ctx = C.c_void_p()
        N.check(N.lib().pimi_context_create(C.c_int32(device), C.byref(ctx)))
        _CONTEXTS[device] = ctx
    return _CONTEXTS[device]


def _model_desc(model: PIModel):
    keep = []
    terms = (N.Term * max(len(model.terms), 1))()
    for k, t in enumerate(model.terms):
        op2 = None
        if isinstance(t.operator, CollectiveProduct):
            op = np.asfortranarray(t.operator.X, dtype=np.complex128)
            op2 = np.asfortranarray(t.operator.Y, dtype=np.complex128)
        else:
            op = np.asfortranarray(t.operator, dtype=np.complex128)
        op = np.ascontiguousarray(op.reshape(-1, order="F"))
        keep.append(op)
        terms[k].kind = t.kind
        terms[k].op = N.ptr(op)
        if op2 is not None:
            op2 = np.ascontiguousarray(op2.reshape(-1, order="F"))
            keep.append(op2)
            terms[k].op2 = N.ptr(op2)
        terms[k].rate = float(t.rate)
        terms[k].hbar = float(getattr(t, "hbar", 1.0))
    desc = N.ModelDesc(model.basis.N, model.basis.d, len(model.terms), 0, terms)
    keep.append(terms)
    return desc, keep


class CompiledPIModel:
    """compile(model; backend=:mi355x): host lowering plus one upload of the
    stencil tables.  Workspaces (one HIP stream + Krylov basis each) are made
    on demand per (precision, krylovdim)."""

    def __init__(self, model: PIModel, device=0, host_only=False):
        if not model.basis.complete:
            raise ValueError("the mi355x backend compiles complete PI bases only")
        self.model = model
        self.basis = model.basis
        self.device = device
        desc, keep = _model_desc(model)
        plan = C.c_void_p()
        ctx = None if host_only else _context(device)
        N.check(N.lib().pimi_compile(ctx, C.byref(desc), C.byref(plan)))
        self._plan = plan
        self._ws = {}
        info = N.PlanInfo()
        N.check(N.lib().pimi_plan_info_get(plan, C.byref(info)))
        self.info = {k: getattr(info, k) for k, _ in N.PlanInfo._fields_}
        if self.info["n"] != len(self.basis):
            raise RuntimeError("plan dimension disagrees with the basis")

    def __del__(self):
        try:
            for ws in self._ws.values():
                N.lib().pimi_workspace_destroy(ws)
            N.lib().pimi_plan_destroy(self._plan)
        except Exception:
            pass

    def __len__(self):
        return len(self.basis)

    def workspace(self, precision=N.PIMI_C64, krylovdim=30):
        key = (precision, krylovdim)
        if key not in self._ws:
            ws = C.c_void_p()
            N.check(N.lib().pimi_workspace_create(self._plan, C.c_int32(precision), C.c_int32(krylovdim),
                                                  C.byref(ws)))
            self._ws[key] = ws
        return self._ws[key]

    def release_workspaces(self):
        """Free the device workspaces made so far (they are re-made on demand)."""
        for ws in self._ws.values():
            N.lib().pimi_workspace_destroy(ws)
        self._ws.clear()

    def prepare(self):
        """Build the stationary-solver tables (sector preconditioner factors,
        sector weighting) now rather than at the first solve."""
        N.check(N.lib().pimi_plan_prepare(self._plan))
        return self

    def trace_functional(self):
        nt = sum(self.basis.dims)
        idx = np.zeros(nt, dtype=np.int64)
        val = np.zeros(nt, dtype=np.float64)
        N.check(N.lib().pimi_trace_functional(self._plan, N.ptr(idx), N.ptr(val)))
        return idx, val

    def collective_operator(self, X) -> PIOperator:
        X = np.ascontiguousarray(np.asarray(X, dtype=np.complex128).reshape(-1, order="F"))
        out = np.zeros(len(self.basis), dtype=np.complex128)
        N.check(N.lib().pimi_collective_operator(self._plan, N.ptr(X), N.ptr(out)))
        return PIOperator(self.basis, out)

    def apply(self, x, precision=N.PIMI_C64, adjoint=False):
        """y = L x (or L^dagger x, the apply_adjoint! of src/liouvillian.jl:3275)
        on the device (host vectors in and out; one or more columns)."""
        x = np.asarray(x, dtype=np.complex128)
        cols = 1 if x.ndim == 1 else x.shape[1]
        xf = np.ascontiguousarray(x.reshape(len(self.basis), cols).T)  # column-major blocks
        y = np.empty_like(xf)
        ws = self.workspace(precision)
        fn = N.lib().pimi_apply_adjoint if adjoint else N.lib().pimi_apply
        N.check(fn(ws, N.ptr(xf), N.ptr(y), C.c_int64(cols), C.c_int32(0)))
        y = y.T
        return y.reshape(x.shape)

    def autotune(self, precision=N.PIMI_C64, ncols=8, reps=50):
        """Pick the banded apply's tile widths for this plan by timing 4-, 8-
        and 16-column tiles on the device (one vector and `ncols` vectors);
        the choice holds for every later apply in `precision`.  Returns the
        (single-vector, multi-column) widths, (0, 0) for ELL-form plans."""
        out = (C.c_int32 * 2)()
        N.check(N.lib().pimi_workspace_autotune_apply(self.workspace(precision), C.c_int64(ncols), C.c_int32(reps),
                                                       out))
        return int(out[0]), int(out[1])

    def last_apply_ms(self, precision=N.PIMI_C64):
        ms = C.c_double()
        N.check(N.lib().pimi_workspace_last_apply_ms(self.workspace(precision), C.byref(ms)))
        return ms.value

    def export_fused(self):
        """The lowered fused plan as flat arrays (same content as the
        reference's FusedStaticPIKernel)."""
        sz = np.zeros(12, dtype=np.int64)
        N.check(N.lib().pimi_plan_export_sizes(self._plan, N.ptr(sz)))
        nblk, ncp, nnz, ncg, nob, nbr, nsup, ndense, has_h, has_q, ns, n = (int(v) for v in sz)
        a = dict(cp_base=np.zeros(max(nblk, 1), np.int64), colptr=np.zeros(max(ncp, 1), np.int64),
                 rows=np.zeros(max(nnz, 1), np.int32), vals=np.zeros(2 * max(nnz, 1)),
                 cg_scale=np.zeros(2 * max(ncg, 1)), ob_scale=np.zeros(2 * max(nob, 1)),
                 ob_br_base=np.zeros(nob + 1, np.int64), br_out=np.zeros(max(nbr, 1), np.int32),
                 br_in=np.zeros(max(nbr, 1), np.int32), br_r=np.zeros(max(nbr, 1)),
                 br_use_support=np.zeros(max(nbr, 1), np.int32), br_sup_base=np.zeros(nbr + 1, np.int64),
                 sup_rows=np.zeros(max(nsup, 1), np.int32), sup_cols=np.zeros(max(nsup, 1), np.int32),
                 sup_vals=np.zeros(2 * max(nsup, 1)), br_dense_base=np.zeros(max(nbr, 1), np.int64),
                 dense=np.zeros(2 * max(ndense, 1)))
        N.check(N.lib().pimi_plan_export(self._plan, *(N.ptr(a[k]) for k in (
            "cp_base", "colptr", "rows", "vals", "cg_scale", "ob_scale", "ob_br_base", "br_out", "br_in", "br_r",
            "br_use_support", "br_sup_base", "sup_rows", "sup_cols", "sup_vals", "br_dense_base", "dense"))))
        a.update(nblk=nblk, ncg=ncg, nob=nob, nbr=nbr, has_h=has_h, has_q=has_q, ns=ns, n=n)
        return a


def compile(model: PIModel, backend="mi355x", device=0, host_only=False):
    if backend not in ("mi355x", ":mi355x"):
        raise ValueError("this package provides the :mi355x backend only")
    return CompiledPIModel(model, device=device, host_only=host_only)


def collective_operator(x, X):
    if isinstance(x, CompiledPIModel):
        return x.collective_operator(X)
    if isinstance(x, PIBasis):
        return CollectiveOperator(x, X)
    raise TypeError("collective_operator needs a compiled model or a PIBasis")


# ---------------------------------------------------------------------------
# stationary_state / solve_dynamics
# ---------------------------------------------------------------------------

@dataclass
class SteadyStateResult:
    state: PIState
    info: dict
    algorithm: object


def stationary_state(compiled: CompiledPIModel, algorithm=None, atol=1e-10, rtol=1e-8, operator_scale=None,
                     return_info=False, precision=N.PIMI_C64, inner_rtol=1e-5, max_refinements=50,
                     criterion="reference", weighting=None, initial_state=None):
    """Trace-fixed stationary state by restarted GMRES on the device
    (src/krylov.jl:797-891).  Raises NotConverged like the reference.
    initial_state (a PIState or coefficient vector) replaces the default
    start v = t/(t.t), as the reference's keyword of that name.

    weighting="sector" iterates on u = S x with S = sqrt(f^nu) on sector nu,
    so that residual norms weigh sectors by multiplicity as observables do
    (needs criterion="backward" or "weighted"; "weighted" stops and accepts on
    the weighted relative residual |S L x|/(s |S x|), which weighs sectors as
    observables do); None iterates on the PI coefficients as the
    reference does."""
    algorithm = algorithm or GMRESAlgorithm()
    if not isinstance(algorithm, GMRESAlgorithm):
        raise ValueError("the mi355x backend solves stationary states with GMRESAlgorithm")
    ws = compiled.workspace(precision, algorithm.krylovdim)
    pc = {None: N.PRECOND_NONE, "sector": N.PRECOND_SECTOR,
          "sector_inner": N.PRECOND_SECTOR_INNER}[algorithm.preconditioner]
    if criterion not in ("reference", "backward", "weighted"):
        raise ValueError("criterion must be \"reference\", \"backward\" or \"weighted\"")
    crit = {"reference": N.CRITERION_REFERENCE, "backward": N.CRITERION_BACKWARD,
            "weighted": N.CRITERION_WEIGHTED}[criterion]
    if weighting not in (None, "sector"):
        raise ValueError("weighting must be None or \"sector\"")
    wt = N.WEIGHT_SECTOR if weighting == "sector" else N.WEIGHT_NONE
    opts = N.GmresOpts(algorithm.krylovdim, pc, algorithm.maxiter, float(atol), float(rtol),
                       float(operator_scale or 0.0), float(inner_rtol), int(max_refinements), crit,
                       int(algorithm.inner_steps), wt, 0)
    rho = np.zeros(len(compiled.basis), dtype=np.complex128)
    x0 = None
    if initial_state is not None:
        data = initial_state.data if isinstance(initial_state, PIState) else initial_state
        x0 = np.ascontiguousarray(np.asarray(data, dtype=np.complex128))
        if x0.shape != (len(compiled.basis),):
            raise ValueError("initial_state has wrong length")
    info = N.SolveInfo()
    N.check(N.lib().pimi_steady_gmres(ws, C.byref(opts), None if x0 is None else N.ptr(x0), N.ptr(rho),
                                      C.byref(info)))
    state = PIState(compiled.basis, rho)
    res = SteadyStateResult(state, info.as_dict(), algorithm)
    return res if return_info else state


def _range_decimal(step):
    """Small denominator q with step*q integral, as Julia's float ranges use
    for an exactly decimal-looking step."""
    for q in (1, 2, 4, 5, 8, 10, 16, 20, 25, 32, 40, 50, 64, 100, 125, 128, 200, 250, 256, 500, 1000, 10000,
              100000, 1000000):
        k = round(step * q)
        if k != 0 and k / q == step:
            return k, q
    return None


def saved_times(tspan, saveat=None):
    """src/highlevel.jl:434-465: t0:saveat:t1 with t1 appended if missing."""
    t0, t1 = float(tspan[0]), float(tspan[1])
    if t1 < t0:
        raise ValueError("tspan must be ordered")
    if saveat is None:
        return np.array([t0, t1])
    if np.ndim(saveat) == 0:
        step = float(saveat)
        if not (math.isfinite(step) and step > 0):
            raise ValueError("saveat must be finite and positive")
        rs, rt0 = _range_decimal(step), _range_decimal(t0) if t0 != 0 else (0, 1)
        if rs is not None and rt0 is not None:
            q = math.lcm(rs[1], rt0[1])
            a, s = round(t0 * q), round(step * q)
            count = int(math.floor((t1 * q - a) / s + 1e-9)) + 1
            ts = [(a + i * s) / q for i in range(count)]
        else:
            count = int(math.floor((t1 - t0) / step + 1e-9)) + 1
            ts = [t0 + i * step for i in range(count)]
        ts = [t for t in ts if t <= t1]
        if not ts or ts[-1] < t1:
            ts.append(t1)
        return np.array(ts)
    ts = np.asarray(saveat, dtype=float)
    if ts.size == 0:
        raise ValueError("saveat cannot be empty")
    if ts[0] != t0 or ts[-1] != t1:
        raise ValueError("explicit saveat times must include both endpoints of tspan")
    if np.any(np.diff(ts) < 0):
        raise ValueError("saveat times must be nondecreasing")
    return ts


@dataclass
class DynamicsStreamResult:
    """Saved times, observable series and the final state; `states` holds
    the state at every saved time when save_states was requested (the
    reference's DynamicsResult / DynamicsStreamResult states), else None."""
    t: np.ndarray
    values: dict
    final_state: PIState
    info: dict
    states: list | None = None


def _dynamics_inputs(compiled, rho0, tspan, saveat, algorithm, observables):
    """Saved times, observable names, the observables as rows of PI vectors
    (or None) and the native options, validated as solve_dynamics does
    (src/highlevel.jl:559-608)."""
    if not isinstance(algorithm, ExpvAlgorithm):
        raise ValueError("the mi355x backend propagates with ExpvAlgorithm")
    if rho0.basis != compiled.basis:
        raise ValueError("Liouvillian source and initial state use incompatible PI bases")
    ts = saved_times(tspan, saveat)
    names, vecs = [], []
    items = (observables.items() if isinstance(observables, dict) else [("observable", observables)]) \
        if observables else []
    for name, op in items:
        names.append(name)
        if isinstance(op, PIState):
            vecs.append(np.asarray(op.data, dtype=np.complex128))
        else:
            vecs.append(compiled.collective_operator(op).data)
    obs = np.ascontiguousarray(np.stack(vecs)) if vecs else None
    opts = N.ExpvOpts(algorithm.krylovdim, 0, float(algorithm.atol), float(algorithm.rtol), float(algorithm.safety),
                      int(algorithm.max_steps))
    return ts, names, obs, opts


def solve_dynamics(compiled: CompiledPIModel, rho0: PIState, tspan, saveat=None, algorithm=None,
                   observables=None, save_states=False):
    """Expv dynamics (src/highlevel.jl:559-733): observables streamed as
    device dots at every saved time; with save_states=True the state at every
    saved time is also returned (one device-to-host copy per saved time, the
    reference's per-interval krylov_expv! with copy(current)).  As in the
    reference, save_states=False needs at least one observable."""
    algorithm = algorithm or ExpvAlgorithm()
    if not observables and not save_states:
        raise ValueError("save_states=false requires at least one observable")
    ts, names, obs, opts = _dynamics_inputs(compiled, rho0, tspan, saveat, algorithm, observables)
    out = np.zeros((len(ts), len(names)), dtype=np.complex128)
    rho = np.ascontiguousarray(rho0.data, dtype=np.complex128)
    final = np.zeros_like(rho)
    ws = compiled.workspace(N.PIMI_C64, algorithm.krylovdim)
    tsc = np.ascontiguousarray(ts, dtype=np.float64)

    def run(t_seg, x_in, obs_seg):
        info = N.SolveInfo()
        N.check(N.lib().pimi_expv_series(ws, N.ptr(x_in), N.ptr(t_seg), C.c_int64(len(t_seg)), C.byref(opts),
                                         C.c_int32(len(names)), None if obs is None else N.ptr(obs),
                                         None if obs is None else N.ptr(obs_seg), N.ptr(final), C.byref(info)))
        return info.as_dict()

    if not save_states:
        info = run(tsc, rho, out)
        states = None
    else:
        # every interval restarts krylov_expv! in the reference too, so
        # splitting the series at the saved times changes no arithmetic
        states = [PIState(compiled.basis, rho.copy())]
        info = None
        seg_out = np.zeros((2, len(names)), dtype=np.complex128)
        cur = rho.copy()
        for k in range(1, len(ts)):
            d = run(np.ascontiguousarray(tsc[k - 1:k + 1]), cur, seg_out)
            if k == 1:
                out[0] = seg_out[0]
            out[k] = seg_out[1]
            cur = final.copy()
            states.append(PIState(compiled.basis, cur.copy()))
            if info is None:
                info = d
            else:
                for key in ("accepted", "rejected", "applications", "factorizations", "iterations"):
                    info[key] += d[key]
                info["estimated_error"] += d["estimated_error"]
                info["seconds"] += d["seconds"]
                info["converged"] = info["converged"] and d["converged"]
        if info is None:  # a single saved time
            info = run(tsc, rho, out)
        final = cur
    values = {name: out[:, k].copy() for k, name in enumerate(names)}
    return DynamicsStreamResult(ts, values, PIState(compiled.basis, final), info, states)
