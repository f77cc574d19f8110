"""The allocation-free entry point (pimi_conditional_action_noalloc) against
the entry point it varies (pimi_conditional_action), through the C ABI, on
two calls in a row (the second with fewer columns): the drifts must be
bitwise equal (same kernels, same arithmetic) and the intensities equal."""
import ctypes as C

import numpy as np
import pytest

import pimi
import pimi._native as N

pytestmark = pytest.mark.gpu
SP2 = pimi.spin_matrices(2)


def _model(N_):
    return pimi.qubit_ensemble_model(N_, hamiltonian=SP2.jz + SP2.jx, emission=0.1, dephasing=0.2, pumping=0.02,
                                     collective_emission=0.1)


def _physical(comp):
    """Density matrices with coherences: |e><e|^N, |g><g|^N and two states of
    the driven dynamics (the intensities of random vectors need not be
    positive, which the sign check rejects)."""
    b = comp.basis
    e, g = pimi.computational_product_state(b, 2), pimi.computational_product_state(b, 1)
    dyn = pimi.solve_dynamics(comp, e, (0.0, 0.7), saveat=0.35, algorithm=pimi.ExpvAlgorithm(30, 0.0, 1e-12),
                              save_states=True)
    return [e.data, g.data, dyn.states[1].data, dyn.states[2].data]


def _states(basis_states, ncols, seed):
    """Row c = column c of the n x ncols matrix: a random convex mixture."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ncols):
        w = rng.uniform(0.1, 1.0, len(basis_states))
        w /= w.sum()
        out.append(sum(wi * s for wi, s in zip(w, basis_states)))
    return np.ascontiguousarray(np.array(out, dtype=np.complex128))


class _DeviceBuffer:
    """Device memory for the test's inputs and outputs, straight from the HIP
    runtime (the caller allocates its buffers; the entry point does not)."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            N.lib()  # the runtime libpimi already loaded: the same library, not a second copy
            for name in ("libamdhip64.so.7", "libamdhip64.so"):
                try:
                    cls._hip = C.CDLL(name)
                    break
                except OSError:
                    continue
        return cls._hip

    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip().hipMemcpy(self.p, N.ptr(a), C.c_size_t(a.nbytes), 1) == 0  # host to device

    def get(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert self.hip().hipMemcpy(N.ptr(out), self.p, C.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def free(self):
        self.hip().hipFree(self.p)


def test_noalloc_conditional_action_matches_allocating_entry_point():
    comp = pimi.compile(_model(20))
    n = comp.info["n"]
    ws = comp.workspace(pimi.PIMI_C64, 1)
    lib = N.lib()
    phys = _physical(comp)
    nmax = 9
    dX, dY = _DeviceBuffer(16 * n * nmax), _DeviceBuffer(16 * n * nmax)
    dI = _DeviceBuffer(8 * nmax)
    try:
        # before the plan's trajectory tables exist the entry point refuses (building them allocates)
        X0 = _states(phys, 2, 5)
        dX.put(np.concatenate([X0.reshape(-1), np.zeros(n * (nmax - 2), dtype=np.complex128)]))
        st = lib.pimi_conditional_action_noalloc(ws, dX.p, dY.p, C.c_int64(2), dI.p)
        assert st == N.PIMI_EINVAL
        for ncols, seed in ((nmax, 1), (4, 2)):
            print(f"noalloc conditional action ncols={ncols}", flush=True)
            X = _states(phys, ncols, seed)
            Y_ref = np.empty_like(X)
            lam_ref = np.empty(ncols)
            N.check(lib.pimi_conditional_action(ws, N.ptr(X), N.ptr(Y_ref), C.c_int64(ncols), N.ptr(lam_ref),
                                                C.c_int32(0)))
            pad = np.zeros(n * (nmax - ncols), dtype=np.complex128)
            dX.put(np.concatenate([X.reshape(-1), pad]))
            dY.put(np.full(n * nmax, np.nan + 0j))
            dI.put(np.full(nmax, np.nan))
            N.check(lib.pimi_conditional_action_noalloc(ws, dX.p, dY.p, C.c_int64(ncols), dI.p))
            Y = dY.get(np.complex128, n * nmax)[:n * ncols].reshape(ncols, n)
            lam = dI.get(np.float64, nmax)[:ncols]
            np.testing.assert_array_equal(Y, Y_ref)
            np.testing.assert_array_equal(lam, lam_ref)
            assert np.all(lam > 0)
        # more columns than the fixed scratch holds: refused before anything runs
        st = lib.pimi_conditional_action_noalloc(ws, dX.p, dY.p, C.c_int64(257), dI.p)
        assert st == N.PIMI_EINVAL
    finally:
        dX.free()
        dY.free()
        dI.free()
