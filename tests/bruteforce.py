"""Brute-force full-Hilbert-space reference for small N (test helper).

Builds the d^N-dimensional Lindbladian of a PI model directly from Kronecker
products, with no Schur transform, so it shares no code with the oracle's
lowering.  PI coordinates are linked to the full space only through physical
pairings Tr(O^dagger rho) with O ranging over products of collective
operators: by Schur--Weyl duality these span the whole permutation-invariant
operator space, and in the orthonormal PI coordinates the pairing is the
Euclidean inner product sum(conj(O_pi) * rho_pi).
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def local_op(X, i, N, d):
    out = np.eye(1, dtype=complex)
    for k in range(N):
        out = np.kron(out, X if k == i else np.eye(d))
    return out


def collective(X, N, d):
    return sum(local_op(X, i, N, d) for i in range(N))


def superop_left(A):
    return np.kron(np.eye(A.shape[0]), A)


def superop_right(B):
    return np.kron(B.T, np.eye(B.shape[0]))


def dissipator(A):
    AdA = A.conj().T @ A
    return np.kron(A.conj(), A) - 0.5 * (superop_left(AdA) + superop_right(AdA))


def lindbladian(N, d, terms):
    """terms: list of (kind, op, rate) with kind in local_h, collective_h,
    local_jump, collective_jump, full_h (op already a d^N matrix)."""
    D = d ** N
    Lsup = np.zeros((D * D, D * D), dtype=complex)
    for kind, op, rate in terms:
        if kind in ("local_h", "collective_h"):
            H = rate * collective(op, N, d)
            Lsup += -1j * (superop_left(H) - superop_right(H))
        elif kind == "full_h":
            H = rate * op
            Lsup += -1j * (superop_left(H) - superop_right(H))
        elif kind == "local_jump":
            for i in range(N):
                Lsup += rate * dissipator(local_op(op, i, N, d))
        elif kind == "collective_jump":
            Lsup += rate * dissipator(collective(op, N, d))
        else:
            raise ValueError(kind)
    return Lsup


def vec(X):
    return X.reshape(-1, order="F")


def unvec(v, D):
    return v.reshape(D, D, order="F")


def dicke_state(counts, d):
    """Normalized symmetric state with the given level occupations (levels
    ordered as the local basis)."""
    N = sum(counts)
    labels = [l for l in range(d) for _ in range(counts[l])]
    psi = np.zeros(d ** N, dtype=complex)
    perms = set(itertools.permutations(labels))
    for p in perms:
        idx = 0
        for l in p:
            idx = idx * d + l
        psi[idx] += 1.0
    return psi / math.sqrt(len(perms))


def unit(d, i, j):
    E = np.zeros((d, d), dtype=complex)
    E[i, j] = 1.0
    return E


def collective_monomials(d, degree):
    """Words (tuples of (i,j)) in the collective units E_ij up to `degree`."""
    gens = [(i, j) for i in range(d) for j in range(d)]
    words = [()]
    for k in range(1, degree + 1):
        words += list(itertools.product(gens, repeat=k))
    return words
