"""Test helpers: compare an exported libpimi fused plan with the oracle's."""
import numpy as np


def _cplx(a):
    return a[0::2] + 1j * a[1::2]


def exported_blocks(e, dims):
    """Split the exported CSC sets into dense blocks: H, Q, K[g]."""
    vals = _cplx(e["vals"])
    ns = len(dims)
    blocks = []
    for k in range(e["nblk"]):
        base = e["cp_base"][k]
        m = dims[k % ns] if True else None
        blocks.append((base, m))
    out = []
    for k, (base, _) in enumerate(blocks):
        s = k % ns
        m = dims[s]
        cp = e["colptr"][base:base + m + 1]
        A = np.zeros((m, m), dtype=complex)
        for c in range(m):
            for p in range(cp[c], cp[c + 1]):
                A[e["rows"][p], c] = vals[p]
        out.append(A)
    i = 0
    H = out[i:i + ns] if e["has_h"] else None
    i += ns if e["has_h"] else 0
    Q = out[i:i + ns] if e["has_q"] else None
    i += ns if e["has_q"] else 0
    K = [out[i + g * ns:i + (g + 1) * ns] for g in range(e["ncg"])]
    return H, Q, K


def exported_branches(e, dims):
    sup_vals = _cplx(e["sup_vals"])
    dense = _cplx(e["dense"])
    gains = []
    for g in range(e["nob"]):
        brs = []
        for b in range(e["ob_br_base"][g], e["ob_br_base"][g + 1]):
            li, ni = int(e["br_out"][b]), int(e["br_in"][b])
            ml, mn = dims[li], dims[ni]
            use = bool(e["br_use_support"][b])
            if use:
                sup = [(int(e["sup_rows"][k]), int(e["sup_cols"][k]), sup_vals[k])
                       for k in range(e["br_sup_base"][b], e["br_sup_base"][b + 1])]
                C = np.zeros((ml, mn), dtype=complex)
                for r, c, v in sup:
                    C[r, c] = v
            else:
                o = e["br_dense_base"][b]
                C = dense[o:o + ml * mn].reshape(ml, mn, order="F")
                sup = None
            brs.append((li, ni, float(e["br_r"][b]), C, use, sup))
        gains.append((complex(_cplx(e["ob_scale"])[g]), brs))
    return gains
