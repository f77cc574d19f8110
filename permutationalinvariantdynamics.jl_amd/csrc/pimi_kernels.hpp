// Device kernels for the PI Liouvillian apply and the Krylov BLAS-1/2 steps.
//
// Complex vectors are arrays of cx<R> (R = double or float), 8/16-byte pairs.
// All reductions are fixed-shape trees (wave shuffles + LDS), so results are
// deterministic from run to run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace pimi {

template <typename R>
struct alignas(2 * sizeof(R)) cx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// conj(a) * b
template <typename R>
__device__ __forceinline__ cx<R> cjmul(cx<R> a, cx<R> b) { return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x}; }

// Address spaces: tables that no kernel writes while it runs are read through
// the constant address space (wave-uniform reads become scalar loads, also
// when the table pointer itself was loaded from a descriptor); vectors
// through the global one (global_load, not flat_load).
// (The attributes apply to the device pass only; the host pass sees plain
// pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
#define PIMI_CONST __attribute__((address_space(4)))
#define PIMI_GLOBAL __attribute__((address_space(1)))
#else
#define PIMI_CONST
#define PIMI_GLOBAL
#endif
template <typename T>
__device__ __forceinline__ const PIMI_CONST T *cptr(const T *p) { return (const PIMI_CONST T *)p; }
template <typename T>
__device__ __forceinline__ const PIMI_GLOBAL T *gptr(const T *p) { return (const PIMI_GLOBAL T *)p; }
template <typename T>
__device__ __forceinline__ PIMI_GLOBAL T *gptr_w(T *p) { return (PIMI_GLOBAL T *)p; }

constexpr int kApplyBlock = 256;
constexpr int kRedBlock = 256;
constexpr int kNumXcd = 8;

// One stencil entry of an output sector: its contribution to Y[a, b] is
//   s * A[a] * conj(B[b]) * X_in[colA(a), colB(b)],
// where the A and B factors are ELL slots: row r of the slot starting at q
// is (col[q + r], val[q + r]).  An identity factor points at the identity
// slot (col[id + r] = r, val[id + r] = 1), so every entry has the same form:
// a stencil term A X_in B^H with ELL widths wa, wb becomes wa * wb entries.
// Each sector's list is padded with zero-scale entries to a multiple of
// kEntryChunk, so a wave walks it in whole chunks.
struct ApplyEntry {
    int64_t in_off;     // offset of the input block
    int32_t qa, qb;     // A-side and B-side slot starts
    int32_t mn;         // input block size
    int32_t in_sector;
    double s_re, s_im;  // term scale
};
constexpr int kEntryChunk = 8;

// An apply tile: rows a0..a0+63 of columns b0..b0+nb-1 (nb <= kApplyCols) of
// sector `sector`'s output block, one wave; each lane computes its row in the
// nb columns.  Entries are processed a chunk at a time: all slot-table loads
// of the chunk, then all input gathers, then the arithmetic, so a wave keeps
// a chunk's worth of independent loads in flight instead of walking a chain
// of dependent loads term by term (the single-vector apply is latency-bound).
// Entry descriptors and the B-side (column) slots are wave-uniform (scalar
// loads); the A-side (row) slots are loaded once per entry for all nb
// columns of a lane.  Every output sums its entries in list order, as
// (s * conj(B)) * A * X.
struct ApplyTile {
    int32_t sector, b0, a0, nb;
};
constexpr int kApplyWave = 64;
constexpr int kApplyCols = 4;  // columns per lane of the batched apply (one for a single vector)

template <typename R, int CB>
__device__ __forceinline__ void apply_tile(const cx<R> *xc_, cx<R> *__restrict__ yc, R out_scale, ApplyTile tile,
                                           const int64_t *__restrict__ offs, const int32_t *dims_,
                                           const int32_t *entry_base_, const ApplyEntry *entries_,
                                           const int32_t *col_, const cx<R> *val_) {
    constexpr int CH = CB == 1 ? kEntryChunk : kEntryChunk / 2;
    // wave-uniform tables (entries, B-side slots) as scalar loads, per-lane
    // ones (A-side slots, input gathers) as global loads
    const PIMI_CONST ApplyEntry *entries = cptr(entries_);
    const PIMI_CONST int32_t *colS = cptr(col_);
    const PIMI_CONST cx<R> *valS = cptr(val_);
    const PIMI_GLOBAL int32_t *col = gptr(col_);
    const PIMI_GLOBAL cx<R> *val = gptr(val_);
    const PIMI_GLOBAL cx<R> *xc = gptr(xc_);
    const PIMI_CONST int32_t *entry_base = cptr(entry_base_);
    const PIMI_CONST int32_t *dims = cptr(dims_);
    const int s = __builtin_amdgcn_readfirstlane(tile.sector);
    const int b0 = __builtin_amdgcn_readfirstlane(tile.b0);
    const int nb = __builtin_amdgcn_readfirstlane(tile.nb);
    const int m = dims[s];
    const int arow = tile.a0 + (int)(threadIdx.x & (kApplyWave - 1));
    const bool active = arow < m;
    const int a = active ? arow : 0;  // inactive lanes evaluate row 0 (in bounds) and store nothing
    cx<R> acc[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) acc[j] = cx<R>{0, 0};
    const int e0 = entry_base[s], e1 = entry_base[s + 1];
    for (int e = e0; e < e1; e += CH) {
        int ca[CH];
        cx<R> va[CH];
        cx<R> xv[CH][CB];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int q = entries[e + k].qa + a;
            ca[k] = col[q];
            va[k] = val[q];
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const ApplyEntry E = entries[e + k];
            const PIMI_GLOBAL cx<R> *Xn = xc + E.in_off;
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                const int cb = colS[E.qb + b0 + (j < nb ? j : 0)];
                xv[k][j] = Xn[(uint32_t)(ca[k] + cb * E.mn)];
            }
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const ApplyEntry E = entries[e + k];
            const cx<R> sc = {(R)E.s_re, (R)E.s_im};
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                const cx<R> vb = valS[E.qb + b0 + (j < nb ? j : 0)];
                const cx<R> coef = cmul(cjmul(vb, sc), va[k]);
                acc[j] = cadd(acc[j], cmul(coef, xv[k][j]));
            }
        }
    }
    if (active) {
        const int64_t o = offs[s] + a;
#pragma unroll
        for (int j = 0; j < CB; ++j)
            if (j < nb) yc[o + (int64_t)(b0 + j) * m] = cx<R>{out_scale * acc[j].x, out_scale * acc[j].y};
    }
}

// y = out_scale * L x for ncols column vectors; one wave per single-column
// tile (the most waves for one vector's latency hiding), tiles spread so
// that XCD k (workgroup id mod 8) handles one contiguous slab.
template <typename R>
__global__ __launch_bounds__(kApplyWave) void k_apply(
    const cx<R> *__restrict__ x, cx<R> *__restrict__ y, int64_t n, const ApplyTile *__restrict__ tiles, int64_t ntiles,
    int64_t tiles_per_xcd, const int64_t *__restrict__ offs, const int32_t *__restrict__ dims,
    const int32_t *__restrict__ entry_base, const ApplyEntry *__restrict__ entries, const int32_t *__restrict__ col,
    const cx<R> *__restrict__ val, R out_scale, int64_t ncols) {
    const int64_t bid = blockIdx.x;
    const int64_t t = (bid % kNumXcd) * tiles_per_xcd + bid / kNumXcd;
    if (t >= ntiles) return;
    const ApplyTile tile = cptr(tiles)[t];
    for (int64_t c = blockIdx.y; c < ncols; c += gridDim.y)
        apply_tile<R, 1>(x + c * n, y + c * n, out_scale, tile, offs, dims, entry_base, entries, col, val);
}

// ---------------------------------------------------------------------------
// Banded (diagonal) form of the stencil apply.
//
// When every factor of every term is a sum of a few diagonals,
//   A = sum_dA diag-shift(cA_dA),  A[a, a + dA] = cA_dA[a],
// a term s A X_in B^H adds to Y[a, b] the products
//   s cA_dA[a] conj(cB_dB[b]) X_in[a + dA, b + dB].
// The qubit sectors are of this form (H, Q and J- tridiagonal or narrower,
// every one-body contraction one diagonal).  Entries are grouped by (input
// sector, dA); a group's lanes read one strip X_in[a + dA, b0 + c0 + k],
// k = 0 .. TB + span - 1, with plain coalesced loads (no column tables), and
// every entry of the group combines the strip at its own column offset
// dB - c0 in 0..2.  Coefficient vectors are zero wherever a + dA or b + dB
// leaves the input block, so clamped load indices never change a result.
// Kinds:  left (B = I): Y += cA'[a] X;  right (A = I): Y += w[b] X;
//         sandwich: Y += (cA'[a] w[b]) X;  real sandwich: the same with real
//         cA', w (two real FMAs per product instead of four complex ones);
//         imaginary left / right: a left / right entry whose coefficients
//         are purely imaginary (i c X: two real FMAs), stored as the real c;
//         left + right: a left and a right entry on the same strip column,
//         (cA'[a] + w[b]) X;  real sandwich pair: two real sandwiches on the
//         same strip column, (cA1[a] w1[b] + cA2[a] w2[b]) X.
// cA' carries the term scale, w[b] = conj(cB[b]) (times the scale for right).
// For qubit plans (tridiagonal H, diagonal loss, one-diagonal contractions)
// the builder turns the off-diagonal left/right entries into imaginary
// ones, merges the two diagonal ones, and pairs the two children's
// sandwiches of each branch, which takes about a third of the arithmetic.
// Every output sums its groups in list order and each group's entries in
// list order with explicit FMAs, so the result of an output does not depend
// on the tile width TB.
struct DiaGroup {
    int64_t in_off;     // offset of the input block
    int32_t mn;         // input block size
    int32_t dA;         // row shift of the strip
    int32_t c0;         // first strip column relative to the tile's first column
    int32_t span;       // strip columns beyond TB (0..2)
    int32_t e0, e1;     // entry range
};
struct DiaEntry {
    int32_t kind;       // 0 left, 1 right, 2 sandwich, 3 real sandwich, 4 imaginary left, 5 imaginary right,
                        // 6 left + right on one strip column, 7 two real sandwiches on one strip column
    int32_t dB;         // strip column offset, 0..2
    int32_t qa, qb;     // coefficient vectors: per output row (A side), per output column (B side)
    int32_t qa2, qb2;   // kind 7: the second sandwich's vectors
};
constexpr int kDiaWaves = 4;  // tiles (waves) per workgroup
constexpr int kNumDiaWidths = 3;
constexpr int kDiaWidths[kNumDiaWidths] = {4, 8, 16};  // tile columns TB
constexpr int kDiaCoefPad = 16;  // coefficient vectors: aligned and zero-padded to this many elements

__device__ __forceinline__ float rfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double rfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename R>
__device__ __forceinline__ cx<R> cfma(cx<R> a, cx<R> b, cx<R> c) {
    c.x = rfma(a.x, b.x, c.x);
    c.x = rfma(-a.y, b.y, c.x);
    c.y = rfma(a.x, b.y, c.y);
    c.y = rfma(a.y, b.x, c.y);
    return c;
}
template <typename R>
__device__ __forceinline__ cx<R> cmulf(cx<R> a, cx<R> b) {
    return cx<R>{rfma(a.x, b.x, -a.y * b.y), rfma(a.x, b.y, a.y * b.x)};
}


// TB consecutive column coefficients, read as one block (complex, or the
// packed reals of a real sandwich)
template <typename R, int TB>
struct DiaCols {
    cx<R> v[TB];
};
template <typename R, int TB>
struct DiaColsR {
    R v[TB];
};

// Per-lane (row) coefficients come through a raw buffer over the pool: the
// lane's offset (a times the element size) is fixed for the tile, the
// vector's start is a scalar offset, so a load needs no address arithmetic.
template <typename R>
__device__ __forceinline__ cx<R> dia_row(__amdgpu_buffer_rsrc_t pool, int voff, int q) {
    if constexpr (sizeof(R) == 8)
        return __builtin_bit_cast(cx<R>, __builtin_amdgcn_raw_buffer_load_b128(pool, voff, q * 16, 0));
    else
        return __builtin_bit_cast(cx<R>, __builtin_amdgcn_raw_buffer_load_b64(pool, voff, q * 8, 0));
}
template <typename R>
__device__ __forceinline__ R dia_row_real(__amdgpu_buffer_rsrc_t pool, int voff_real, int q) {
    if constexpr (sizeof(R) == 8)
        return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(pool, voff_real, q * 16, 0));
    else
        return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(pool, voff_real, q * 8, 0));
}

template <typename R, int TB, int KIND, int DB>
__device__ __forceinline__ void dia_entry(cx<R> (&acc)[TB], const cx<R> (&st)[TB + 2], const DiaEntry &E,
                                          __amdgpu_buffer_rsrc_t pool, const PIMI_CONST cx<R> *coefs, int va,
                                          int vr, int b0) {
    if constexpr (KIND == 4) {  // i c[a] X
        const R c = dia_row_real<R>(pool, vr, E.qa);
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            acc[j].x = rfma(-c, st[j + DB].y, acc[j].x);
            acc[j].y = rfma(c, st[j + DB].x, acc[j].y);
        }
    } else if constexpr (KIND == 5) {  // i w[b] X
        const DiaColsR<R, TB> w = *(const PIMI_CONST DiaColsR<R, TB> *)((const PIMI_CONST R *)(coefs + E.qb) + b0);
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            acc[j].x = rfma(-w.v[j], st[j + DB].y, acc[j].x);
            acc[j].y = rfma(w.v[j], st[j + DB].x, acc[j].y);
        }
    } else if constexpr (KIND == 6) {  // (cA[a] + w[b]) X
        const cx<R> ca = dia_row<R>(pool, va, E.qa);
        const DiaCols<R, TB> w = *(const PIMI_CONST DiaCols<R, TB> *)(coefs + E.qb + b0);
#pragma unroll
        for (int j = 0; j < TB; ++j) acc[j] = cfma(cx<R>{ca.x + w.v[j].x, ca.y + w.v[j].y}, st[j + DB], acc[j]);
    } else if constexpr (KIND == 7) {  // (cA1[a] w1[b] + cA2[a] w2[b]) X, all real
        const R c1 = dia_row_real<R>(pool, vr, E.qa), c2 = dia_row_real<R>(pool, vr, E.qa2);
        const DiaColsR<R, TB> w1 = *(const PIMI_CONST DiaColsR<R, TB> *)((const PIMI_CONST R *)(coefs + E.qb) + b0);
        const DiaColsR<R, TB> w2 = *(const PIMI_CONST DiaColsR<R, TB> *)((const PIMI_CONST R *)(coefs + E.qb2) + b0);
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            const R c = rfma(c1, w1.v[j], c2 * w2.v[j]);
            acc[j].x = rfma(c, st[j + DB].x, acc[j].x);
            acc[j].y = rfma(c, st[j + DB].y, acc[j].y);
        }
    } else if constexpr (KIND == 0) {
        const cx<R> ca = dia_row<R>(pool, va, E.qa);
#pragma unroll
        for (int j = 0; j < TB; ++j) acc[j] = cfma(ca, st[j + DB], acc[j]);
    } else if constexpr (KIND == 3) {
        const DiaColsR<R, TB> w = *(const PIMI_CONST DiaColsR<R, TB> *)((const PIMI_CONST R *)(coefs + E.qb) + b0);
        const R ca = dia_row_real<R>(pool, vr, E.qa);
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            const R c = ca * w.v[j];
            acc[j].x = rfma(c, st[j + DB].x, acc[j].x);
            acc[j].y = rfma(c, st[j + DB].y, acc[j].y);
        }
    } else {
        const DiaCols<R, TB> w = *(const PIMI_CONST DiaCols<R, TB> *)(coefs + E.qb + b0);
        if constexpr (KIND == 1) {
#pragma unroll
            for (int j = 0; j < TB; ++j) acc[j] = cfma(w.v[j], st[j + DB], acc[j]);
        } else {
            const cx<R> ca = dia_row<R>(pool, va, E.qa);
#pragma unroll
            for (int j = 0; j < TB; ++j) acc[j] = cfma(cmulf(ca, w.v[j]), st[j + DB], acc[j]);
        }
    }
}

// the strip X_in[r, cfirst + k], k < NK, of one group through a raw buffer
// over the input block, the whole offset ((clamped) row + column times the
// block size) in voffset, which the hardware range check covers: a column
// outside the block gives an offset outside the resource (a negative one
// reads as a huge unsigned offset), which reads zero -- exactly the zero the
// coefficient vectors hold there.  (The scalar offset operand is not used:
// the range check need not cover it.)  0x00020000: the gfx9-family word 3 of
// a raw buffer resource.
template <typename R, int TB, int NK>
__device__ __forceinline__ void dia_strip(cx<R> (&st)[TB + 2], const cx<R> *Xn, int r, int mn, int cfirst) {
    constexpr int es = (int)sizeof(cx<R>);
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)Xn, (short)0, mn * mn * es, 0x00020000);
    const int vr = r * es;
    const int step = mn * es;
    const int s0 = cfirst * step;
#pragma unroll
    for (int k = 0; k < TB + 2; ++k) {
        const int vo = vr + (s0 + k * step);
        if constexpr (sizeof(R) == 8) {
            st[k] = k < NK ? __builtin_bit_cast(cx<R>, __builtin_amdgcn_raw_buffer_load_b128(rs, vo, 0, 0))
                           : cx<R>{0, 0};
        } else {
            st[k] = k < NK ? __builtin_bit_cast(cx<R>, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, 0, 0))
                           : cx<R>{0, 0};
        }
    }
}

// One banded tile: rows a0..a0+63 (one per lane) of columns b0..b0+nb-1
// (nb <= TB) of sector s's output block.  Everything but the lane's row is
// wave-uniform: descriptors and column coefficients are scalar loads, the
// entry switch a scalar branch.
//
// SPLIT > 1: the tile's groups are dealt to the SPLIT waves of the workgroup
// (wave w takes groups w, w + SPLIT, ...), which shortens each wave's chain of
// dependent strip loads; the partial sums meet in LDS and wave 0 adds them in
// the order w = 0 .. SPLIT-1 (deterministic).  Used for the single-solve
// paths, where one vector has too few tiles to hide the chain's latency.
template <typename R, int TB, int SPLIT = 1>
__device__ __forceinline__ void dia_tile(const cx<R> *xc_, cx<R> *yc_, R out_scale, ApplyTile tile,
                                         const int64_t *offs_, const int32_t *dims_, const int32_t *gbase_,
                                         const DiaGroup *groups_, const DiaEntry *ents_, const cx<R> *coef_) {
    PIMI_GLOBAL cx<R> *yc = gptr_w(yc_);
    const PIMI_CONST int64_t *offs = cptr(offs_);
    const PIMI_CONST int32_t *dims = cptr(dims_);
    const PIMI_CONST int32_t *gbase = cptr(gbase_);
    const PIMI_CONST DiaGroup *groups = cptr(groups_);
    const PIMI_CONST DiaEntry *ents = cptr(ents_);
    const PIMI_CONST cx<R> *coefs = cptr(coef_);
    const int s = __builtin_amdgcn_readfirstlane(tile.sector);
    const int b0 = __builtin_amdgcn_readfirstlane(tile.b0);
    const int a0 = __builtin_amdgcn_readfirstlane(tile.a0);
    const int nb = __builtin_amdgcn_readfirstlane(tile.nb);
    const int m = dims[s];
    const int arow = a0 + (int)(threadIdx.x & (kApplyWave - 1));
    const bool active = arow < m;
    const int a = active ? arow : m - 1;  // inactive lanes evaluate the last row and store nothing
    // the coefficient pool as one raw buffer (every read is in range; the
    // pool is far below 2 GiB), and the lane's byte offsets into a vector:
    // complex element a, and real element a of a packed real vector
    const __amdgpu_buffer_rsrc_t pool =
        __builtin_amdgcn_make_buffer_rsrc((void *)coef_, (short)0, 0x7fffffff, 0x00020000);
    const int va = a * (int)sizeof(cx<R>), vr = a * (int)sizeof(R);
    cx<R> acc[TB];
#pragma unroll
    for (int j = 0; j < TB; ++j) acc[j] = cx<R>{0, 0};
    const int wave = SPLIT > 1 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
    const int g0 = gbase[s], g1 = gbase[s + 1];
    for (int g = g0 + wave; g < g1; g += SPLIT) {
        const DiaGroup G = groups[g];
        cx<R> st[TB + 2];
        const int r = min(max(a + G.dA, 0), G.mn - 1);
        if (G.span == 0) dia_strip<R, TB, TB>(st, xc_ + G.in_off, r, G.mn, b0 + G.c0);
        else dia_strip<R, TB, TB + 2>(st, xc_ + G.in_off, r, G.mn, b0 + G.c0);
        for (int e = G.e0; e < G.e1; ++e) {
            const DiaEntry E = ents[e];
            switch (E.kind * 4 + E.dB) {
                case 0: dia_entry<R, TB, 0, 0>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 1: dia_entry<R, TB, 0, 1>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 2: dia_entry<R, TB, 0, 2>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 4: dia_entry<R, TB, 1, 0>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 5: dia_entry<R, TB, 1, 1>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 6: dia_entry<R, TB, 1, 2>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 8: dia_entry<R, TB, 2, 0>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 9: dia_entry<R, TB, 2, 1>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 10: dia_entry<R, TB, 2, 2>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 12: dia_entry<R, TB, 3, 0>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 13: dia_entry<R, TB, 3, 1>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 14: dia_entry<R, TB, 3, 2>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 16: dia_entry<R, TB, 4, 0>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 17: dia_entry<R, TB, 4, 1>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 18: dia_entry<R, TB, 4, 2>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 20: dia_entry<R, TB, 5, 0>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 21: dia_entry<R, TB, 5, 1>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 22: dia_entry<R, TB, 5, 2>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 24: dia_entry<R, TB, 6, 0>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 25: dia_entry<R, TB, 6, 1>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 26: dia_entry<R, TB, 6, 2>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 28: dia_entry<R, TB, 7, 0>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 29: dia_entry<R, TB, 7, 1>(acc, st, E, pool, coefs, va, vr, b0); break;
                case 30: dia_entry<R, TB, 7, 2>(acc, st, E, pool, coefs, va, vr, b0); break;
                default: break;
            }
        }
    }
    if constexpr (SPLIT > 1) {
        __shared__ cx<R> part[SPLIT - 1][TB][kApplyWave];
        const int lane = threadIdx.x & (kApplyWave - 1);
        if (wave > 0)
#pragma unroll
            for (int j = 0; j < TB; ++j) part[wave - 1][j][lane] = acc[j];
        __syncthreads();
        if (wave == 0)
#pragma unroll
            for (int w = 0; w < SPLIT - 1; ++w)
#pragma unroll
                for (int j = 0; j < TB; ++j) acc[j] = cadd(acc[j], part[w][j][lane]);
        __syncthreads();  // the LDS is free for the workgroup's next column vector
        if (wave > 0) return;
    }
    if (active) {
        const int64_t o = offs[s] + a;
#pragma unroll
        for (int j = 0; j < TB; ++j)
            if (j < nb) yc[o + (int64_t)(b0 + j) * m] = cx<R>{out_scale * acc[j].x, out_scale * acc[j].y};
    }
}

// y = out_scale * L x for ncols column vectors through the banded form, one
// tile per workgroup of kDiaSplit waves that share the tile's groups; tiles
// spread so that XCD k (workgroup id mod 8) handles one contiguous slab of
// the sector-major tile list.
constexpr int kDiaSplit = 4;
template <typename R, int TB>
__global__ __launch_bounds__(kApplyWave *kDiaSplit) void k_apply_dia_split(
    const cx<R> *__restrict__ x, cx<R> *__restrict__ y, int64_t n, const ApplyTile *__restrict__ tiles,
    int64_t ntiles, int64_t blocks_per_xcd, const int64_t *__restrict__ offs, const int32_t *__restrict__ dims,
    const int32_t *__restrict__ gbase, const DiaGroup *__restrict__ groups, const DiaEntry *__restrict__ ents,
    const cx<R> *__restrict__ coef, R out_scale, int64_t ncols) {
    const int64_t bid = blockIdx.x;
    const int64_t t = (bid % kNumXcd) * blocks_per_xcd + bid / kNumXcd;
    if (t >= ntiles) return;  // the whole workgroup leaves together
    const ApplyTile tile = cptr(tiles)[t];
    for (int64_t c = blockIdx.y; c < ncols; c += gridDim.y)
        dia_tile<R, TB, kDiaSplit>(x + c * n, y + c * n, out_scale, tile, offs, dims, gbase, groups, ents, coef);
}

// y[i] = x[i] * w[sector of i], one coordinate per lane over the plan's tiles
template <typename R>
__global__ __launch_bounds__(kApplyBlock) void k_sector_scale(const cx<R> *__restrict__ x, cx<R> *__restrict__ y,
                                                              int64_t n, const int64_t *__restrict__ offs,
                                                              const int32_t *__restrict__ tile_sector,
                                                              const double *__restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * kApplyBlock + threadIdx.x;
    if (i >= n) return;
    int s = tile_sector[blockIdx.x];
    while (i >= offs[s + 1]) ++s;
    const R f = (R)w[s];
    y[i] = cx<R>{x[i].x * f, x[i].y * f};
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the block; the result is valid in every thread.  `sh` must hold
// blockDim.x/64 entries; callers separate consecutive uses with __syncthreads.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    T r = 0;
    for (int k = 0; k < nw; ++k) r += sh[k];
    return r;
}

// block_sum of N values with one barrier (the same summation order, so the
// same results); `sh` must hold N * blockDim.x/64 entries.
template <int N, typename T>
__device__ __forceinline__ void block_sums(T (&v)[N], T *sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < N; ++q) sh[q * nw + wid] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; ++q) {
        T r = 0;
        for (int k = 0; k < nw; ++k) r += sh[q * nw + k];
        v[q] = r;
    }
}

// Grouped per-sector complex GEMM over PI-layout blocks:
//   C_s = op(A_s) op(B_s) [.* D_s * dscale]   for every tile of every sector,
// op = identity (0) or conjugate transpose (1).  A tile is a TS x TS patch of
// one sector's m x m output block; tiles[k] = (sector, tile row, tile col).
constexpr int kGemmTile = 32;

struct GemmTile {
    int32_t sector, ti, tj, pad;
};

template <typename R>
__global__ __launch_bounds__(256) void k_sector_gemm(const cx<R> *__restrict__ A, const cx<R> *__restrict__ B,
                                                     cx<R> *__restrict__ Cm, const int64_t *__restrict__ offs,
                                                     const int32_t *__restrict__ dims,
                                                     const GemmTile *__restrict__ tiles, int transA, int transB,
                                                     const cx<R> *__restrict__ D, R dscale) {
    constexpr int TS = kGemmTile;
    __shared__ cx<R> As[TS][TS + 1];
    __shared__ cx<R> Bs[TS][TS + 1];
    const GemmTile t = tiles[blockIdx.x];
    const int64_t m = dims[t.sector];
    const int64_t base = offs[t.sector];
    const int64_t row0 = (int64_t)t.ti * TS, col0 = (int64_t)t.tj * TS;
    const int tid = threadIdx.x, tx = tid & 31, ty = tid >> 5;
    cx<R> acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = cx<R>{0, 0};
    for (int64_t k0 = 0; k0 < m; k0 += TS) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            int r, k;
            cx<R> v = {0, 0};
            if (transA == 0) {
                r = e & 31; k = e >> 5;
                const int64_t gr = row0 + r, gk = k0 + k;
                if (gr < m && gk < m) v = A[base + gr + gk * m];
            } else {
                k = e & 31; r = e >> 5;
                const int64_t gr = row0 + r, gk = k0 + k;
                if (gr < m && gk < m) { v = A[base + gk + gr * m]; v.y = -v.y; }
            }
            As[r][k] = v;
            int c;
            cx<R> w = {0, 0};
            if (transB == 0) {
                k = e & 31; c = e >> 5;
                const int64_t gk = k0 + k, gc = col0 + c;
                if (gk < m && gc < m) w = B[base + gk + gc * m];
            } else {
                c = e & 31; k = e >> 5;
                const int64_t gk = k0 + k, gc = col0 + c;
                if (gk < m && gc < m) { w = B[base + gc + gk * m]; w.y = -w.y; }
            }
            Bs[k][c] = w;
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < TS; ++k) {
            const cx<R> a = As[tx][k];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = cadd(acc[q], cmul(a, Bs[k][ty + 8 * q]));
        }
        __syncthreads();
    }
    const int64_t gr = row0 + tx;
    if (gr >= m) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t gc = col0 + ty + 8 * q;
        if (gc >= m) continue;
        const int64_t o = base + gr + gc * m;
        cx<R> v = acc[q];
        if (D) {
            const cx<R> dd = D[o];
            v = cmul(v, cx<R>{dd.x * dscale, dd.y * dscale});
        }
        Cm[o] = v;
    }
}

// The same grouped per-sector complex GEMM on the matrix cores: one wave per
// 16x16 output tile (tiles of 16 here), four waves per workgroup, K stepped by
// 4 with v_mfma_*_16x16x4.  Complex products use four real MFMAs:
//   Cr += Ar Br - Ai Bi,   Ci += Ar Bi + Ai Br
// (the split-K sector kernel below uses three, by Gauss's rearrangement).
// Operand lanes: A(i = lane%16, k = lane/16), B(k = lane/16, j = lane%16).
// Accumulator lanes: f32 -> row 4*(lane/16) + r, f64 -> row lane/16 + 4 r;
// column lane%16 (checked on gfx950 with a known product).
constexpr int kMfmaTile = 16;
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef double f64x4_t __attribute__((ext_vector_type(4)));

template <typename R>
struct MfmaOps;
template <>
struct MfmaOps<float> {
    typedef f32x4_t acc_t;
    static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct MfmaOps<double> {
    typedef f64x4_t acc_t;
    static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

// One 16x16 output tile of a sector GEMM, computed by one wave.
template <typename R>
__device__ __forceinline__ void gemm_tile_mfma(const cx<R> *__restrict__ A, const cx<R> *__restrict__ B,
                                               cx<R> *__restrict__ Cm, int64_t m, int64_t base, GemmTile t,
                                               int transA, int transB, const cx<R> *__restrict__ D, R dscale,
                                               int lane) {
    typedef MfmaOps<R> Ops;
    const int64_t row0 = (int64_t)t.ti * kMfmaTile, col0 = (int64_t)t.tj * kMfmaTile;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t gi = row0 + li, gj = col0 + li;
    typename Ops::acc_t cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    auto load_a = [&](int64_t gk) {
        cx<R> a = {0, 0};
        if (gi < m && gk < m) {
            if (transA) { a = A[base + gk + gi * m]; a.y = -a.y; }
            else a = A[base + gi + gk * m];
        }
        return a;
    };
    auto load_b = [&](int64_t gk) {
        cx<R> b = {0, 0};
        if (gk < m && gj < m) {
            if (transB) { b = B[base + gj + gk * m]; b.y = -b.y; }
            else b = B[base + gk + gj * m];
        }
        return b;
    };
    // K in groups of kUnroll MFMA steps: the operands of a whole group are
    // loaded before its first MFMA, so kUnroll load pairs are in flight per
    // wait instead of one (the k-loop is latency-bound at these sizes).
    constexpr int kUnroll = 4;
    for (int64_t k0 = 0; k0 < m; k0 += 4 * kUnroll) {
        cx<R> a[kUnroll], b[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            a[u] = load_a(k0 + 4 * u + lk);
            b[u] = load_b(k0 + 4 * u + lk);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            cr = Ops::mac(a[u].x, b[u].x, cr);
            cr = Ops::mac(-a[u].y, b[u].y, cr);
            ci = Ops::mac(a[u].x, b[u].y, ci);
            ci = Ops::mac(a[u].y, b[u].x, ci);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t i = row0 + Ops::row(lane, r);
        if (i >= m || gj >= m) continue;
        const int64_t o = base + i + gj * m;
        cx<R> v = {cr[r], ci[r]};
        if (D) {
            const cx<R> dd = D[o];
            v = cmul(v, cx<R>{dd.x * dscale, dd.y * dscale});
        }
        Cm[o] = v;
    }
}

// Raw buffer loads of one complex entry: the resource covers one sector block,
// offsets are 32-bit byte offsets, and an offset at or past the block's end
// reads zero (the hardware range check), which is how out-of-range tile lanes
// get their zero operands without a branch or a 64-bit address per load.
// 0x00020000: the gfx9-family word 3 of a raw (stride 0) buffer resource.
constexpr uint32_t kBufOut = 0x80000000u;  // an offset past every block (blocks < 2 GiB)
template <typename R>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t block_rsrc(const cx<R> *p, int64_t entries) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)(entries * (int64_t)sizeof(cx<R>)), 0x00020000);
}
__device__ __forceinline__ cx<double> buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, cx<double> *) {
    return __builtin_bit_cast(cx<double>, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}
__device__ __forceinline__ cx<float> buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, cx<float> *) {
    return __builtin_bit_cast(cx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

// The same 16x16 tile with its operands through raw buffers over the sector
// block: an out-of-range row or k reads zero by the hardware range check (no
// per-lane branch, 32-bit offsets).
template <typename R>
__device__ __forceinline__ void gemm_tile_mfma_buf(const cx<R> *A, const cx<R> *B, cx<R> *__restrict__ Cm,
                                                   int64_t m, int64_t base, GemmTile t, int transA, int transB,
                                                   const cx<R> *__restrict__ D, R dscale, int lane) {
    typedef MfmaOps<R> Ops;
    const int row0 = t.ti * kMfmaTile, col0 = t.tj * kMfmaTile;
    const int mi = (int)m;
    const int li = lane & 15, lk = lane >> 4;
    const int gi = row0 + li, gj = col0 + li;
    constexpr uint32_t E = sizeof(cx<R>);
    cx<R> *const tag = nullptr;
    const __amdgpu_buffer_rsrc_t ra = block_rsrc<R>(A + base, m * m);
    const __amdgpu_buffer_rsrc_t rb = block_rsrc<R>(B + base, m * m);
    typename Ops::acc_t cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    constexpr int kUnroll = 4;
    for (int k0 = 0; k0 < mi; k0 += 4 * kUnroll) {
        cx<R> a[kUnroll], b[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int gk = k0 + 4 * u + lk;
            const uint32_t ia = (uint32_t)(transA ? gk + gi * mi : gi + gk * mi) * E;
            const uint32_t ib = (uint32_t)(transB ? gj + gk * mi : gk + gj * mi) * E;
            a[u] = buf_load(ra, (gi < mi && gk < mi) ? ia : kBufOut, tag);
            b[u] = buf_load(rb, (gk < mi && gj < mi) ? ib : kBufOut, tag);
            if (transA) a[u].y = -a[u].y;
            if (transB) b[u].y = -b[u].y;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            cr = Ops::mac(a[u].x, b[u].x, cr);
            cr = Ops::mac(-a[u].y, b[u].y, cr);
            ci = Ops::mac(a[u].x, b[u].y, ci);
            ci = Ops::mac(a[u].y, b[u].x, ci);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = row0 + Ops::row(lane, r);
        if (i >= mi || gj >= mi) continue;
        const int64_t o = base + i + (int64_t)gj * mi;
        cx<R> v = {cr[r], ci[r]};
        if (D) {
            const cx<R> dd = D[o];
            v = cmul(v, cx<R>{dd.x * dscale, dd.y * dscale});
        }
        Cm[o] = v;
    }
}

// A split-K GEMM tile: the sector's size and offset travel in the descriptor,
// so a workgroup's operand loads wait for one descriptor load only.
// m = 0 marks padding.
struct MTile {
    int64_t base;
    int32_t m, ti, tj, pad;
};

// One 16x16 output tile per workgroup of four waves, K split four ways: wave w
// accumulates k-steps [w K/4, (w+1) K/4), the partials meet in LDS and wave 0
// adds them in the order w = 0..3 (deterministic).  Four times the waves and a
// quarter of each wave's dependent k-chain, for sector sizes (m <= a few
// hundred) whose GEMMs have too few tiles to fill the chip.  A wave loads the
// operands of two k-steps before their MFMAs, and wave 0 loads its D entries
// before the k-loop, so the loads overlap.
template <typename R, bool transA, bool transB>
__global__ __launch_bounds__(256) void k_sector_gemm_mfma(const cx<R> *__restrict__ A, const cx<R> *__restrict__ B,
                                                          cx<R> *__restrict__ Cm, const MTile *__restrict__ tiles,
                                                          int64_t ntiles, const cx<R> *__restrict__ D, R dscale) {
    typedef MfmaOps<R> Ops;
    __shared__ R part[3][2][4][64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t tid = blockIdx.x;
    if (tid >= ntiles) return;  // the whole workgroup leaves together
    const MTile t = tiles[tid];
    if (t.m <= 0) return;       // padding tile
    const int m = t.m;
    const int64_t base = t.base;
    const int row0 = t.ti * kMfmaTile, col0 = t.tj * kMfmaTile;
    const int li = lane & 15, lk = lane >> 4;
    const int gi = row0 + li, gj = col0 + li;
    constexpr uint32_t E = sizeof(cx<R>);
    cx<R> *const tag = nullptr;
    const __amdgpu_buffer_rsrc_t ra = block_rsrc<R>(A + base, (int64_t)m * m);
    const __amdgpu_buffer_rsrc_t rb = block_rsrc<R>(B + base, (int64_t)m * m);
    cx<R> dd[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) dd[r] = cx<R>{0, 0};
    if (D && w == 0) {
        const __amdgpu_buffer_rsrc_t rd = block_rsrc<R>(D + base, (int64_t)m * m);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = row0 + Ops::row(lane, r);
            dd[r] = buf_load(rd, (i < m && gj < m) ? (uint32_t)(i + gj * m) * E : kBufOut, tag);
        }
    }
    constexpr int kUnroll = 4, kStep = 4 * kUnroll;
    const int nsteps = (m + kStep - 1) / kStep;
    const int s0 = nsteps * w / 4, s1 = nsteps * (w + 1) / 4;
    // Three real products per complex one (Gauss): t1 = Re a Re b, t2 = Im a Im b,
    // t3 = (Re a + Im a)(Re b + Im b); Re = t1 - t2, Im = t3 - t1 - t2.
    typename Ops::acc_t t1 = {0, 0, 0, 0}, t2 = {0, 0, 0, 0}, t3 = {0, 0, 0, 0};
    // G k-steps (4 G MFMA k-slices): all operand loads first, then the MFMAs.
    auto group = [&](int k0, auto G) {
        constexpr int NU = decltype(G)::value * kUnroll;
        cx<R> a[NU], b[NU];
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int gk = k0 + 4 * u + lk;
            const uint32_t ia = (uint32_t)(transA ? gk + gi * m : gi + gk * m) * E;
            const uint32_t ib = (uint32_t)(transB ? gj + gk * m : gk + gj * m) * E;
            a[u] = buf_load(ra, (gi < m && gk < m) ? ia : kBufOut, tag);
            b[u] = buf_load(rb, (gk < m && gj < m) ? ib : kBufOut, tag);
            if (transA) a[u].y = -a[u].y;
            if (transB) b[u].y = -b[u].y;
        }
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            t1 = Ops::mac(a[u].x, b[u].x, t1);
            t2 = Ops::mac(a[u].y, b[u].y, t2);
            t3 = Ops::mac(a[u].x + a[u].y, b[u].x + b[u].y, t3);
        }
    };
    int st = s0;
    for (; st + 1 < s1; st += 2) group(st * kStep, std::integral_constant<int, 2>());
    if (st < s1) group(st * kStep, std::integral_constant<int, 1>());
    typename Ops::acc_t cr, ci;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        cr[r] = t1[r] - t2[r];
        ci[r] = t3[r] - t1[r] - t2[r];
    }
    if (w > 0)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            part[w - 1][0][r][lane] = cr[r];
            part[w - 1][1][r][lane] = ci[r];
        }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cr[r] += part[q][0][r][lane];
            ci[r] += part[q][1][r][lane];
        }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = row0 + Ops::row(lane, r);
        if (i >= m || gj >= m) continue;
        const int64_t o = base + i + (int64_t)gj * m;
        cx<R> v = {cr[r], ci[r]};
        if (D) v = cmul(v, cx<R>{dd[r].x * dscale, dd[r].y * dscale});
        Cm[o] = v;
    }
}

}  // namespace pimi
