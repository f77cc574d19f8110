// Banded (diagonal) form of the stencil apply: tables, per-entry
// arithmetic and the tile loop.  Self-contained (pimi_base.hpp only), so the
// same code also compiles at run time, for a plan's compiled programs.
#pragma once

#include "pimi_base.hpp"

namespace pimi {

// When every factor of every term is a sum of a few diagonals,
//   A = sum_dA diag-shift(cA_dA),  A[a, a + dA] = cA_dA[a],
// a term s A X_in B^H adds to Y[a, b] the products
//   s cA_dA[a] conj(cB_dB[b]) X_in[a + dA, b + dB].
// The qubit sectors are of this form (H, Q and J- tridiagonal or narrower,
// every one-body contraction one diagonal).  Entries are grouped into chains
// of one input sector: a chain's lanes read one strip X_in[a + dA, b0 + c0 + k],
// k = 0 .. TB + span - 1, with plain coalesced loads (no column tables), and
// its entries follow in increasing row shift: an entry of row shift dA + sh
// finds the strip moved up sh rows, one wave-wide lane shift (DPP
// wave_shl:1, lane l takes lane l + 1) per row, so the three row shifts of
// a qubit input sector cost one strip load instead of three.  Lane l of a
// tile holds output row a0 + l; the last kDiaShift lanes only feed the
// shifts (a tile stores kDiaRows rows), so every stored row sees rows of the
// strip that were loaded, never the zero shifted in at lane 63.  Each entry
// combines the strip at its own column offset dB in 0..2 (relative to the
// chain's c0).  Coefficient vectors are zero wherever a + dA or b + dB
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
// Every output sums its chains in list order and each chain's entries in
// list order with explicit FMAs, so the result of an output does not depend
// on the tile width TB.
struct DiaChain {
    int64_t in_off;     // offset of the input block
    int32_t mn;         // input block size
    int32_t dA;         // row shift of the loaded strip
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
    int32_t sh;         // row shift beyond the chain's dA, 0..kDiaShift (non-decreasing along a chain)
    int32_t pad;
};
constexpr int kDiaShift = 2;                      // largest row shift within a chain
constexpr int kDiaRows = kApplyWave - kDiaShift;  // output rows of a banded tile
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

// One entry's coefficients for a tile: its row value(s) (per lane: complex,
// or the real one in .x, kind 7 its two reals in .x and .y) and its TB
// column values (wave-uniform: complex as re/im pairs, or the real vector,
// kind 7 its two real vectors one after the other).
template <typename R, int TB>
struct DiaFetch {
    cx<R> row;
    R col[2 * TB];
};

template <typename R, int TB, int KIND>
__device__ __forceinline__ void dia_fetch(DiaFetch<R, TB> &F, const DiaEntry &E, __amdgpu_buffer_rsrc_t pool,
                                          const PIMI_CONST cx<R> *coefs, int va, int vr, int b0) {
    auto cols_c = [&]() {
        const DiaCols<R, TB> w = *(const PIMI_CONST DiaCols<R, TB> *)(coefs + E.qb + b0);
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            F.col[2 * j] = w.v[j].x;
            F.col[2 * j + 1] = w.v[j].y;
        }
    };
    auto cols_r = [&](int q, int at) {
        const DiaColsR<R, TB> w = *(const PIMI_CONST DiaColsR<R, TB> *)((const PIMI_CONST R *)(coefs + q) + b0);
#pragma unroll
        for (int j = 0; j < TB; ++j) F.col[at + j] = w.v[j];
    };
    if constexpr (KIND == 0) {
        F.row = dia_row<R>(pool, va, E.qa);
    } else if constexpr (KIND == 1) {
        cols_c();
    } else if constexpr (KIND == 2 || KIND == 6) {
        F.row = dia_row<R>(pool, va, E.qa);
        cols_c();
    } else if constexpr (KIND == 3) {
        F.row.x = dia_row_real<R>(pool, vr, E.qa);
        cols_r(E.qb, 0);
    } else if constexpr (KIND == 4) {
        F.row.x = dia_row_real<R>(pool, vr, E.qa);
    } else if constexpr (KIND == 5) {
        cols_r(E.qb, 0);
    } else {
        F.row.x = dia_row_real<R>(pool, vr, E.qa);
        F.row.y = dia_row_real<R>(pool, vr, E.qa2);
        cols_r(E.qb, 0);
        cols_r(E.qb2, TB);
    }
}

template <typename R, int TB, int KIND, int DB>
__device__ __forceinline__ void dia_apply(cx<R> (&acc)[TB], const cx<R> (&st)[TB + 2], const DiaFetch<R, TB> &F) {
    auto wc = [&](int j) { return cx<R>{F.col[2 * j], F.col[2 * j + 1]}; };
    if constexpr (KIND == 4) {  // i c[a] X
        const R c = F.row.x;
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            acc[j].x = rfma(-c, st[j + DB].y, acc[j].x);
            acc[j].y = rfma(c, st[j + DB].x, acc[j].y);
        }
    } else if constexpr (KIND == 5) {  // i w[b] X
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            acc[j].x = rfma(-F.col[j], st[j + DB].y, acc[j].x);
            acc[j].y = rfma(F.col[j], st[j + DB].x, acc[j].y);
        }
    } else if constexpr (KIND == 6) {  // (cA[a] + w[b]) X
        const cx<R> ca = F.row;
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            const cx<R> w = wc(j);
            acc[j] = cfma(cx<R>{ca.x + w.x, ca.y + w.y}, st[j + DB], acc[j]);
        }
    } else if constexpr (KIND == 7) {  // (cA1[a] w1[b] + cA2[a] w2[b]) X, all real
        const R c1 = F.row.x, c2 = F.row.y;
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            const R c = rfma(c1, F.col[j], c2 * F.col[TB + j]);
            acc[j].x = rfma(c, st[j + DB].x, acc[j].x);
            acc[j].y = rfma(c, st[j + DB].y, acc[j].y);
        }
    } else if constexpr (KIND == 0) {
        const cx<R> ca = F.row;
#pragma unroll
        for (int j = 0; j < TB; ++j) acc[j] = cfma(ca, st[j + DB], acc[j]);
    } else if constexpr (KIND == 3) {
        const R ca = F.row.x;
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            const R c = ca * F.col[j];
            acc[j].x = rfma(c, st[j + DB].x, acc[j].x);
            acc[j].y = rfma(c, st[j + DB].y, acc[j].y);
        }
    } else if constexpr (KIND == 1) {
#pragma unroll
        for (int j = 0; j < TB; ++j) acc[j] = cfma(wc(j), st[j + DB], acc[j]);
    } else {
        const cx<R> ca = F.row;
#pragma unroll
        for (int j = 0; j < TB; ++j) acc[j] = cfma(cmulf(ca, wc(j)), st[j + DB], acc[j]);
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

// lane l takes lane l + 1's value (lane 63 takes 0): DPP wave_shl:1
__device__ __forceinline__ int lane_up(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
template <int W>
struct DiaWords {
    int v[W];
};
template <typename R>
__device__ __forceinline__ cx<R> lane_up(cx<R> v) {
    constexpr int W = (int)(sizeof(cx<R>) / sizeof(int));
    DiaWords<W> w = __builtin_bit_cast(DiaWords<W>, v);
#pragma unroll
    for (int k = 0; k < W; ++k) w.v[k] = lane_up(w.v[k]);
    return __builtin_bit_cast(cx<R>, w);
}

template <typename R, int TB>
__device__ __forceinline__ void dia_shift(cx<R> (&st)[TB + 2]) {
#pragma unroll
    for (int k = 0; k < TB + 2; ++k) st[k] = lane_up(st[k]);
}

// The banded tables of one stencil, as the kernels take them.  pid: per
// output sector, the compiled program that runs it (-1: the tables are
// interpreted); unused by the interpreting kernels.
struct DiaTables {
    const int32_t *cbase;   // ns + 1: chains of each output sector
    const DiaChain *chains;
    const DiaEntry *ents;
    const int32_t *pid;
};

// What a tile's program works with: the lane's rows, the tile's first
// column, the input vector and the coefficient pool.
template <typename R, int TB>
struct DiaLane {
    const cx<R> *x;
    __amdgpu_buffer_rsrc_t pool;
    const PIMI_CONST cx<R> *coefs;
    int arow;    // the lane's output row a0 + l (may lie past the sector)
    int va, vr;  // byte offsets of the lane's coefficient row (clamped into the sector)
    int b0;
    // the strip row of lane l is row a0 + l + dA, clamped into the block,
    // also past the sector's last row: after k shifts lane l holds lane
    // l + k's row, the clamped row a + dA + k
    template <int NK>
    __device__ __forceinline__ void strip(cx<R> (&st)[TB + 2], const DiaChain &G) const {
        const int r = min(max(arow + G.dA, 0), G.mn - 1);
        dia_strip<R, TB, NK>(st, x + G.in_off, r, G.mn, b0 + G.c0);
    }
    __device__ __forceinline__ void strip_any(cx<R> (&st)[TB + 2], const DiaChain &G) const {
        if (G.span == 0) strip<TB>(st, G);
        else strip<TB + 2>(st, G);
    }
    // fetch and apply one entry of fixed kind and strip column
    template <int KIND, int DB>
    __device__ __forceinline__ void entry(cx<R> (&acc)[TB], const cx<R> (&st)[TB + 2], const DiaEntry &E) const {
        DiaFetch<R, TB> F;
        dia_fetch<R, TB, KIND>(F, E, pool, coefs, va, vr, b0);
        dia_apply<R, TB, KIND, DB>(acc, st, F);
    }
    // the same for an entry whose kind and column are known at run time only
    __device__ __forceinline__ void entry_any(cx<R> (&acc)[TB], const cx<R> (&st)[TB + 2], const DiaEntry &E) const {
        switch (E.kind * 4 + E.dB) {
            case 0: entry<0, 0>(acc, st, E); break;
            case 1: entry<0, 1>(acc, st, E); break;
            case 2: entry<0, 2>(acc, st, E); break;
            case 4: entry<1, 0>(acc, st, E); break;
            case 5: entry<1, 1>(acc, st, E); break;
            case 6: entry<1, 2>(acc, st, E); break;
            case 8: entry<2, 0>(acc, st, E); break;
            case 9: entry<2, 1>(acc, st, E); break;
            case 10: entry<2, 2>(acc, st, E); break;
            case 12: entry<3, 0>(acc, st, E); break;
            case 13: entry<3, 1>(acc, st, E); break;
            case 14: entry<3, 2>(acc, st, E); break;
            case 16: entry<4, 0>(acc, st, E); break;
            case 17: entry<4, 1>(acc, st, E); break;
            case 18: entry<4, 2>(acc, st, E); break;
            case 20: entry<5, 0>(acc, st, E); break;
            case 21: entry<5, 1>(acc, st, E); break;
            case 22: entry<5, 2>(acc, st, E); break;
            case 24: entry<6, 0>(acc, st, E); break;
            case 25: entry<6, 1>(acc, st, E); break;
            case 26: entry<6, 2>(acc, st, E); break;
            case 28: entry<7, 0>(acc, st, E); break;
            case 29: entry<7, 1>(acc, st, E); break;
            case 30: entry<7, 2>(acc, st, E); break;
            default: break;
        }
    }
};

// The interpreting program: walks sector s's chains and entries from the
// tables (wave w of SPLIT takes chains w, w + SPLIT, ...).
struct DiaInterpreted {
    template <typename R, int TB, int SPLIT>
    static __device__ __forceinline__ void run(cx<R> (&acc)[TB], const DiaLane<R, TB> &L, const DiaTables &T, int s,
                                               int wave) {
        const PIMI_CONST int32_t *cbase = cptr(T.cbase);
        const PIMI_CONST DiaChain *chains = cptr(T.chains);
        const PIMI_CONST DiaEntry *ents = cptr(T.ents);
        const int g1 = cbase[s + 1];
        for (int g = cbase[s] + wave; g < g1; g += SPLIT) {
            const DiaChain G = chains[g];
            cx<R> st[TB + 2];
            L.strip_any(st, G);
            int shifted = 0;
            for (int e = G.e0; e < G.e1; ++e) {
                const DiaEntry E = ents[e];
                for (; shifted < E.sh; ++shifted) dia_shift<R, TB>(st);
                L.entry_any(acc, st, E);
            }
        }
    }
};

// One banded tile: rows a0..a0+kDiaRows-1 (one per lane) of columns
// b0..b0+nb-1 (nb <= TB) of sector s's output block, run by program Prog.
// Everything but the lane's row is wave-uniform: descriptors and column
// coefficients are scalar loads, kinds and shift counts scalar branches (or
// fixed in a compiled program).
//
// SPLIT > 1: the tile's chains are dealt to the SPLIT waves of the workgroup
// (wave w takes chains w, w + SPLIT, ...), which shortens each wave's chain of
// dependent strip loads; the partial sums meet in LDS and wave 0 adds them in
// the order w = 0 .. SPLIT-1 (deterministic).  Used for the single-solve
// paths, where one vector has too few tiles to hide the load latency.
template <typename Prog, typename R, int TB, int SPLIT = 1>
__device__ __forceinline__ void dia_tile(const cx<R> *xc_, cx<R> *yc_, R out_scale, ApplyTile tile,
                                         const int64_t *offs_, const int32_t *dims_, const DiaTables &T,
                                         const cx<R> *coef_) {
    PIMI_GLOBAL cx<R> *yc = gptr_w(yc_);
    const PIMI_CONST int64_t *offs = cptr(offs_);
    const PIMI_CONST int32_t *dims = cptr(dims_);
    const int s = __builtin_amdgcn_readfirstlane(tile.sector);
    const int b0 = __builtin_amdgcn_readfirstlane(tile.b0);
    const int a0 = __builtin_amdgcn_readfirstlane(tile.a0);
    const int nb = __builtin_amdgcn_readfirstlane(tile.nb);
    const int m = dims[s];
    const int lane = (int)(threadIdx.x & (kApplyWave - 1));
    const int arow = a0 + lane;
    const bool active = lane < kDiaRows && arow < m;
    const int a = min(arow, m - 1);  // the lane's coefficient row (lanes that store nothing read the last one)
    // the coefficient pool as one raw buffer (every read is in range; the
    // pool is far below 2 GiB), and the lane's byte offsets into a vector:
    // complex element a, and real element a of a packed real vector
    const DiaLane<R, TB> L{xc_, __builtin_amdgcn_make_buffer_rsrc((void *)coef_, (short)0, 0x7fffffff, 0x00020000),
                           cptr(coef_), arow, a * (int)sizeof(cx<R>), a * (int)sizeof(R), b0};
    cx<R> acc[TB];
#pragma unroll
    for (int j = 0; j < TB; ++j) acc[j] = cx<R>{0, 0};
    const int wave = SPLIT > 1 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
    Prog::template run<R, TB, SPLIT>(acc, L, T, s, wave);
    if constexpr (SPLIT > 1) {
        __shared__ cx<R> part[SPLIT - 1][TB][kApplyWave];
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
// tile per workgroup of kDiaSplit waves that share the tile's chains (a
// qubit sector has three: its own block and its two neighbours); tiles
// spread so that XCD k (workgroup id mod 8) handles one contiguous slab of
// the sector-major tile list.
constexpr int kDiaSplit = 3;
template <typename Prog, typename R, int TB>
__device__ __forceinline__ void dia_split_body(const cx<R> *x, cx<R> *y, int64_t n, const ApplyTile *tiles,
                                               int64_t ntiles, int64_t blocks_per_xcd, const int64_t *offs,
                                               const int32_t *dims, const DiaTables &T, const cx<R> *coef,
                                               R out_scale, int64_t ncols) {
    const int64_t bid = blockIdx.x;
    const int64_t t = (bid % kNumXcd) * blocks_per_xcd + bid / kNumXcd;
    if (t >= ntiles) return;  // the whole workgroup leaves together
    const ApplyTile tile = cptr(tiles)[t];
    for (int64_t c = blockIdx.y; c < ncols; c += gridDim.y)
        dia_tile<Prog, R, TB, kDiaSplit>(x + c * n, y + c * n, out_scale, tile, offs, dims, T, coef);
}

// The batched form: one descriptor per job, the jobs' tiles laid end to end.
struct BDia {
    int64_t first;
    const void *x;
    void *y;
    const int64_t *offs;
    const int32_t *dims;
    const ApplyTile *tiles;
    DiaTables tab;
    const void *coef;
    int64_t n;
    double out_scale;
};

// kDiaWaves tiles (waves) per workgroup, workgroups split into one
// contiguous slab per XCD; each wave finds its job.
template <typename Prog, typename R, int TB>
__device__ __forceinline__ void dia_batch_body(const BDia *J, int nj, int64_t total, int64_t per_xcd) {
    const int64_t bid = blockIdx.x;
    const int64_t blk = (bid % kNumXcd) * per_xcd + bid / kNumXcd;
    const int64_t g = blk * kDiaWaves + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (g >= total) return;
    const PIMI_CONST BDia *Jc = cptr(J);
    int lo = 0, hi = nj - 1;  // the wave's job: binary search on first (scalar)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (Jc[mid].first <= g) lo = mid;
        else hi = mid - 1;
    }
    const BDia A = Jc[lo];
    dia_tile<Prog, R, TB>((const cx<R> *)A.x, (cx<R> *)A.y, (R)A.out_scale, cptr(A.tiles)[g - A.first], A.offs,
                          A.dims, A.tab, (const cx<R> *)A.coef);
}

}  // namespace pimi
