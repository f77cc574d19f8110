// Device kernels for the PI Liouvillian apply and the Krylov BLAS-1/2 steps.
//
// All reductions are fixed-shape trees (wave shuffles + LDS), so results are
// deterministic from run to run.
#pragma once

#include <type_traits>

#include "pimi_base.hpp"
#include "pimi_dia.hpp"

namespace pimi {

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

// the banded apply through the stencil's tables (programs compiled for a
// plan run the same body, see the plan's compiled banded module)
template <typename R, int TB>
__global__ __launch_bounds__(kApplyWave *kDiaSplit) void k_apply_dia_split(
    const cx<R> *__restrict__ x, cx<R> *__restrict__ y, int64_t n, const ApplyTile *__restrict__ tiles,
    int64_t ntiles, int64_t blocks_per_xcd, const int64_t *__restrict__ offs, const int32_t *__restrict__ dims,
    DiaTables T, const cx<R> *__restrict__ coef, R out_scale, int64_t ncols) {
    dia_split_body<DiaInterpreted, R, TB>(x, y, n, tiles, ntiles, blocks_per_xcd, offs, dims, T, coef, out_scale,
                                          ncols);
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
__device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, double *) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, float *) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t elem_rsrc(const T *p, int64_t entries) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)(entries * (int64_t)sizeof(T)), 0x00020000);
}

// A sector GEMM operand is complex (cx<R>) or real (R: the preconditioner's
// eigenvector factor of a real Hamiltonian block).  A complex product takes
// three real MFMAs (Gauss), a complex-by-real one two.
template <typename T>
struct is_cx_t : std::false_type {};
template <typename R>
struct is_cx_t<cx<R>> : std::true_type {};
__device__ __forceinline__ double op_conj(double a) { return a; }
__device__ __forceinline__ float op_conj(float a) { return a; }
template <typename R>
__device__ __forceinline__ cx<R> op_conj(cx<R> a) { return cx<R>{a.x, -a.y}; }

// The same 16x16 tile with its operands through raw buffers over the sector
// block: an out-of-range row or k reads zero by the hardware range check (no
// per-lane branch, 32-bit offsets).
template <typename R, typename TA = cx<R>, typename TB = cx<R>>
__device__ __forceinline__ void gemm_tile_mfma_buf(const TA *A, const TB *B, cx<R> *__restrict__ Cm,
                                                   int64_t m, int64_t base, GemmTile t, int transA, int transB,
                                                   const cx<R> *__restrict__ D, R dscale, int lane) {
    typedef MfmaOps<R> Ops;
    const int row0 = t.ti * kMfmaTile, col0 = t.tj * kMfmaTile;
    const int mi = (int)m;
    const int li = lane & 15, lk = lane >> 4;
    const int gi = row0 + li, gj = col0 + li;
    constexpr uint32_t EA = sizeof(TA), EB = sizeof(TB);
    TA *const tagA = nullptr;
    TB *const tagB = nullptr;
    const __amdgpu_buffer_rsrc_t ra = elem_rsrc(A + base, m * m);
    const __amdgpu_buffer_rsrc_t rb = elem_rsrc(B + base, m * m);
    typename Ops::acc_t cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    constexpr int kUnroll = 4;
    for (int k0 = 0; k0 < mi; k0 += 4 * kUnroll) {
        TA a[kUnroll];
        TB b[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int gk = k0 + 4 * u + lk;
            const uint32_t ia = (uint32_t)(transA ? gk + gi * mi : gi + gk * mi) * EA;
            const uint32_t ib = (uint32_t)(transB ? gj + gk * mi : gk + gj * mi) * EB;
            a[u] = buf_load(ra, (gi < mi && gk < mi) ? ia : kBufOut, tagA);
            b[u] = buf_load(rb, (gk < mi && gj < mi) ? ib : kBufOut, tagB);
            if (transA) a[u] = op_conj(a[u]);
            if (transB) b[u] = op_conj(b[u]);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if constexpr (is_cx_t<TA>::value && is_cx_t<TB>::value) {
                cr = Ops::mac(a[u].x, b[u].x, cr);
                cr = Ops::mac(-a[u].y, b[u].y, cr);
                ci = Ops::mac(a[u].x, b[u].y, ci);
                ci = Ops::mac(a[u].y, b[u].x, ci);
            } else if constexpr (is_cx_t<TA>::value) {
                cr = Ops::mac(a[u].x, b[u], cr);
                ci = Ops::mac(a[u].y, b[u], ci);
            } else {
                cr = Ops::mac(a[u], b[u].x, cr);
                ci = Ops::mac(a[u], b[u].y, ci);
            }
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
template <typename R, bool transA, bool transB, typename TA = cx<R>, typename TB = cx<R>>
__global__ __launch_bounds__(256) void k_sector_gemm_mfma(const TA *__restrict__ A, const TB *__restrict__ B,
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
    constexpr uint32_t E = sizeof(cx<R>), EA = sizeof(TA), EB = sizeof(TB);
    cx<R> *const tag = nullptr;
    TA *const tagA = nullptr;
    TB *const tagB = nullptr;
    const __amdgpu_buffer_rsrc_t ra = elem_rsrc(A + base, (int64_t)m * m);
    const __amdgpu_buffer_rsrc_t rb = elem_rsrc(B + base, (int64_t)m * m);
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
    // Complex operands: three real products per complex one (Gauss): t1 = Re a Re b,
    // t2 = Im a Im b, t3 = (Re a + Im a)(Re b + Im b); Re = t1 - t2, Im = t3 - t1 - t2.
    // One real operand: t1 and t2 are the real and imaginary parts.
    constexpr bool kGauss = is_cx_t<TA>::value && is_cx_t<TB>::value;
    typename Ops::acc_t t1 = {0, 0, 0, 0}, t2 = {0, 0, 0, 0}, t3 = {0, 0, 0, 0};
    // G k-steps (4 G MFMA k-slices): all operand loads first, then the MFMAs.
    auto group = [&](int k0, auto G) {
        constexpr int NU = decltype(G)::value * kUnroll;
        TA a[NU];
        TB b[NU];
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int gk = k0 + 4 * u + lk;
            const uint32_t ia = (uint32_t)(transA ? gk + gi * m : gi + gk * m) * EA;
            const uint32_t ib = (uint32_t)(transB ? gj + gk * m : gk + gj * m) * EB;
            a[u] = buf_load(ra, (gi < m && gk < m) ? ia : kBufOut, tagA);
            b[u] = buf_load(rb, (gk < m && gj < m) ? ib : kBufOut, tagB);
            if (transA) a[u] = op_conj(a[u]);
            if (transB) b[u] = op_conj(b[u]);
        }
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            if constexpr (kGauss) {
                t1 = Ops::mac(a[u].x, b[u].x, t1);
                t2 = Ops::mac(a[u].y, b[u].y, t2);
                t3 = Ops::mac(a[u].x + a[u].y, b[u].x + b[u].y, t3);
            } else if constexpr (is_cx_t<TA>::value) {
                t1 = Ops::mac(a[u].x, b[u], t1);
                t2 = Ops::mac(a[u].y, b[u], t2);
            } else {
                t1 = Ops::mac(a[u], b[u].x, t1);
                t2 = Ops::mac(a[u], b[u].y, t2);
            }
        }
    };
    int st = s0;
    for (; st + 1 < s1; st += 2) group(st * kStep, std::integral_constant<int, 2>());
    if (st < s1) group(st * kStep, std::integral_constant<int, 1>());
    typename Ops::acc_t cr, ci;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if constexpr (kGauss) {
            cr[r] = t1[r] - t2[r];
            ci[r] = t3[r] - t1[r] - t2[r];
        } else {
            cr[r] = t1[r];
            ci[r] = t2[r];
        }
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
