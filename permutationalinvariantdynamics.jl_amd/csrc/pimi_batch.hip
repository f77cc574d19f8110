// Batched stationary solves: many independent PI models advanced together,
// one kernel launch per operation for all of them.
//
// A batch engine owns one HIP stream and one device arena.  Jobs (a compiled
// plan plus GMRES options) wait in a FIFO queue and are admitted while the
// arena has room for their state (the HBM footprint of their Krylov basis and
// work vectors) and a slot is free.  Admitted jobs advance in ticks: in one
// tick every job takes one step of its own state machine, and each kind of
// step is one launch over all jobs taking it (descriptor arrays give every
// job its own vectors, tables and sizes).  Each job solves the bordered
// stationary system (L/s + v t^H) x = v (src/krylov.jl:797-891) by
// right-preconditioned GMRES cycles inside iterative refinement:
//
//   r = v - A x (FP64)  ->  GMRES on A d = r in the engine precision, sector
//   preconditioner on the right, until |r_k| <= max(inner_rtol |r|,
//   (atol + rtol |x|) / 2) or the Krylov dimension is used up  ->  x += d
//
// With PIMI_C32 the Arnoldi basis, apply and preconditioner are complex FP32
// (dot products accumulate in FP64); residuals, the solution, Hessenberg
// least squares and the trace normalization stay FP64.  The stopping rule is
// the scale-invariant one (t^H L = 0 bounds |L x|/s by 2 |r|), so a finished
// job has |L x| / (s |x|) <= rtol for atol = 0.
//
// Every job's arithmetic depends only on its own data: reductions run over a
// per-job block count fixed by its dimension, so a job's result does not
// depend on which other jobs share its launches.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>

#include "pimi_internal.hpp"
#include "pimi_expv.hpp"

namespace pimi {

// ---------------------------------------------------------------------------
// launch descriptors: one entry per job taking part in a launch, `first` =
// the job's first block (tile, wave) in the launch, ascending
// ---------------------------------------------------------------------------
struct BApply {
    int64_t first;
    const void *x;
    void *y;
    const int64_t *offs;
    const int32_t *dims;
    const int32_t *tile_sector;  // 256-coordinate tiles (kb_sector_scale)
    const ApplyTile *tiles;      // apply tiles (kb_apply)
    const int32_t *entry_base;
    const ApplyEntry *entries;
    const int32_t *col;
    const void *val;
    int64_t n;
    double out_scale;
};

struct BTrace {  // one block per job
    const void *z;    // input
    void *y;          // bordered: y += (t . z) v
    cx<double> *dot;  // optional output: t . z
    const int64_t *tidx;
    const double *tval, *vval;
    int64_t nt;
    int32_t bordered, pad;
};

struct BGemm {
    int64_t first;
    const void *A, *B;
    void *C;
    const void *D;
    const int64_t *offs;
    const int32_t *dims;
    const GemmTile *tiles;
    double dscale;
    int32_t real_a, real_b;  // A (or B) is an array of R: the real factor U
};

struct BVec {
    int64_t first;
    const void *a, *b;
    void *c, *d;
    int64_t n;
    int32_t nblk, ncol;
    double s0;
    uint64_t u0;
};

struct BScatter {  // one block per job
    cx<double> *b, *x;
    const int64_t *tidx;
    const double *vval;
    int64_t nt;
};

struct BCopy {  // one block per entry
    const cx<double> *src;
    cx<double> *dst;
    int64_t count;
};

constexpr int kDotChunks = 16;  // reduction chunks per workgroup in kb_dots2_partial

template <typename T>
__device__ __forceinline__ int seg_of(const T *__restrict__ seg, int nseg, int64_t b) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid].first <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// batched kernels
// ---------------------------------------------------------------------------
// y = out_scale * L x for every job; apply tiles of all jobs laid end to end,
// one wave each, the global tile space split into one contiguous slab per XCD.
template <typename R>
__global__ __launch_bounds__(kApplyWave) void kb_apply(const BApply *__restrict__ J, int nj, int64_t total,
                                                       int64_t per_xcd) {
    const int64_t bid = blockIdx.x;
    const int64_t g = (bid % kNumXcd) * per_xcd + bid / kNumXcd;
    if (g >= total) return;
    const PIMI_CONST BApply *Jc = cptr(J);
    int lo = 0, hi = nj - 1;  // the wave's job: binary search on first (scalar)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (Jc[mid].first <= g) lo = mid;
        else hi = mid - 1;
    }
    const BApply A = Jc[lo];
    apply_tile<R, kApplyCols>((const cx<R> *)A.x, (cx<R> *)A.y, (R)A.out_scale, cptr(A.tiles)[g - A.first], A.offs,
                              A.dims, A.entry_base, A.entries, A.col, (const cx<R> *)A.val);
}

// The same through the banded form, interpreting the tables
template <typename R, int TB>
__global__ __launch_bounds__(kApplyWave *kDiaWaves) void kb_apply_dia(const BDia *__restrict__ J, int nj,
                                                                      int64_t total, int64_t per_xcd) {
    dia_batch_body<DiaInterpreted, R, TB>(J, nj, total, per_xcd);
}

// S: a stencil whose compiled programs every job's main stencil shares
// (null: interpret the tables)
template <typename R>
static void launch_kb_dia(hipStream_t st, const BDia *d, int nj, int64_t tot, int w, const DevStencil *S) {
    const int64_t nblk = (tot + kDiaWaves - 1) / kDiaWaves;
    const int64_t per_xcd = (nblk + kNumXcd - 1) / kNumXcd;
    const dim3 grid((unsigned)(per_xcd * kNumXcd)), block(kApplyWave * kDiaWaves);
    if (hipFunction_t f = S ? dia_program(*S, 1, sizeof(R) == 8 ? 0 : 1, w, st) : nullptr) {
        void *args[] = {(void *)&d, (void *)&nj, (void *)&tot, (void *)&per_xcd};
        HIPCHK(hipModuleLaunchKernel(f, grid.x, 1, 1, block.x, 1, 1, 0, st, args, nullptr));
        return;
    }
    switch (kDiaWidths[w]) {
        case 4: hipLaunchKernelGGL((kb_apply_dia<R, 4>), grid, block, 0, st, d, nj, tot, per_xcd); break;
        case 8: hipLaunchKernelGGL((kb_apply_dia<R, 8>), grid, block, 0, st, d, nj, tot, per_xcd); break;
        default: hipLaunchKernelGGL((kb_apply_dia<R, 16>), grid, block, 0, st, d, nj, tot, per_xcd); break;
    }
}

// a main stencil whose compiled programs those of every banded plan in
// `plans` equal (same structure, same device), or null
template <typename It>
static const DevStencil *common_programs(It first, It last) {
    const DevStencil *M = nullptr;
    for (It it = first; it != last; ++it) {
        const pimi_plan *P = *it;
        if (!P->main.dia_ok) continue;
        const DevStencil &S = P->main;
        if (S.prog_sig.empty()) return nullptr;
        if (M && (S.prog_sig != M->prog_sig || S.prog_device != M->prog_device)) return nullptr;
        if (!M) M = &S;
    }
    return M;
}

// banded-apply descriptor of one plan (coefficients in precision R, weighted or not)
template <typename R>
static BDia dia_desc(int64_t first, const pimi_plan *P, const void *x, void *y, bool weighted, double out_scale,
                     int w) {
    const DevStencil &S = P->main;
    const void *coef;
    if constexpr (sizeof(R) == 8) coef = weighted ? (const void *)S.coefw64.p : (const void *)S.coef64.p;
    else coef = weighted ? (const void *)S.coefw32.p : (const void *)S.coef32.p;
    return BDia{first, x, y, P->offs.p, P->dims.p, P->dtiles[w].p, S.tables(), coef, P->n,
                out_scale};
}

// y = x * w[sector] (w = 1/sigma: back from the weighted coordinates), one
// coordinate per lane over all jobs' tiles (FP64)
__global__ __launch_bounds__(kApplyBlock) void kb_sector_scale(const BApply *__restrict__ J, int nj, int64_t total) {
    const int64_t g = blockIdx.x;
    if (g >= total) return;
    const BApply A = J[seg_of(J, nj, g)];
    const int64_t lt = g - A.first;
    const int64_t i = lt * kApplyBlock + threadIdx.x;
    if (i >= A.n) return;
    int s = A.tile_sector[lt];
    while (i >= A.offs[s + 1]) ++s;
    const double f = ((const double *)A.val)[s];
    const cx<double> v = ((const cx<double> *)A.x)[i];
    ((cx<double> *)A.y)[i] = cx<double>{v.x * f, v.y * f};
}

// t . z per job (FP64 accumulation); bordered: y += (t . z) v on the trace support
template <typename R>
__global__ __launch_bounds__(256) void kb_trace(const BTrace *__restrict__ T) {
    __shared__ double sh[256 / 64];
    const BTrace J = T[blockIdx.x];
    const cx<R> *z = (const cx<R> *)J.z;
    double re = 0, im = 0;
    for (int64_t k = threadIdx.x; k < J.nt; k += 256) {
        const cx<R> zz = z[J.tidx[k]];
        re += J.tval[k] * (double)zz.x;
        im += J.tval[k] * (double)zz.y;
    }
    re = block_sum(re, sh);
    __syncthreads();
    im = block_sum(im, sh);
    if (J.dot && threadIdx.x == 0) *J.dot = cx<double>{re, im};
    if (J.bordered) {
        cx<R> *y = (cx<R> *)J.y;
        for (int64_t k = threadIdx.x; k < J.nt; k += 256) {
            const cx<R> yy = y[J.tidx[k]];
            y[J.tidx[k]] = cx<R>{(R)(yy.x + re * J.vval[k]), (R)(yy.y + im * J.vval[k])};
        }
    }
}

// grouped sector GEMM tiles of all jobs, one wave per 16x16 tile
template <typename R>
__global__ __launch_bounds__(256) void kb_gemm(const BGemm *__restrict__ G, int ng, int64_t total, int transA,
                                               int transB, int64_t per_xcd) {
    const int lane = threadIdx.x & 63;
    // the wave's tile and job are wave-uniform: scalar loads of the
    // descriptors (constant address space), operands through raw buffers.
    // Workgroup b runs on XCD b % 8, which takes one contiguous slab of the
    // tile space, so a sector's tiles share their operand panels in one L2.
    const int64_t g = (int64_t)(blockIdx.x % kNumXcd) * per_xcd + blockIdx.x / kNumXcd;
    const int64_t w = g * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (w >= total) return;  // whole wave leaves together; no block barriers below
    const PIMI_CONST BGemm *Gc = cptr(G);
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (Gc[mid].first <= w) lo = mid;
        else hi = mid - 1;
    }
    const BGemm J = Gc[lo];
    const GemmTile t = cptr(J.tiles)[w - J.first];
    const int64_t m = cptr(J.dims)[t.sector], base = cptr(J.offs)[t.sector];
    // a job whose factor U is real passes it as a real operand (wave-uniform branch)
    if (J.real_a)
        gemm_tile_mfma_buf<R, R, cx<R>>((const R *)J.A, (const cx<R> *)J.B, (cx<R> *)J.C, m, base, t, transA, transB,
                                        (const cx<R> *)J.D, (R)J.dscale, lane);
    else if (J.real_b)
        gemm_tile_mfma_buf<R, cx<R>, R>((const cx<R> *)J.A, (const R *)J.B, (cx<R> *)J.C, m, base, t, transA, transB,
                                        (const cx<R> *)J.D, (R)J.dscale, lane);
    else
        gemm_tile_mfma_buf<R>((const cx<R> *)J.A, (const cx<R> *)J.B, (cx<R> *)J.C, m, base, t, transA, transB,
                              (const cx<R> *)J.D, (R)J.dscale, lane);
}

// The BLAS-1 kernels below run over "virtual blocks": entry e of the
// descriptor array owns virtual blocks [first, first + its count), and a
// launch of at most kMaxGrid workgroups loops over all of them, so that the
// grid stays within the device's work-item limit for any batch.
constexpr int64_t kMaxGrid = 1 << 20;

// c[lb] = sum over block lb's elements of |a|^2
template <typename R>
__global__ __launch_bounds__(kRedBlock) void kb_norm_partial(const BVec *__restrict__ S, int ns, int64_t total) {
    __shared__ double sh[kRedBlock / 64];
    for (int64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const BVec J = S[seg_of(S, ns, vb)];
        const int64_t lb = vb - J.first;
        const cx<R> *a = (const cx<R> *)J.a;
        double s = 0;
        for (int64_t k = lb * kRedBlock + threadIdx.x; k < J.n; k += (int64_t)J.nblk * kRedBlock) {
            const double x = a[k].x, y = a[k].y;
            s += x * x + y * y;
        }
        s = block_sum(s, sh);
        if (threadIdx.x == 0) ((double *)J.c)[lb] = s;
        __syncthreads();
    }
}

// *d = sqrt(sum of the nblk partials in c); one wave per entry
__global__ __launch_bounds__(64) void kb_reduce_norm(const BVec *__restrict__ S) {
    const BVec J = S[blockIdx.x];
    double s = 0;
    for (int k = threadIdx.x; k < J.nblk; k += 64) s += ((const double *)J.c)[k];
    s = wave_sum(s);
    if (threadIdx.x == 0) *(double *)J.d = sqrt(s);
}

// mode 0: c -= V[:, 0:ncol] b ; mode 1: c = V[:, 0:ncol] b   (b: FP64 coefficients)
template <typename R>
__global__ __launch_bounds__(kRedBlock) void kb_gemv(const BVec *__restrict__ S, int ns, int64_t total, int mode) {
    __shared__ cx<double> hs[kColChunk];
    for (int64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const BVec J = S[seg_of(S, ns, vb)];
        const int64_t lb = vb - J.first;
        const cx<R> *V = (const cx<R> *)J.a;
        const cx<double> *h = (const cx<double> *)J.b;
        cx<R> *w = (cx<R> *)J.c;
        for (int64_t k0 = lb * kRedBlock; k0 < J.n; k0 += (int64_t)J.nblk * kRedBlock) {
            const int64_t k = k0 + threadIdx.x;
            double re = 0, im = 0;
            for (int c0 = 0; c0 < J.ncol; c0 += kColChunk) {
                const int cn = min(kColChunk, J.ncol - c0);
                __syncthreads();
                for (int c = threadIdx.x; c < cn; c += kRedBlock) hs[c] = h[c0 + c];
                __syncthreads();
                if (k < J.n)
                    for (int c = 0; c < cn; ++c) {
                        const cx<R> v = V[k + (int64_t)(c0 + c) * J.n];
                        re += (double)v.x * hs[c].x - (double)v.y * hs[c].y;
                        im += (double)v.x * hs[c].y + (double)v.y * hs[c].x;
                    }
            }
            if (k < J.n) {
                if (mode == 0) w[k] = cx<R>{(R)((double)w[k].x - re), (R)((double)w[k].y - im)};
                else w[k] = cx<R>{(R)re, (R)im};
            }
        }
        __syncthreads();
    }
}

// Delayed reorthogonalization (one reduction and one update pass per Arnoldi
// step).  With Q = V[:, 0:j] final, p = V[:, j] the once-orthogonalized
// candidate and w = B p:
//   s = Q^H p, pi = p^H p, t = Q^H w, tau = p^H w                 (one pass)
//   rho = sqrt(pi - |s|^2), e = (tau - s^H t) / rho^2
//   q_j = (p - Q s) / rho,  u_{j+1} = (w - e p - Q (t - e s)) / rho  (one pass)
// u_{j+1} is B q_j projected once against [Q, q_j]: B Q s lies in span[Q, q_j].

// The sum of v over the 64 lanes of a wave, in lane 63, by DPP moves (no
// LDS traffic): pairs, quads, half rows and rows by quad permutes and
// mirrors, then rows 0+1 and 2+3 by row broadcasts (the fixed tree of the
// gfx9 DPP reduction).  Lanes a step does not write add zero.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double v) {
    int2 w = __builtin_bit_cast(int2, v);
    w.x = __builtin_amdgcn_update_dpp(0, w.x, CTRL, ROW_MASK, 0xf, false);
    w.y = __builtin_amdgcn_update_dpp(0, w.y, CTRL, ROW_MASK, 0xf, false);
    return __builtin_bit_cast(double, w);
}
__device__ __forceinline__ double wave_total_lane63(double v) {
    v += dpp_f64<0xB1>(v);        // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);        // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);       // row_half_mirror
    v += dpp_f64<0x140>(v);       // row_mirror: every lane holds its row's sum
    v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
    v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3: lane 63 holds the total
    return v;
}

// block sums of N values with one barrier: each wave's total (lane 63) to
// LDS, then the waves' totals added in wave order; valid in every thread.
// `sh` holds N * blockDim.x / 64 doubles.
template <int N>
__device__ __forceinline__ void block_sums_dpp(double (&v)[N], double *sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = wave_total_lane63(v[q]);
    if (lane == 63)
#pragma unroll
        for (int q = 0; q < N; ++q) sh[q * nw + wid] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; ++q) {
        double r = 0;
        for (int k = 0; k < nw; ++k) r += sh[q * nw + k];
        v[q] = r;
    }
}

// partial[(col * nblk + chunk) * 2 + {0, 1}] = chunk's part of conj(V_col) . p
// and conj(V_col) . w for col = 0..ncol-1, p = V[:, ncol-1]; a = V, b = w.
// A workgroup takes J.u0 consecutive chunks and kDotCols consecutive columns:
// p and w are read once per element for all its columns.
constexpr int kDotCols = 4;
template <typename R>
__global__ __launch_bounds__(kRedBlock) void kb_dots2_partial(const BVec *__restrict__ S, int ns, int64_t total) {
    __shared__ double sh[4 * kDotCols * (kRedBlock / 64)];
    // workgroup b runs on XCD b % 8 (the grid is a multiple of 8); each XCD
    // takes a contiguous slab of every grid-stride pass, so the column groups
    // that share a chunk of p and w (consecutive vb) share one L2
    const int64_t slab = gridDim.x / kNumXcd;
    const int64_t b0 = (int64_t)(blockIdx.x % kNumXcd) * slab + blockIdx.x / kNumXcd;
    for (int64_t vb = b0; vb < total; vb += gridDim.x) {
        const BVec J = S[seg_of(S, ns, vb)];
        const int64_t lb = vb - J.first;
        const int64_t cpw = J.u0 ? (int64_t)J.u0 : 1;
        const int ngrp = (J.ncol + kDotCols - 1) / kDotCols;
        const int64_t grp = lb / ngrp;
        const int c0 = (int)(lb % ngrp) * kDotCols;
        const int nc = min(kDotCols, J.ncol - c0);
        const cx<R> *v = (const cx<R> *)J.a + (int64_t)c0 * J.n;
        const cx<R> *p = (const cx<R> *)J.a + (int64_t)(J.ncol - 1) * J.n;
        const cx<R> *w = (const cx<R> *)J.b;
        const int64_t ch1 = min((int64_t)J.nblk, (grp + 1) * cpw);
        for (int64_t chunk = grp * cpw; chunk < ch1; ++chunk) {
            double sums[4 * kDotCols];
#pragma unroll
            for (int q = 0; q < 4 * kDotCols; ++q) sums[q] = 0;
            for (int64_t k = chunk * kRedBlock + threadIdx.x; k < J.n; k += (int64_t)J.nblk * kRedBlock) {
                const cx<R> b = p[k], d = w[k];
                // columns past the job's last (a partial group) repeat the last
                // column's loads; their sums are not written
                cx<R> a[kDotCols];
#pragma unroll
                for (int u = 0; u < kDotCols; ++u) a[u] = v[k + (int64_t)min(u, nc - 1) * J.n];
#pragma unroll
                for (int u = 0; u < kDotCols; ++u) {
                    sums[4 * u + 0] += (double)a[u].x * (double)b.x + (double)a[u].y * (double)b.y;
                    sums[4 * u + 1] += (double)a[u].x * (double)b.y - (double)a[u].y * (double)b.x;
                    sums[4 * u + 2] += (double)a[u].x * (double)d.x + (double)a[u].y * (double)d.y;
                    sums[4 * u + 3] += (double)a[u].x * (double)d.y - (double)a[u].y * (double)d.x;
                }
            }
            block_sums_dpp(sums, sh);
            if (threadIdx.x < nc) {
                const int u = threadIdx.x;
                cx<double> *out = (cx<double> *)J.c + ((int64_t)(c0 + u) * J.nblk + chunk) * 2;
                out[0] = cx<double>{sums[4 * u + 0], sums[4 * u + 1]};
                out[1] = cx<double>{sums[4 * u + 2], sums[4 * u + 3]};
            }
            __syncthreads();
        }
    }
}

// c[col] = V_col^H p, c[ncol + col] = V_col^H w from the partials (a)
__global__ __launch_bounds__(64) void kb_reduce_cols2(const BVec *__restrict__ S, int ns, int64_t total) {
    for (int64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const BVec J = S[seg_of(S, ns, vb)];
        const int64_t c = vb - J.first;
        const cx<double> *pa = (const cx<double> *)J.a + c * J.nblk * 2;
        double r0 = 0, i0 = 0, r1 = 0, i1 = 0;
        for (int k = threadIdx.x; k < J.nblk; k += 64) {
            r0 += pa[2 * k].x;
            i0 += pa[2 * k].y;
            r1 += pa[2 * k + 1].x;
            i1 += pa[2 * k + 1].y;
        }
        r0 = wave_sum(r0);
        i0 = wave_sum(i0);
        r1 = wave_sum(r1);
        i1 = wave_sum(i1);
        if (threadIdx.x == 0) {
            cx<double> *o = (cx<double> *)J.c;
            o[c] = cx<double>{r0, i0};
            o[J.ncol + c] = cx<double>{r1, i1};
        }
    }
}

// Scalar slots (doubles) of the delayed step in a job's mailbox.
constexpr int kSlotRho = 10, kSlotInvRho = 11, kSlotE = 12;  // e: complex at 12, 13

// rho, 1/rho, e per job (one wave each) and the update coefficients
// coef[0:j] = s, coef[j:2j] = t - e s; a = reductions (V^H p, V^H w),
// c = coef, d = scalar slots, ncol = j + 1
// rho, 1/rho and e of a delayed step from its reductions sp = V^H p, tw = V^H w
// (ncol = j + 1 entries each); one wave, every lane gets them
__device__ __forceinline__ void dcgs_scalars(const cx<double> *sp, const cx<double> *tw, int j, double &rho,
                                             double &inv, cx<double> &e) {
    double ss = 0, str = 0, sti = 0;
    for (int c = (int)(threadIdx.x & 63); c < j; c += 64) {
        const cx<double> a = sp[c], b = tw[c];
        ss += a.x * a.x + a.y * a.y;
        str += a.x * b.x + a.y * b.y;
        sti += a.x * b.y - a.y * b.x;
    }
    ss = wave_sum(ss);
    str = wave_sum(str);
    sti = wave_sum(sti);
    const double r2 = sp[j].x - ss;
    rho = r2 > 0 ? sqrt(r2) : 0.0;
    inv = rho > 0 ? 1.0 / rho : 0.0;
    e = rho > 0 ? cx<double>{(tw[j].x - str) / r2, (tw[j].y - sti) / r2} : cx<double>{0, 0};
}

__global__ __launch_bounds__(64) void kb_dcgs_coef(const BVec *__restrict__ S) {
    const BVec J = S[blockIdx.x];
    const int j = J.ncol - 1;
    const cx<double> *sp = (const cx<double> *)J.a, *tw = sp + J.ncol;
    double rho, inv;
    cx<double> e;
    dcgs_scalars(sp, tw, j, rho, inv, e);
    cx<double> *coef = (cx<double> *)J.c;
    for (int c = threadIdx.x; c < j; c += 64) {
        const cx<double> a = sp[c], b = tw[c];
        coef[c] = a;
        coef[j + c] = cx<double>{b.x - (e.x * a.x - e.y * a.y), b.y - (e.x * a.y + e.y * a.x)};
    }
    if (threadIdx.x == 0) {
        double *slot = (double *)J.d;
        slot[kSlotRho] = rho;
        slot[kSlotInvRho] = inv;
        slot[kSlotE] = e.x;
        slot[kSlotE + 1] = e.y;
    }
}

// V[:, j] <- q_j and V[:, j+1] <- u_{j+1} (a = V, b = coef, c = w, d = slots, ncol = j + 1);
// kUpdCols basis columns' loads in flight per thread (measured on the 512-model
// FP32 batch: 8 -> 16 columns took the update pass from 3.9 to 4.3 TB/s; 24 and
// 32 were slower)
constexpr int kUpdCols = 16;
template <typename R>
__global__ __launch_bounds__(kRedBlock) void kb_dcgs_update(const BVec *__restrict__ S, int ns, int64_t total) {
    __shared__ cx<double> ca[kColChunk], cb[kColChunk];
    for (int64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const BVec J = S[seg_of(S, ns, vb)];
        const int64_t lb = vb - J.first;
        const int j = J.ncol - 1;
        cx<R> *V = (cx<R> *)J.a;
        const cx<double> *coef = (const cx<double> *)J.b;
        const cx<R> *w = (const cx<R> *)J.c;
        const double *slot = (const double *)J.d;
        const double inv = slot[kSlotInvRho], er = slot[kSlotE], ei = slot[kSlotE + 1];
        cx<R> *pcol = V + (int64_t)j * J.n, *ucol = V + (int64_t)(j + 1) * J.n;
        // one coordinate per thread (ceil(n / kRedBlock) workgroups per job);
        // kUpdCols basis columns' loads in flight, accumulated in column order
        const int64_t k = lb * kRedBlock + threadIdx.x;
        double ar = 0, ai = 0, br = 0, bi = 0;
        auto acc = [&](const cx<R> &v, int c) {
            ar += (double)v.x * ca[c].x - (double)v.y * ca[c].y;
            ai += (double)v.x * ca[c].y + (double)v.y * ca[c].x;
            br += (double)v.x * cb[c].x - (double)v.y * cb[c].y;
            bi += (double)v.x * cb[c].y + (double)v.y * cb[c].x;
        };
        for (int c0 = 0; c0 < j; c0 += kColChunk) {
            const int cn = min(kColChunk, j - c0);
            __syncthreads();
            for (int c = threadIdx.x; c < cn; c += kRedBlock) {
                ca[c] = coef[c0 + c];
                cb[c] = coef[j + c0 + c];
            }
            __syncthreads();
            if (k < J.n) {
                const cx<R> *vk = V + k + (int64_t)c0 * J.n;
                int c = 0;
                for (; c + kUpdCols <= cn; c += kUpdCols) {
                    cx<R> v[kUpdCols];
#pragma unroll
                    for (int u = 0; u < kUpdCols; ++u) v[u] = vk[(int64_t)(c + u) * J.n];
#pragma unroll
                    for (int u = 0; u < kUpdCols; ++u) acc(v[u], c + u);
                }
                for (; c < cn; ++c) acc(vk[(int64_t)c * J.n], c);
            }
        }
        if (k < J.n) {
            const double px = pcol[k].x, py = pcol[k].y, wx = w[k].x, wy = w[k].y;
            pcol[k] = cx<R>{(R)((px - ar) * inv), (R)((py - ai) * inv)};
            ucol[k] = cx<R>{(R)((wx - (er * px - ei * py) - br) * inv),
                            (R)((wy - (er * py + ei * px) - bi) * inv)};
        }
        __syncthreads();
    }
}

// kb_dcgs_coef and kb_dcgs_update of one job in one launch.  Every
// workgroup's first wave forms rho, 1/rho and e from the reductions (the
// coefficient kernel's arithmetic), the update coefficients s and t - e s are
// formed per column chunk in LDS, and workgroup 0 writes the scalar slots and,
// when the coefficient descriptor's b is set, mirrors the reductions and the
// scalars into that pinned host memory (scalars from entry u0).  S[0] is the
// step's coefficient descriptor, S[1] its update descriptor.  Bitwise the
// results of the two launches.
template <typename R>
__global__ __launch_bounds__(kRedBlock) void kb_dcgs_coef_update(const BVec *__restrict__ S, int64_t total) {
    __shared__ cx<double> ca[kColChunk], cb[kColChunk];
    __shared__ double scal[3];
    const BVec C = S[0], J = S[1];
    const int j = C.ncol - 1;
    const cx<double> *sp = (const cx<double> *)C.a, *tw = sp + C.ncol;
    if (threadIdx.x < 64) {
        double rho, inv;
        cx<double> e;
        dcgs_scalars(sp, tw, j, rho, inv, e);
        if (threadIdx.x == 0) {
            scal[0] = inv;
            scal[1] = e.x;
            scal[2] = e.y;
        }
        if (blockIdx.x == 0) {
            if (threadIdx.x == 0) {
                double *slot = (double *)C.d;
                slot[kSlotRho] = rho;
                slot[kSlotInvRho] = inv;
                slot[kSlotE] = e.x;
                slot[kSlotE + 1] = e.y;
            }
            if (C.b) {
                cx<double> *hb = (cx<double> *)const_cast<void *>(C.b);
                for (int c = threadIdx.x; c < 2 * C.ncol; c += 64) hb[c] = sp[c];
                if (threadIdx.x == 0) {
                    double *hs = (double *)(hb + C.u0);
                    hs[kSlotRho] = rho;
                    hs[kSlotInvRho] = inv;
                    hs[kSlotE] = e.x;
                    hs[kSlotE + 1] = e.y;
                }
            }
        }
    }
    __syncthreads();
    const double inv = scal[0], er = scal[1], ei = scal[2];
    cx<R> *V = (cx<R> *)J.a;
    const cx<R> *w = (const cx<R> *)J.c;
    cx<R> *pcol = V + (int64_t)j * J.n, *ucol = V + (int64_t)(j + 1) * J.n;
    for (int64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const int64_t k = vb * kRedBlock + threadIdx.x;
        double ar = 0, ai = 0, br = 0, bi = 0;
        auto acc = [&](const cx<R> &v, int c) {
            ar += (double)v.x * ca[c].x - (double)v.y * ca[c].y;
            ai += (double)v.x * ca[c].y + (double)v.y * ca[c].x;
            br += (double)v.x * cb[c].x - (double)v.y * cb[c].y;
            bi += (double)v.x * cb[c].y + (double)v.y * cb[c].x;
        };
        for (int c0 = 0; c0 < j; c0 += kColChunk) {
            const int cn = min(kColChunk, j - c0);
            __syncthreads();
            for (int c = threadIdx.x; c < cn; c += kRedBlock) {
                const cx<double> a = sp[c0 + c], b = tw[c0 + c];
                ca[c] = a;
                cb[c] = cx<double>{b.x - (er * a.x - ei * a.y), b.y - (er * a.y + ei * a.x)};
            }
            __syncthreads();
            if (k < J.n) {
                const cx<R> *vk = V + k + (int64_t)c0 * J.n;
                int c = 0;
                for (; c + kUpdCols <= cn; c += kUpdCols) {
                    cx<R> v[kUpdCols];
#pragma unroll
                    for (int u = 0; u < kUpdCols; ++u) v[u] = vk[(int64_t)(c + u) * J.n];
#pragma unroll
                    for (int u = 0; u < kUpdCols; ++u) acc(v[u], c + u);
                }
                for (; c < cn; ++c) acc(vk[(int64_t)c * J.n], c);
            }
        }
        if (k < J.n) {
            const double px = pcol[k].x, py = pcol[k].y, wx = w[k].x, wy = w[k].y;
            pcol[k] = cx<R>{(R)((px - ar) * inv), (R)((py - ai) * inv)};
            ucol[k] = cx<R>{(R)((wx - (er * px - ei * py) - br) * inv),
                            (R)((wy - (er * py + ei * px) - bi) * inv)};
        }
        __syncthreads();
    }
}

// c = a / *d (a zero denominator, an exact breakdown, gives 0), converting precision
template <typename Rd, typename Rs>
__global__ __launch_bounds__(kRedBlock) void kb_scale_div(const BVec *__restrict__ S, int ns, int64_t total) {
    for (int64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const BVec J = S[seg_of(S, ns, vb)];
        const int64_t lb = vb - J.first;
        const double den = *(const double *)J.d;
        const cx<Rs> *a = (const cx<Rs> *)J.a;
        cx<Rd> *c = (cx<Rd> *)J.c;
        for (int64_t k = lb * kRedBlock + threadIdx.x; k < J.n; k += (int64_t)J.nblk * kRedBlock)
            c[k] = den == 0.0 ? cx<Rd>{0, 0} : cx<Rd>{(Rd)((double)a[k].x / den), (Rd)((double)a[k].y / den)};
    }
}

// c = a - b (FP64)
__global__ __launch_bounds__(kRedBlock) void kb_sub(const BVec *__restrict__ S, int ns, int64_t total) {
    for (int64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const BVec J = S[seg_of(S, ns, vb)];
        const int64_t lb = vb - J.first;
        const cx<double> *a = (const cx<double> *)J.a, *b = (const cx<double> *)J.b;
        cx<double> *c = (cx<double> *)J.c;
        for (int64_t k = lb * kRedBlock + threadIdx.x; k < J.n; k += (int64_t)J.nblk * kRedBlock)
            c[k] = cx<double>{a[k].x - b[k].x, a[k].y - b[k].y};
    }
}

// c += a, a in the engine precision, c FP64
template <typename Rs>
__global__ __launch_bounds__(kRedBlock) void kb_add(const BVec *__restrict__ S, int ns, int64_t total) {
    for (int64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const BVec J = S[seg_of(S, ns, vb)];
        const int64_t lb = vb - J.first;
        const cx<Rs> *a = (const cx<Rs> *)J.a;
        cx<double> *c = (cx<double> *)J.c;
        for (int64_t k = lb * kRedBlock + threadIdx.x; k < J.n; k += (int64_t)J.nblk * kRedBlock)
            c[k] = cx<double>{c[k].x + (double)a[k].x, c[k].y + (double)a[k].y};
    }
}

__global__ __launch_bounds__(kRedBlock) void kb_zero(const BVec *__restrict__ S, int ns, int64_t total) {
    for (int64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const BVec J = S[seg_of(S, ns, vb)];
        const int64_t lb = vb - J.first;
        cx<double> *c = (cx<double> *)J.c;
        for (int64_t k = lb * kRedBlock + threadIdx.x; k < J.n; k += (int64_t)J.nblk * kRedBlock)
            c[k] = cx<double>{0, 0};
    }
}

// Probe vector for the operator-scale estimate: element e of the vector
// starting at counter u0 takes outputs u0 + 2e + 1 and u0 + 2e + 2 of the
// splitmix64 stream seeded 0x51ca1e, through Box-Muller; splitmix64's k-th
// output depends only on k, so every element is formed independently.
__device__ __forceinline__ double probe_uniform(uint64_t k) {
    uint64_t z = 0x51ca1eULL + k * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(kRedBlock) void kb_probe(const BVec *__restrict__ S, int ns, int64_t total) {
    for (int64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const BVec J = S[seg_of(S, ns, vb)];
        const int64_t lb = vb - J.first;
        cx<double> *c = (cx<double> *)J.c;
        for (int64_t k = lb * kRedBlock + threadIdx.x; k < J.n; k += (int64_t)J.nblk * kRedBlock) {
            const double u = probe_uniform(J.u0 + 2 * (uint64_t)k + 1),
                         v = probe_uniform(J.u0 + 2 * (uint64_t)k + 2);
            const double r = sqrt(-2.0 * log(u)), a = 6.283185307179586 * v;
            c[k] = cx<double>{r * cos(a) * 0.7071067811865476, r * sin(a) * 0.7071067811865476};
        }
    }
}

// b = x = v on the trace support (b and x zeroed before)
__global__ __launch_bounds__(256) void kb_scatter(const BScatter *__restrict__ S) {
    const BScatter J = S[blockIdx.x];
    for (int64_t k = threadIdx.x; k < J.nt; k += 256) {
        J.b[J.tidx[k]] = cx<double>{J.vval[k], 0};
        J.x[J.tidx[k]] = cx<double>{J.vval[k], 0};
    }
}

__global__ __launch_bounds__(256) void kb_gather(const BCopy *__restrict__ S) {
    const BCopy J = S[blockIdx.x];
    for (int64_t k = threadIdx.x; k < J.count; k += 256) J.dst[k] = J.src[k];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// First-fit sub-allocator over one device reservation (no hipMalloc/hipFree
// per job: those synchronize the device).
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    void init(size_t bytes) {
        HIPCHK(hipMalloc((void **)&base, bytes));
        size = bytes;
        free_.clear();
        free_[0] = bytes;
    }
    bool alloc(size_t bytes, size_t &off) {
        bytes = align_up(bytes, 256);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                off = it->first;
                const size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used += bytes;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(bytes, 256);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    ~Arena() {
        if (base) (void)hipFree(base);
    }
};

enum Phase { PH_PROBE = 0, PH_INIT, PH_ARNOLDI, PH_END, PH_FINAL };
// phases of a dynamics job (adaptive Krylov exponential, src/krylov_extensions.jl:1452-1568):
// observe (observables of x at a save time, |x|), Arnoldi ticks of one
// factorization, and projections (trial = beta V exp(h S) e1, |trial|)
enum DynPhase { PX_OBSERVE = 10, PX_ARNOLDI, PX_PROJECT };
enum JobKind { JK_STEADY = 0, JK_DYNAMICS = 1 };
enum JobState { JS_QUEUED = 0, JS_ACTIVE, JS_DONE, JS_FAILED };

// mailbox slots (complex doubles) at the head of each job's mail region:
// slots 0-7 hold 16 doubles (see the S_* indices)
constexpr int kMailTrace = 8;     // t . x
constexpr int kMailHacc = 9;      // Hessenberg column h[0..j]
enum {
    S_WNORM = 0,  // |w| of the Arnoldi step
    S_RNORM = 1,  // |r| (FP64 residual of the bordered system)
    S_XNORM = 2,  // |x|
    S_LXNORM = 3, // |L x|
    S_PNORM = 4,  // |p_k|, k = 0..2 (probe vectors)
    S_LPNORM = 7  // |L p_k|
};

struct BJob {
    int64_t id = 0;
    pimi_plan *plan = nullptr;
    pimi_gmres_opts o{};
    int state = JS_QUEUED, phase = PH_PROBE;
    int kdim = 0, nblk = 1;
    int64_t n = 0, nt = 0;
    size_t off = 0, bytes = 0;
    // device state (inside the arena)
    cx<double> *x = nullptr, *b = nullptr, *r = nullptr, *w64 = nullptr, *xo = nullptr;
    char *V = nullptr, *w = nullptr, *z = nullptr, *t1 = nullptr, *t2 = nullptr;  // engine precision
    cx<double> *partial = nullptr, *h = nullptr, *mail = nullptr;
    double *npart = nullptr, *vval = nullptr;
    double *scal() const { return (double *)mail; }
    // host Krylov state
    std::vector<cd> H, Hraw, cprov, g, sn, y;  // H: rotated (least squares); Hraw: final columns
    std::vector<double> cs;
    std::vector<double> vv;
    int j = 0, kend = 0;
    bool begin_cycle = false, candidate = false, weighted = false;
    double table_bytes64 = 0, table_bytes32 = 0, gemm_flops = 0;  // per apply / per preconditioner application
    double tight = 1.0;  // threshold factor, lowered when the coefficient-space check fails
    double scale = 0, beta = 0, beta0 = 0, xnorm = 0;
    int64_t iterations = 0, cycles = 0, applications = 0;
    std::chrono::steady_clock::time_point t_submit, t_admit;
    std::vector<cx<double>> result;
    pimi_solve_info info{};
    std::string error;
    // dynamics jobs (kind JK_DYNAMICS): save times, observables (nobs vectors
    // plus one zero column on the device), the state's samples, and the
    // exponential's factorization and step control
    int kind = JK_STEADY;
    pimi_expv_opts eo{};
    std::vector<double> ts;
    int64_t sample = 0;  // next save time to record
    int32_t nobs = 0;
    std::vector<cx<double>> rho0, obs_host, obs_out;
    cx<double> *obs = nullptr;
    Factorization F;
    ExpvControl X;
    DelayedArnoldi arn;
    bool have = false;
    double err = 0;
    std::vector<cx<double>> coeffs;
    int64_t accepted = 0, rejected = 0, factorizations = 0, attempts = 0;
    double error_sum = 0;
    // a factorization's Arnoldi steps run in one tick, step i reporting in
    // mail region i (dyn_mail_stride complex each)
    int arn_steps = 0;
    int64_t mail_stride = 0;
};

static int64_t dyn_mail_stride(int kdim, int nobs) {
    return kMailHacc + 2 * (int64_t)std::max(kdim + 3, nobs + 2);
}

// Workgroups of a job's BLAS-1 passes and reduction partials: one per 2048
// coordinates (so a lone large job still fills the chip), at most 1024; fixed
// by the job's dimension alone, so its reductions do not depend on the batch.
static int blocks_for(int64_t n) {
    return (int)std::min<int64_t>(std::max<int64_t>((n + 2047) / 2048, 1), 1024);
}

// Device bytes of one job's state: FP64 x, b, r, w64, xo; the Krylov basis and four
// work vectors in the engine precision; reduction partials, Hessenberg
// column, mailbox, coefficients and the bordered-system weights.
static size_t job_bytes(int64_t n, int64_t nt, int kdim, int prec) {
    const size_t es = prec == PIMI_C32 ? sizeof(cx<float>) : sizeof(cx<double>);
    const int nblk = blocks_for(n);
    size_t b = 0;
    b += 5 * align_up(sizeof(cx<double>) * n, 256);
    b += align_up(es * n * (size_t)(kdim + 2), 256) + 4 * align_up(es * n, 256);
    b += align_up(sizeof(cx<double>) * nblk * 2 * (size_t)(kdim + 1), 256);
    b += 2 * align_up(sizeof(cx<double>) * (size_t)(2 * (kdim + 2) + kMailHacc), 256);
    b += align_up(sizeof(double) * 4 * nblk, 256) + align_up(sizeof(double) * nt, 256);
    return b;
}

// A pipeline lane: its own stream, descriptor staging, mailbox and share of
// the active jobs.  The engine runs two lanes: while the device executes one
// lane's tick, the host reads back the other lane's results and stages its
// next tick, so host work overlaps device work (and the two lanes' kernels
// may overlap on the device).  A job stays in the lane that admitted it.
struct BLane {
    hipStream_t stream = nullptr;
    char *hstage = nullptr;
    size_t hstage_cap = 0;
    DevBuf<char> dstage;
    cx<double> *hmail = nullptr;
    size_t hmail_cap = 0;
    DevBuf<cx<double>> dmail;
    std::vector<hipEvent_t> events;
    std::vector<BJob *> active;
    // the tick in flight (launched, not yet read back)
    bool inflight = false;
    std::vector<size_t> moff;       // mailbox offset of each active job
    std::vector<int> op_class;      // per launch (timing)
    double bytes[8] = {}, flops[8] = {};
    // captured ticks: one executable graph per tick shape (the launch count
    // and each launch's operation class); a tick of a known shape updates
    // that graph's launch parameters in place (hipGraphExecUpdate) instead of
    // instantiating a new one
    std::map<std::string, hipGraphExec_t> graphs;
    int64_t graph_ticks = 0, graph_instantiations = 0, graph_updates = 0;
    void drop_graphs() {
        for (auto &kv : graphs)
            if (kv.second) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    ~BLane() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (auto ev : events) (void)hipEventDestroy(ev);
        if (hstage) (void)hipHostFree(hstage);
        if (hmail) (void)hipHostFree(hmail);
        if (stream) (void)hipStreamDestroy(stream);
    }
};
constexpr int kLanes = 2;

// ---------------------------------------------------------------------------
// The delayed-reorthogonalization step on one vector (the single-solve
// flexible GMRES): the engine's kernels with one descriptor each.
// ---------------------------------------------------------------------------
int dcgs_single_nblk(int64_t n) { return blocks_for(n); }

size_t dcgs_single_desc_bytes(int m) { return sizeof(BVec) * 4 * (size_t)m; }

void dcgs_single_descs(void *host, int m, void *V, const void *w, void *partial, void *mail, void *h, double *slot,
                       int64_t n, cx<double> *look, int64_t look_stride) {
    BVec *d = static_cast<BVec *>(host);
    const int nblk = blocks_for(n);
    for (int j = 0; j < m; ++j) {
        const int nc = j + 1;
        d[4 * j + 0] = BVec{0, V, w, partial, nullptr, n, nblk, nc, 0.0, 1};
        d[4 * j + 1] = BVec{0, partial, nullptr, mail, nullptr, n, nblk, nc, 0.0, 0};
        d[4 * j + 2] = BVec{0, mail, look + (size_t)(j & 1) * (size_t)look_stride, h, slot, n, nblk, nc, 0.0,
                            (uint64_t)(2 * ((int64_t)m + 1))};
        d[4 * j + 3] = BVec{0, V, h, const_cast<void *>(w), slot, n, nblk, nc, 0.0, 0};
    }
}

static_assert(kArnMailHead == kMailHacc && kArnSlotRho == kSlotRho && kArnSlotE == kSlotE,
              "the Arnoldi mail layout is the engine mailbox's");

void dcgs_arnoldi_descs(void *host, int steps, void *V, const void *w, void *partial, cx<double> *mail,
                        int64_t stride, void *h, int64_t n) {
    BVec *d = static_cast<BVec *>(host);
    const int nblk = blocks_for(n);
    for (int j = 0; j < steps; ++j) {
        const int nc = j + 1;
        cx<double> *mb = mail + (size_t)j * (size_t)stride;
        double *slot = reinterpret_cast<double *>(mb);
        d[4 * j + 0] = BVec{0, V, w, partial, nullptr, n, nblk, nc, 0.0, 1};
        d[4 * j + 1] = BVec{0, partial, nullptr, mb + kMailHacc, nullptr, n, nblk, nc, 0.0, 0};
        d[4 * j + 2] = BVec{0, mb + kMailHacc, nullptr, h, slot, n, nblk, nc, 0.0, 0};
        d[4 * j + 3] = BVec{0, V, h, const_cast<void *>(w), slot, n, nblk, nc, 0.0, 0};
    }
}

void dcgs_single_step(hipStream_t st, const void *desc_dev, int j, int64_t n) {
    const BVec *d = static_cast<const BVec *>(desc_dev) + 4 * (size_t)j;
    const int64_t nblk = blocks_for(n), nc = j + 1;
    auto grid8 = [](int64_t tot) {
        return (unsigned)std::min<int64_t>((tot + kNumXcd - 1) / kNumXcd * kNumXcd, kMaxGrid);
    };
    const int64_t tdots = nblk * ((nc + kDotCols - 1) / kDotCols);  // one chunk per workgroup (u0 = 1)
    hipLaunchKernelGGL(kb_dots2_partial<double>, dim3(grid8(tdots)), dim3(kRedBlock), 0, st, d + 0, 1, tdots);
    hipLaunchKernelGGL(kb_reduce_cols2, dim3(grid8(nc)), dim3(64), 0, st, d + 1, 1, nc);
    const int64_t tupd = (n + kRedBlock - 1) / kRedBlock;
    hipLaunchKernelGGL(kb_dcgs_coef_update<double>, dim3(grid8(tupd)), dim3(kRedBlock), 0, st, d + 2, tupd);
}

}  // namespace pimi

struct pimi_batch {
    pimi_context *ctx = nullptr;
    int prec = PIMI_C64;
    int max_jobs = 1024;
    hipStream_t stream = nullptr;  // engine-level work outside ticks (apply bench)
    pimi::Arena arena;
    int64_t next_id = 1;
    std::deque<pimi::BJob *> queue;
    pimi::BLane lanes[pimi::kLanes];
    std::map<int64_t, std::unique_ptr<pimi::BJob>> jobs;
    std::deque<int64_t> finished;
    std::atomic<int> busy{0};
    int64_t ticks = 0;
    int dia_w = 1;  // banded-apply tile width index (kDiaWidths) of the engine's launches
    // the autotuner's width per plan size (PI dimension, sector count); plans
    // of other sizes take dia_w
    std::map<std::pair<int64_t, int>, int> dia_w_size;
    int width_for(const pimi_plan *P) const {
        const auto it = dia_w_size.find({P->n, P->ns});
        return it == dia_w_size.end() ? dia_w : it->second;
    }
    bool trace = false;
    bool timing = false;  // PIMI_BATCH_TIMING: HIP events between launches, per operation class
    bool use_graphs = true;  // ticks captured as HIP graphs (PIMI_NO_GRAPH=1 or timing: eager launches)
    double stat_ms[8] = {}, stat_bytes[8] = {}, stat_flops[8] = {};
    size_t nactive() const {
        size_t k = 0;
        for (const auto &L : lanes) k += L.active.size();
        return k;
    }
    bool any_inflight() const {
        for (const auto &L : lanes)
            if (L.inflight) return true;
        return false;
    }
    ~pimi_batch() {
        for (auto &L : lanes)
            if (L.stream) (void)hipStreamSynchronize(L.stream);
        if (stream) (void)hipStreamSynchronize(stream);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace pimi {

// ---------------------------------------------------------------------------
// one tick: descriptors for every launch are staged on the host, uploaded in
// one copy, then the launches run; one gather + one copy bring back every
// job's scalars and Hessenberg column; one synchronization per tick.
// ---------------------------------------------------------------------------
// Operation classes of a tick, for the optional per-class timing (HIP events
// between launches) and the algorithmic bytes / flops they move.
enum OpClass { OC_GEMM = 0, OC_APPLY, OC_DOTS, OC_UPDATE, OC_CYCLE_END, OC_FP64, OC_OTHER, OC_COUNT };

struct Tick {
    pimi_batch *E = nullptr;
    hipStream_t stream = nullptr;
    std::vector<char> buf;
    std::vector<std::function<void(const char *)>> ops;
    std::vector<int> op_class;
    int cls = OC_OTHER;
    double bytes[OC_COUNT] = {}, flops[OC_COUNT] = {};
    void push(std::function<void(const char *)> f) {
        ops.push_back(std::move(f));
        op_class.push_back(cls);
    }
    // pointer fields (offset in buf) that must hold the device address of
    // stage data (offset in buf): patched once the device stage is placed
    std::vector<std::pair<size_t, size_t>> relocs;

    template <typename T>
    size_t put(const std::vector<T> &v) {
        const size_t off = align_up(buf.size(), 256);
        buf.resize(off + sizeof(T) * v.size());
        if (!v.empty()) std::memcpy(buf.data() + off, v.data(), sizeof(T) * v.size());
        return off;
    }

    // BLAS-1 launch over per-job entries; blocks(e) = blocks of entry e
    template <typename K>
    void vec(std::vector<BVec> v, K kernel, std::function<int64_t(const BVec &)> blocks, unsigned bdim = kRedBlock) {
        if (v.empty()) return;
        int64_t tot = 0;
        for (auto &e : v) {
            e.first = tot;
            tot += blocks(e);
        }
        const size_t off = put(v);
        const int ns = (int)v.size();
        // a multiple of the XCD count (kMaxGrid is one), for kernels that map
        // workgroups to XCD slabs; workgroups past tot exit at once
        const unsigned grid = (unsigned)std::min<int64_t>((tot + kNumXcd - 1) / kNumXcd * kNumXcd, kMaxGrid);
        hipStream_t st = stream;
        push([=](const char *d) {
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(bdim), 0, st, (const BVec *)(d + off), ns, tot);
        });
    }
    static int64_t nblk_of(const BVec &e) { return e.nblk; }

    // unit: y = L x (the coefficient-space operator); otherwise the job's
    // solve operator y = L x / s, in weighted coordinates for a weighted job
    template <typename R>
    void apply(const std::vector<BJob *> &js, std::function<const void *(BJob *)> in, std::function<void *(BJob *)> out,
               bool unit) {
        if (js.empty()) return;
        std::vector<BApply> v;
        // banded jobs grouped by their plan size's tile width: one launch per width
        std::vector<BDia> vd[kNumDiaWidths];
        std::vector<const pimi_plan *> vp[kNumDiaWidths];
        int64_t tot = 0, totd[kNumDiaWidths] = {};
        for (BJob *J : js) {
            pimi_plan *P = J->plan;
            if (P->main.dia_ok) {
                const int w = E->width_for(P);
                vd[w].push_back(dia_desc<R>(totd[w], P, in(J), out(J), !unit && J->weighted,
                                            unit ? 1.0 : 1.0 / J->scale, w));
                vp[w].push_back(P);
                totd[w] += P->ndtiles[w];
                bytes[cls] += 2.0 * J->n * sizeof(cx<R>) + (sizeof(R) == 8 ? J->table_bytes64 : J->table_bytes32);
                continue;
            }
            const void *val = sizeof(R) == 8 ? (const void *)P->main.val64.p : (const void *)P->main.val32.p;
            const ApplyEntry *entries = (!unit && J->weighted) ? P->main.entries_w.p : P->main.entries.p;
            v.push_back(BApply{tot, in(J), out(J), P->offs.p, P->dims.p, P->tile_sector.p, P->atiles4.p,
                               P->main.entry_base.p, entries, P->main.col.p, val, J->n,
                               unit ? 1.0 : 1.0 / J->scale});
            tot += P->natiles4;
            bytes[cls] += 2.0 * J->n * sizeof(cx<R>) + (sizeof(R) == 8 ? J->table_bytes64 : J->table_bytes32);
        }
        hipStream_t st = stream;
        for (int w = 0; w < kNumDiaWidths; ++w) {
            if (vd[w].empty()) continue;
            const size_t offd = put(vd[w]);
            const int njd = (int)vd[w].size();
            const int64_t td = totd[w];
            const DevStencil *M = common_programs(vp[w].begin(), vp[w].end());
            if (M) (void)dia_program(*M, 1, sizeof(R) == 8 ? 0 : 1, w, st);  // compiled now, not under capture
            push([=](const char *d) { launch_kb_dia<R>(st, (const BDia *)(d + offd), njd, td, w, M); });
        }
        if (v.empty()) return;
        const size_t off = put(v);
        const int nj = (int)v.size();
        const int64_t per_xcd = (tot + kNumXcd - 1) / kNumXcd;
        push([=](const char *d) {
            hipLaunchKernelGGL(kb_apply<R>, dim3((unsigned)(per_xcd * kNumXcd)), dim3(kApplyWave), 0, st,
                               (const BApply *)(d + off), nj, tot, per_xcd);
        });
    }

    template <typename R>
    void trace(const std::vector<BJob *> &js, std::function<const void *(BJob *)> z, std::function<void *(BJob *)> y,
               bool bordered, bool dot) {
        if (js.empty()) return;
        std::vector<BTrace> v;
        for (BJob *J : js)  // the solve operator of a weighted job has t -> ones (v already S v)
            v.push_back(BTrace{z(J), y ? y(J) : nullptr, dot ? J->mail + kMailTrace : nullptr, J->plan->d_tidx.p,
                               (bordered && J->weighted) ? J->plan->d_ones.p : J->plan->d_tval.p, J->vval, J->nt,
                               bordered ? 1 : 0, 0});
        const size_t off = put(v);
        const unsigned nj = (unsigned)v.size();
        hipStream_t st = stream;
        push([=](const char *d) {
            hipLaunchKernelGGL(kb_trace<R>, dim3(nj), dim3(256), 0, st, (const BTrace *)(d + off));
        });
    }

    // out = scale * U ((U^H in U) .* inv_den) U^H per sector, t1/t2 as temporaries
    template <typename R>
    void precond(const std::vector<BJob *> &js, std::function<const void *(BJob *)> in,
                 std::function<void *(BJob *)> out) {
        if (js.empty()) return;
        struct G {
            int ta, tb;
        };
        const G shapes[4] = {{1, 0}, {0, 0}, {0, 0}, {0, 1}};
        const int saved = cls;
        cls = OC_GEMM;
        for (BJob *J : js) flops[OC_GEMM] += J->gemm_flops;
        for (int stage = 0; stage < 4; ++stage) {
            std::vector<BGemm> v;
            int64_t tot = 0;
            for (BJob *J : js) {
                pimi_plan *P = J->plan;
                const bool ur = P->u_real;
                const void *U = sizeof(R) == 8 ? (ur ? (const void *)P->pUr64.p : (const void *)P->pU64.p)
                                               : (ur ? (const void *)P->pUr32.p : (const void *)P->pU32.p);
                const void *Inv = sizeof(R) == 8 ? (const void *)P->pInv64.p : (const void *)P->pInv32.p;
                BGemm e{tot, nullptr, nullptr, nullptr, nullptr, P->offs.p, P->dims.p, P->mtiles.p, 1.0, 0, 0};
                switch (stage) {
                    case 0: e.A = U; e.B = in(J); e.C = J->t1; e.real_a = ur; break;             // U^H in
                    case 1:                                                                      // (. U) .* inv
                        e.A = J->t1; e.B = U; e.C = J->t2; e.D = Inv; e.dscale = J->scale; e.real_b = ur;
                        break;
                    case 2: e.A = U; e.B = J->t2; e.C = J->t1; e.real_a = ur; break;            // U .
                    default: e.A = J->t1; e.B = U; e.C = out(J); e.real_b = ur; break;          // . U^H
                }
                v.push_back(e);
                tot += P->nmtiles;
            }
            const size_t off = put(v);
            const int ng = (int)v.size();
            const int ta = shapes[stage].ta, tb = shapes[stage].tb;
            hipStream_t st = stream;
            push([=](const char *d) {
                const int64_t nwg = (tot + 3) / 4, per_xcd = (nwg + kNumXcd - 1) / kNumXcd;
                hipLaunchKernelGGL(kb_gemm<R>, dim3((unsigned)(per_xcd * kNumXcd)), dim3(256), 0, st,
                                   (const BGemm *)(d + off), ng, tot, ta, tb, per_xcd);
            });
        }
        cls = saved;
    }

    template <typename R>
    void norm(const std::vector<BJob *> &js, std::function<const void *(BJob *)> a, int slot, int pslot = 0) {
        if (js.empty()) return;
        std::vector<BVec> v;
        for (BJob *J : js) {
            v.push_back(BVec{0, a(J), nullptr, J->npart + (size_t)pslot * J->nblk, J->scal() + slot, J->n, J->nblk, 1,
                             0.0, 0});
            bytes[cls] += (double)J->n * sizeof(cx<R>);
        }
        vec(v, kb_norm_partial<R>, nblk_of);
        const size_t off = put(v);
        const unsigned nj = (unsigned)v.size();
        hipStream_t st = stream;
        push([=](const char *d) {
            hipLaunchKernelGGL(kb_reduce_norm, dim3(nj), dim3(64), 0, st, (const BVec *)(d + off));
        });
    }

    // One delayed-reorthogonalization Arnoldi step (see kb_dcgs_update):
    // V[:, j] becomes q_j and V[:, j+1] the next candidate; the reductions
    // V^H p, V^H w land in the mailbox for the host's Hessenberg update.
    // Step j of job J is J->j (or step_of(J)); its reductions and scalars go to
    // the mail region mail_of(J) (default: the head of J's mailbox).
    template <typename R>
    void dcgs(const std::vector<BJob *> &js, std::function<int(BJob *)> step_of = nullptr,
              std::function<cx<double> *(BJob *)> mail_of = nullptr) {
        if (js.empty()) return;
        const int saved = cls;
        std::vector<BVec> d, r, c, u;
        for (BJob *J : js) {
            const int jj = step_of ? step_of(J) : J->j;
            cx<double> *mb = mail_of ? mail_of(J) : J->mail;
            const int nc = jj + 1;
            bytes[OC_DOTS] += (double)(nc + 1) * J->n * sizeof(cx<R>);
            bytes[OC_UPDATE] += (double)(jj + 4) * J->n * sizeof(cx<R>);
            d.push_back(BVec{0, J->V, J->w, J->partial, nullptr, J->n, J->nblk, nc, 0.0, 0});
            r.push_back(BVec{0, J->partial, nullptr, mb + kMailHacc, nullptr, J->n, J->nblk, nc, 0.0, 0});
            c.push_back(BVec{0, mb + kMailHacc, nullptr, J->h, (double *)mb, J->n, J->nblk, nc, 0.0, 0});
            u.push_back(BVec{0, J->V, J->h, J->w, (double *)mb, J->n, J->nblk, nc, 0.0, 0});
        }
        cls = OC_DOTS;
        // chunks per workgroup: kDotChunks when the launch has workgroups
        // enough to fill the chip, fewer when few jobs share it (each chunk's
        // partial is summed on its own, so results do not depend on this)
        int64_t cpw = kDotChunks;
        auto groups = [&](int64_t c) {
            int64_t g = 0;
            for (const BVec &e : d) g += ((e.nblk + c - 1) / c) * ((e.ncol + kDotCols - 1) / kDotCols);
            return g;
        };
        while (cpw > 1 && groups(cpw) < 2048) cpw /= 2;  // eight workgroups per CU
        for (auto &e : d) e.u0 = (uint64_t)cpw;
        vec(d, kb_dots2_partial<R>, [](const BVec &e) {
            return (int64_t)((e.nblk + (int64_t)e.u0 - 1) / (int64_t)e.u0) * ((e.ncol + kDotCols - 1) / kDotCols);
        });
        cls = OC_OTHER;
        vec(r, kb_reduce_cols2, [](const BVec &e) { return (int64_t)e.ncol; }, 64);
        {
            const size_t off = put(c);
            const unsigned nj = (unsigned)c.size();
            hipStream_t st = stream;
            push([=](const char *dd) {
                hipLaunchKernelGGL(kb_dcgs_coef, dim3(nj), dim3(64), 0, st, (const BVec *)(dd + off));
            });
        }
        cls = OC_UPDATE;
        vec(u, kb_dcgs_update<R>, [](const BVec &e) { return (e.n + kRedBlock - 1) / kRedBlock; });
        cls = saved;
    }

    template <typename R>
    size_t gemv(std::vector<BVec> u, int mode) {
        if (u.empty()) return 0;
        int64_t tot = 0;
        for (auto &e : u) {
            e.first = tot;
            tot += e.nblk;
        }
        const size_t off = put(u);
        const int ns = (int)u.size();
        const unsigned grid = (unsigned)std::min<int64_t>(tot, kMaxGrid);
        hipStream_t st = stream;
        push([=](const char *d) {
            hipLaunchKernelGGL(kb_gemv<R>, dim3(grid), dim3(kRedBlock), 0, st, (const BVec *)(d + off), ns, tot,
                               mode);
        });
        return off;
    }
};

// Device bytes of one dynamics job (FP64 engine): x and the trial vector, the
// basis with room for one enlargement step, w, reduction partials (for the
// basis or the observables, whichever has more columns), Hessenberg
// coefficients, mailbox, norm partials and the observables.
static int64_t dyn_mail_stride(int kdim, int nobs);
static size_t dyn_job_bytes(int64_t n, int kdim, int nobs) {
    const int nblk = blocks_for(n);
    const int cols = std::max(kdim + 2, nobs + 1);
    size_t b = 0;
    b += 2 * align_up(sizeof(cx<double>) * n, 256);
    b += align_up(sizeof(cx<double>) * n * (size_t)(kdim + 3), 256) + align_up(sizeof(cx<double>) * n, 256);
    b += align_up(sizeof(cx<double>) * nblk * 2 * (size_t)cols, 256);
    b += align_up(sizeof(cx<double>) * (size_t)(2 * (cols + 1) + kMailHacc), 256);
    b += align_up(sizeof(cx<double>) * (size_t)(dyn_mail_stride(kdim, nobs) * (kdim + 2)), 256);
    b += align_up(sizeof(double) * 4 * nblk, 256);
    b += align_up(sizeof(cx<double>) * n * (size_t)(nobs + 1), 256);
    return b;
}

template <typename R>
static char *vcol(BJob *J, int c) {
    return J->V + sizeof(cx<R>) * (size_t)J->n * (size_t)c;
}

static void carve(pimi_batch *E, BJob *J) {
    const size_t es = E->prec == PIMI_C32 ? sizeof(cx<float>) : sizeof(cx<double>);
    char *p = E->arena.base + J->off;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += align_up(bytes, 256);
        return q;
    };
    const int64_t n = J->n;
    if (J->kind == JK_DYNAMICS) {
        const int cols = std::max(J->kdim + 2, J->nobs + 1);
        J->x = (cx<double> *)take(sizeof(cx<double>) * n);
        J->r = (cx<double> *)take(sizeof(cx<double>) * n);
        J->V = take(sizeof(cx<double>) * n * (size_t)(J->kdim + 3));
        J->w = take(sizeof(cx<double>) * n);
        J->partial = (cx<double> *)take(sizeof(cx<double>) * J->nblk * 2 * (size_t)cols);
        J->h = (cx<double> *)take(sizeof(cx<double>) * (size_t)(2 * (cols + 1) + kMailHacc));
        J->mail_stride = dyn_mail_stride(J->kdim, J->nobs);
        J->mail = (cx<double> *)take(sizeof(cx<double>) * (size_t)(J->mail_stride * (J->kdim + 2)));
        J->npart = (double *)take(sizeof(double) * 4 * J->nblk);
        J->obs = (cx<double> *)take(sizeof(cx<double>) * n * (size_t)(J->nobs + 1));
        if ((size_t)(p - (E->arena.base + J->off)) > J->bytes)
            throw Error(PIMI_EINVAL, "job layout exceeds its footprint");
        return;
    }
    J->x = (cx<double> *)take(sizeof(cx<double>) * n);
    J->b = (cx<double> *)take(sizeof(cx<double>) * n);
    J->r = (cx<double> *)take(sizeof(cx<double>) * n);
    J->w64 = (cx<double> *)take(sizeof(cx<double>) * n);
    J->xo = (cx<double> *)take(sizeof(cx<double>) * n);
    J->V = take(es * n * (size_t)(J->kdim + 2));
    J->w = take(es * n);
    J->z = take(es * n);
    J->t1 = take(es * n);
    J->t2 = take(es * n);
    J->partial = (cx<double> *)take(sizeof(cx<double>) * J->nblk * 2 * (size_t)(J->kdim + 1));
    J->h = (cx<double> *)take(sizeof(cx<double>) * (size_t)(2 * (J->kdim + 2) + kMailHacc));
    J->mail = (cx<double> *)take(sizeof(cx<double>) * (size_t)(2 * (J->kdim + 2) + kMailHacc));
    J->npart = (double *)take(sizeof(double) * 4 * J->nblk);
    J->vval = (double *)take(sizeof(double) * J->nt);
    if ((size_t)(p - (E->arena.base + J->off)) > J->bytes) throw Error(PIMI_EINVAL, "job layout exceeds its footprint");
}

// Admit queued jobs into lane L (FIFO) while memory and slots allow, up to
// L's even share of all active and queued jobs (so the lanes stay balanced);
// called when L stages a tick, so a job's first work is ordered on its lane's
// stream.
static void admit(pimi_batch *E, BLane &L) {
    const size_t share = (E->nactive() + E->queue.size() + kLanes - 1) / kLanes;
    while (!E->queue.empty() && (int)E->nactive() < E->max_jobs && L.active.size() < share) {
        BJob *J = E->queue.front();
        size_t off;
        if (!E->arena.alloc(J->bytes, off)) break;  // FIFO: wait for memory to free up
        E->queue.pop_front();
        J->off = off;
        carve(E, J);
        if (J->kind == JK_DYNAMICS) {
            // the initial state and the observables (host copies kept until
            // the job ends: the copies are asynchronous on the lane's stream)
            HIPCHK(hipMemcpyAsync(J->x, J->rho0.data(), sizeof(cx<double>) * J->n, hipMemcpyHostToDevice, L.stream));
            if (J->nobs > 0)
                HIPCHK(hipMemcpyAsync(J->obs, J->obs_host.data(), sizeof(cx<double>) * J->n * J->nobs,
                                      hipMemcpyHostToDevice, L.stream));
            HIPCHK(hipMemsetAsync(J->obs + (size_t)J->n * J->nobs, 0, sizeof(cx<double>) * J->n, L.stream));
            J->state = JS_ACTIVE;
            J->phase = PX_OBSERVE;
            J->t_admit = std::chrono::steady_clock::now();
            L.active.push_back(J);
            continue;
        }
        HIPCHK(hipMemcpyAsync(J->vval, J->vv.data(), sizeof(double) * J->nt, hipMemcpyHostToDevice, L.stream));
        J->state = JS_ACTIVE;
        J->phase = J->o.operator_scale > 0 ? PH_INIT : PH_PROBE;
        J->scale = J->o.operator_scale;
        J->t_admit = std::chrono::steady_clock::now();
        L.active.push_back(J);
    }
}

static void finish(pimi_batch *E, BJob *J, bool ok) {
    J->state = ok ? JS_DONE : JS_FAILED;
    if (J->kind == JK_DYNAMICS) {
        if ((int64_t)J->result.size() != J->n) J->result.assign((size_t)J->n, cx<double>{0, 0});
        pimi_solve_info &I = J->info;
        std::memset(&I, 0, sizeof I);
        I.converged = ok ? 1 : 0;
        I.accepted = J->accepted;
        I.rejected = J->rejected;
        I.applications = J->applications;
        I.factorizations = J->factorizations;
        I.iterations = J->attempts;
        I.estimated_error = J->error_sum;
        I.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - J->t_submit).count();
        E->arena.release(J->off, J->bytes);
        J->x = J->r = nullptr;
        J->V = J->w = nullptr;
        J->obs = nullptr;
        E->finished.push_back(J->id);
        return;
    }
    if ((int64_t)J->result.size() != J->n) {  // failed before its final phase
        J->result.assign((size_t)J->n, cx<double>{0, 0});
        std::memset(&J->info, 0, sizeof J->info);
        J->info.iterations = J->iterations;
        J->info.applications = J->applications;
        J->info.operator_scale = J->scale;
    }
    E->arena.release(J->off, J->bytes);
    J->x = J->b = J->r = J->w64 = J->xo = nullptr;
    J->V = J->w = J->z = J->t1 = J->t2 = nullptr;
    E->finished.push_back(J->id);
}

// Stage and launch lane L's tick (no wait); complete_tick reads it back.
// The tick's launches as one HIP graph on the lane's stream.  The launches
// are captured (kernels only: the stage upload before and the mailbox and
// result copies after stay ordinary stream operations); the lane keeps one
// executable graph per tick shape and updates its parameters in place when a
// tick of that shape recurs -- the Arnoldi ticks of a stable job set -- and
// instantiates anew when the update is refused (a different kernel at some
// position).  Every kernel a tick uses is resolved before the capture (the
// compiled banded programs are compiled outside it).
static void launch_graph(BLane &L, Tick &T, const char *dbase) {
    std::string key = std::to_string(T.ops.size()) + ":";
    for (int c : T.op_class) key += (char)('a' + c);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
    try {
        for (auto &op : T.ops) op(dbase);
    } catch (...) {
        (void)hipStreamEndCapture(L.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIPCHK(hipStreamEndCapture(L.stream, &g));
    if (L.graphs.size() >= 64 && !L.graphs.count(key)) L.drop_graphs();
    hipGraphExec_t &ex = L.graphs[key];
    bool updated = false;
    if (ex) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult r = hipGraphExecUpdateError;
        updated = hipGraphExecUpdate(ex, g, &err_node, &r) == hipSuccess && r == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(ex);
            ex = nullptr;
        }
    }
    if (!updated) {
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            ex = nullptr;
            L.graphs.erase(key);
            (void)hipGraphDestroy(g);
            HIPCHK(e);
        }
        ++L.graph_instantiations;
    } else {
        ++L.graph_updates;
    }
    (void)hipGraphDestroy(g);
    HIPCHK(hipGraphLaunch(ex, L.stream));
    ++L.graph_ticks;
}

template <typename R>
static void launch_tick(pimi_batch *E, BLane &L) {
    admit(E, L);
    if (L.active.empty()) return;
    std::vector<BJob *> probe, init, arn, begin, end, fin;
    std::vector<BJob *> xobs, xbegin, xarn, xproj;  // dynamics jobs
    for (BJob *J : L.active) switch (J->phase) {
            case PH_PROBE: probe.push_back(J); break;
            case PH_INIT: init.push_back(J); break;
            case PH_ARNOLDI:
                arn.push_back(J);
                if (J->begin_cycle) begin.push_back(J);
                break;
            case PH_END: end.push_back(J); break;
            case PX_OBSERVE: xobs.push_back(J); break;
            case PX_ARNOLDI:
                xarn.push_back(J);
                if (J->j == 0) xbegin.push_back(J);
                break;
            case PX_PROJECT: xproj.push_back(J); break;
            default: fin.push_back(J); break;
        }
    Tick T;
    T.E = E;
    T.stream = L.stream;
    auto W64 = [](BJob *J) -> void * { return J->w64; };
    auto cX = [](BJob *J) -> const void * { return J->x; };
    auto cW64 = [](BJob *J) -> const void * { return J->w64; };
    auto cR = [](BJob *J) -> const void * { return J->r; };
    auto vecs = [](const std::vector<BJob *> &js, std::function<BVec(BJob *)> f) {
        std::vector<BVec> v;
        for (BJob *J : js) v.push_back(f(J));
        return v;
    };
    // -- operator scale: max over three probe vectors of |L p| / |p| (src/krylov.jl:156-169)
    T.cls = OC_FP64;
    for (int k = 0; k < 3 && !probe.empty(); ++k) {
        T.vec(vecs(probe, [&](BJob *J) {
                  return BVec{0, nullptr, nullptr, J->r, nullptr, J->n, J->nblk, 1, 0.0, (uint64_t)(2 * J->n * k)};
              }),
              kb_probe, Tick::nblk_of);
        T.norm<double>(probe, cR, S_PNORM + k);
        T.apply<double>(probe, cR, W64, true);
        T.norm<double>(probe, cW64, S_LPNORM + k, 1);
    }
    // -- initial residual of x0 = v
    if (!init.empty()) {
        std::vector<BVec> z;
        for (BJob *J : init) {
            z.push_back(BVec{0, nullptr, nullptr, J->b, nullptr, J->n, J->nblk, 1, 0.0, 0});
            z.push_back(BVec{0, nullptr, nullptr, J->x, nullptr, J->n, J->nblk, 1, 0.0, 0});
        }
        T.vec(z, kb_zero, Tick::nblk_of);
        std::vector<BScatter> sc;
        for (BJob *J : init) sc.push_back(BScatter{J->b, J->x, J->plan->d_tidx.p, J->vval, J->nt});
        const size_t off = T.put(sc);
        const unsigned nj = (unsigned)sc.size();
        hipStream_t st = L.stream;
        T.push([=](const char *d) {
            hipLaunchKernelGGL(kb_scatter, dim3(nj), dim3(256), 0, st, (const BScatter *)(d + off));
        });
    }
    // -- Arnoldi step (first the new cycle's start vector V0 = r / |r|)
    T.cls = OC_OTHER;
    T.vec(vecs(begin, [](BJob *J) {
              return BVec{0, J->r, nullptr, J->V, J->scal() + S_RNORM, J->n, J->nblk, 1, 0.0, 0};
          }),
          kb_scale_div<R, double>, Tick::nblk_of);
    if (!arn.empty()) {
        T.precond<R>(arn, [](BJob *J) -> const void * { return vcol<R>(J, J->j); },
                     [](BJob *J) -> void * { return J->z; });
        T.cls = OC_APPLY;
        T.apply<R>(arn, [](BJob *J) -> const void * { return J->z; }, [](BJob *J) -> void * { return J->w; }, false);
        T.cls = OC_OTHER;
        T.trace<R>(arn, [](BJob *J) -> const void * { return J->z; }, [](BJob *J) -> void * { return J->w; }, true,
                   false);
        T.dcgs<R>(arn);
    }
    // -- end of a cycle: x += M^{-1} V y, then the FP64 residual
    if (!end.empty()) {
        T.cls = OC_CYCLE_END;
        std::vector<BVec> u;
        std::vector<size_t> yoffs;
        for (BJob *J : end) {
            T.bytes[OC_CYCLE_END] += (double)(J->kend + 1) * J->n * (E->prec == PIMI_C32 ? 8 : 16);
            std::vector<cx<double>> yc(J->kend);
            for (int i = 0; i < J->kend; ++i) yc[i] = cx<double>{J->y[i].real(), J->y[i].imag()};
            yoffs.push_back(T.put(yc));  // coefficients travel with the descriptors
            u.push_back(BVec{0, J->V, nullptr, J->w, nullptr, J->n, J->nblk, J->kend, 0.0, 0});
        }
        const size_t uoff = T.gemv<R>(u, 1);
        for (size_t k = 0; k < yoffs.size(); ++k)
            T.relocs.push_back({uoff + k * sizeof(BVec) + offsetof(BVec, b), yoffs[k]});
        T.precond<R>(end, [](BJob *J) -> const void * { return J->w; }, [](BJob *J) -> void * { return J->z; });
        T.vec(vecs(end, [](BJob *J) { return BVec{0, J->z, nullptr, J->x, nullptr, J->n, J->nblk, 1, 0.0, 0}; }),
              kb_add<R>, Tick::nblk_of);
    }
    // -- FP64 bordered residual r = b - (L x / s + (t . x) v): new jobs and cycle ends
    std::vector<BJob *> resid(init);
    resid.insert(resid.end(), end.begin(), end.end());
    T.cls = OC_FP64;
    if (!resid.empty()) {
        T.apply<double>(resid, cX, W64, false);
        T.trace<double>(resid, cX, W64, true, false);
        T.vec(vecs(resid, [](BJob *J) { return BVec{0, J->b, J->w64, J->r, nullptr, J->n, J->nblk, 1, 0.0, 0}; }),
              kb_sub, Tick::nblk_of);
        T.norm<double>(resid, cR, S_RNORM);
        T.norm<double>(resid, cX, S_XNORM, 1);
    }
    // -- dynamics jobs: observables <obs_q, x> at a save time and |x|; an
    // Arnoldi tick of a factorization (V0 = x / |x| on its first); a
    // projection trial = V c (c = beta exp(h S) e1, staged) and |trial|
    if (!xobs.empty()) {
        T.cls = OC_OTHER;
        std::vector<BJob *> withobs;
        for (BJob *J : xobs)
            if (J->nobs > 0) withobs.push_back(J);
        if (!withobs.empty()) {
            std::vector<BVec> d, r;
            for (BJob *J : withobs) {
                const int nc = J->nobs + 1;  // the observables and a zero column in the "p" role
                d.push_back(BVec{0, J->obs, J->x, J->partial, nullptr, J->n, J->nblk, nc, 0.0, (uint64_t)kDotChunks});
                r.push_back(BVec{0, J->partial, nullptr, J->mail + kMailHacc, nullptr, J->n, J->nblk, nc, 0.0, 0});
            }
            T.vec(d, kb_dots2_partial<double>, [](const BVec &e) {
                return (int64_t)((e.nblk + kDotChunks - 1) / kDotChunks) * ((e.ncol + kDotCols - 1) / kDotCols);
            });
            T.vec(r, kb_reduce_cols2, [](const BVec &e) { return (int64_t)e.ncol; }, 64);
        }
        T.norm<double>(xobs, cX, S_RNORM);
    }
    T.cls = OC_OTHER;
    T.vec(vecs(xbegin, [](BJob *J) {
              return BVec{0, J->x, nullptr, J->V, J->scal() + S_RNORM, J->n, J->nblk, 1, 0.0, 0};
          }),
          kb_scale_div<R, double>, Tick::nblk_of);
    if (!xarn.empty()) {
        // a new factorization stages its steps 0..m (column m-1 is final after
        // step m); the enlargement after a near breakdown is one more step
        int most = 0;
        for (BJob *J : xarn) {
            J->arn_steps = J->arn.enlarging ? 1 : J->kdim + 1;
            most = std::max(most, J->arn_steps);
        }
        for (int i = 0; i < most; ++i) {
            std::vector<BJob *> at;
            for (BJob *J : xarn)
                if (i < J->arn_steps) at.push_back(J);
            T.cls = OC_APPLY;
            T.apply<R>(at, [i](BJob *J) -> const void * { return vcol<R>(J, J->j + i); },
                       [](BJob *J) -> void * { return J->w; }, true);
            T.dcgs<R>(at, [i](BJob *J) { return J->j + i; },
                      [i](BJob *J) { return J->mail + (size_t)i * (size_t)J->mail_stride; });
        }
    }
    if (!xproj.empty()) {
        T.cls = OC_CYCLE_END;
        std::vector<BVec> u, z;
        std::vector<size_t> coffs;
        std::vector<BJob *> proj;
        for (BJob *J : xproj) {
            if (J->F.zero_state) {
                z.push_back(BVec{0, nullptr, nullptr, J->r, nullptr, J->n, J->nblk, 1, 0.0, 0});
                continue;
            }
            T.bytes[OC_CYCLE_END] += (double)(J->F.coeff_count + 1) * J->n * sizeof(cx<R>);
            coffs.push_back(T.put(J->coeffs));  // coefficients travel with the descriptors
            u.push_back(BVec{0, J->V, nullptr, J->r, nullptr, J->n, J->nblk, J->F.coeff_count, 0.0, 0});
        }
        const size_t uoff = T.gemv<R>(u, 1);
        for (size_t k = 0; k < coffs.size(); ++k)
            T.relocs.push_back({uoff + k * sizeof(BVec) + offsetof(BVec, b), coffs[k]});
        T.vec(z, kb_zero, Tick::nblk_of);
        T.cls = OC_OTHER;
        T.norm<double>(xproj, cR, S_RNORM);
    }
    // -- final: back to coefficients (x = S^{-1} u for weighted jobs), |L x|, |x| and t . x
    auto cXo = [](BJob *J) -> const void * { return J->weighted ? (const void *)J->xo : (const void *)J->x; };
    if (!fin.empty()) {
        std::vector<BApply> v;
        int64_t tot = 0;
        for (BJob *J : fin)
            if (J->weighted) {
                pimi_plan *P = J->plan;
                v.push_back(BApply{tot, J->x, J->xo, P->offs.p, P->dims.p, P->tile_sector.p, nullptr, nullptr,
                                   nullptr, nullptr, P->d_inv_sigma.p, J->n, 1.0});
                tot += P->ntiles;
            }
        if (!v.empty()) {
            const size_t off = T.put(v);
            const int nj = (int)v.size();
            hipStream_t st = L.stream;
            T.push([=](const char *d) {
                hipLaunchKernelGGL(kb_sector_scale, dim3((unsigned)tot), dim3(kApplyBlock), 0, st,
                                   (const BApply *)(d + off), nj, tot);
            });
        }
        T.apply<double>(fin, cXo, W64, true);
        T.norm<double>(fin, cW64, S_LXNORM);
        T.norm<double>(fin, cXo, S_XNORM, 1);
        T.trace<double>(fin, cXo, nullptr, false, true);
    }
    // -- mailbox gather
    T.cls = OC_OTHER;
    std::vector<BCopy> g;
    size_t mtot = 0;
    std::vector<size_t> &moff = L.moff;
    moff.assign(L.active.size(), 0);
    for (size_t k = 0; k < L.active.size(); ++k) {
        BJob *J = L.active[k];
        const int64_t cnt = J->phase == PX_ARNOLDI  ? J->arn_steps * J->mail_stride
                            : J->phase == PH_ARNOLDI ? kMailHacc + 2 * (J->j + 1)
                            : J->phase == PX_OBSERVE ? kMailHacc + 2 * (J->nobs + 1)
                                                     : kMailHacc;
        moff[k] = mtot;
        g.push_back(BCopy{J->mail, nullptr, cnt});
        mtot += (size_t)cnt;
    }
    if (L.dmail.count < mtot) L.dmail.alloc(std::max<size_t>(mtot, 2 * L.dmail.count));
    for (size_t k = 0; k < g.size(); ++k) g[k].dst = L.dmail.p + moff[k];
    {
        const size_t off = T.put(g);
        const unsigned ng = (unsigned)g.size();
        hipStream_t st = L.stream;
        T.push([=](const char *d) {
            hipLaunchKernelGGL(kb_gather, dim3(ng), dim3(256), 0, st, (const BCopy *)(d + off));
        });
    }
    // -- upload, launch, download
    if (L.hstage_cap < T.buf.size()) {
        if (L.hstage) HIPCHK(hipHostFree(L.hstage));
        L.hstage_cap = std::max<size_t>(T.buf.size(), 2 * L.hstage_cap);
        HIPCHK(hipHostMalloc((void **)&L.hstage, L.hstage_cap, hipHostMallocDefault));
    }
    if (L.dstage.count < T.buf.size()) L.dstage.alloc(std::max<size_t>(T.buf.size(), 2 * L.dstage.count));
    const char *dbase = L.dstage.p;
    const bool timing = E->timing;
    if (timing)
        while (L.events.size() < T.ops.size() + 1) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            L.events.push_back(ev);
        }
    for (auto &rl : T.relocs) {
        const void *addr = dbase + rl.second;
        std::memcpy(T.buf.data() + rl.first, &addr, sizeof addr);
    }
    std::memcpy(L.hstage, T.buf.data(), T.buf.size());
    HIPCHK(hipMemcpyAsync(L.dstage.p, L.hstage, T.buf.size(), hipMemcpyHostToDevice, L.stream));
    if (E->use_graphs) {
        launch_graph(L, T, dbase);
    } else {
        if (timing) HIPCHK(hipEventRecord(L.events[0], L.stream));
        for (size_t k = 0; k < T.ops.size(); ++k) {
            T.ops[k](dbase);
            if (timing) HIPCHK(hipEventRecord(L.events[k + 1], L.stream));
        }
    }
    HIPCHK(hipGetLastError());
    if (L.hmail_cap < mtot) {
        if (L.hmail) HIPCHK(hipHostFree(L.hmail));
        L.hmail_cap = std::max<size_t>(mtot, 2 * L.hmail_cap);
        HIPCHK(hipHostMalloc((void **)&L.hmail, sizeof(cx<double>) * L.hmail_cap, hipHostMallocDefault));
    }
    HIPCHK(hipMemcpyAsync(L.hmail, L.dmail.p, sizeof(cx<double>) * mtot, hipMemcpyDeviceToHost, L.stream));
    for (BJob *J : fin) {
        J->result.resize((size_t)J->n);
        HIPCHK(hipMemcpyAsync(J->result.data(), J->weighted ? J->xo : J->x, sizeof(cx<double>) * J->n,
                              hipMemcpyDeviceToHost, L.stream));
    }
    for (BJob *J : xobs)  // a dynamics job's last save time: its final state
        if (J->sample + 1 == (int64_t)J->ts.size()) {
            J->result.resize((size_t)J->n);
            HIPCHK(hipMemcpyAsync(J->result.data(), J->x, sizeof(cx<double>) * J->n, hipMemcpyDeviceToHost,
                                  L.stream));
        }
    L.op_class = T.op_class;
    for (int c = 0; c < OC_COUNT; ++c) {
        L.bytes[c] = T.bytes[c];
        L.flops[c] = T.flops[c];
    }
    L.inflight = true;
}

// Dynamics jobs: the next attempt of the current interval (a new
// factorization when the last one was used by an accepted step, else a
// projection with the new step), as the loop of krylov_expv!
// (src/krylov_extensions.jl:1452-1568).  Returns false when the job has used
// its max_steps attempts.
static bool dyn_attempt(BJob *J) {
    if (J->attempts >= J->eo.max_steps) return false;
    J->attempts++;
    if (!J->have) {
        J->F = Factorization();
        J->factorizations++;
        J->have = true;
        if (J->xnorm == 0) {  // zero state: exp(h A) 0 = 0
            J->F.zero_state = true;
            J->F.completed = true;
        } else {
            J->F.beta = J->xnorm;
            J->j = 0;
            J->arn.reset(J->kdim, J->n);
            J->phase = PX_ARNOLDI;
            return true;
        }
    }
    J->err = J->F.zero_state ? 0.0 : expv_coefficients(J->F, J->kdim + 1, J->X.step(), J->coeffs);
    if (J->F.zero_state) J->X.step();
    for (auto &c : J->coeffs) c = cx<double>{J->F.beta * c.x, J->F.beta * c.y};
    J->phase = PX_PROJECT;
    return true;
}

// After an observation: the next interval, or the end of the job.  Returns
// false when the job cannot continue (max_steps).
static bool dyn_next_interval(BJob *J, bool &done) {
    if (J->sample == (int64_t)J->ts.size()) {
        done = true;
        return true;
    }
    const double dt = J->ts[(size_t)J->sample] - J->ts[(size_t)J->sample - 1];
    if (!J->X.begin(dt)) {  // a zero-length interval: the same state again
        J->phase = PX_OBSERVE;
        return true;
    }
    return dyn_attempt(J);
}

// One Arnoldi step of a dynamics job's factorization (its reductions at mb,
// scalars at sc): the Hessenberg column it finalizes, and at the end of the
// factorization the first projection.
static void dyn_arnoldi(BJob *J, const cx<double> *mb, const double *sc, bool &done, bool &ok) {
    const int jj = J->j;
    const cx<double> *sv = mb + kMailHacc;
    const int r = J->arn.step(jj, sv, sv + (jj + 1), sc[kSlotRho], cd(sc[kSlotE], sc[kSlotE + 1]));
    if (r < 0) {
        J->error = "zero or non-finite start vector of a Krylov factorization";
        done = true;
        ok = false;
        return;
    }
    if (r == 0) {
        J->j = jj + 1;
        return;
    }
    const double beta = J->F.beta;
    J->arn.finish(jj, beta, J->F);
    J->applications += J->F.used;
    J->err = expv_coefficients(J->F, J->kdim + 1, J->X.step(), J->coeffs);
    for (auto &c : J->coeffs) c = cx<double>{beta * c.x, beta * c.y};
    J->phase = PX_PROJECT;
}

// Wait for lane L's tick and advance every one of its jobs' state machines.
template <typename R>
static void complete_tick(pimi_batch *E, BLane &L) {
    if (!L.inflight) return;
    HIPCHK(hipStreamSynchronize(L.stream));
    L.inflight = false;
    E->ticks++;
    for (int c = 0; c < OC_COUNT; ++c) {
        E->stat_bytes[c] += L.bytes[c];
        E->stat_flops[c] += L.flops[c];
    }
    if (E->timing)
        for (size_t k = 0; k < L.op_class.size(); ++k) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, L.events[k], L.events[k + 1]));
            E->stat_ms[L.op_class[k]] += ms;
        }
    const std::vector<size_t> &moff = L.moff;
    // -- host updates
    std::vector<BJob *> still;
    for (size_t k = 0; k < L.active.size(); ++k) {
        BJob *J = L.active[k];
        const cx<double> *mb = L.hmail + moff[k];
        const double *sc = (const double *)mb;
        auto thr = [&]() { return 0.5 * J->tight * (J->o.atol + J->o.rtol * J->xnorm); };
        bool done = false, ok = false;
        switch (J->phase) {
            case PH_PROBE: {
                double s = 0;
                for (int q = 0; q < 3; ++q) s = std::max(s, sc[S_LPNORM + q] / sc[S_PNORM + q]);
                J->applications += 3;
                if (!(s > 0) || !std::isfinite(s)) {
                    J->error = "operator scale estimate is not finite and positive";
                    done = true;
                    break;
                }
                J->scale = s;
                J->phase = PH_INIT;
                break;
            }
            case PH_INIT:
            case PH_END: {
                if (J->phase == PH_END) J->cycles++;
                J->applications++;
                J->beta = sc[S_RNORM];
                J->xnorm = sc[S_XNORM];
                if (!std::isfinite(J->beta) || !std::isfinite(J->xnorm)) {
                    J->error = "non-finite residual";
                    J->phase = PH_FINAL;
                    J->candidate = false;
                } else if (J->beta <= thr()) {
                    J->phase = PH_FINAL;
                    J->candidate = true;
                } else if (J->iterations >= J->o.maxiter || J->cycles >= J->o.max_refinements) {
                    J->phase = PH_FINAL;
                    J->candidate = false;
                } else {
                    J->phase = PH_ARNOLDI;
                    J->begin_cycle = true;
                    J->j = 0;
                    J->beta0 = J->beta;
                    std::fill(J->H.begin(), J->H.end(), cd(0, 0));
                    std::fill(J->Hraw.begin(), J->Hraw.end(), cd(0, 0));
                    std::fill(J->g.begin(), J->g.end(), cd(0, 0));
                    J->g[0] = J->beta;
                }
                break;
            }
            case PH_ARNOLDI: {
                // Tick j finalized Hessenberg column j-1 (first-pass coefficients
                // c_{j-1} + reorthogonalization s, subdiagonal rho) and produced the
                // first-pass coefficients of column j:
                //   c_j = ((t - H[0:j, 0:j] s) / rho ; e - s_{j-1})
                J->begin_cycle = false;
                J->applications++;
                const int m = J->kdim, jj = J->j;
                auto Hr = [&](int i, int c) -> cd & { return J->H[(size_t)i + (size_t)c * (m + 1)]; };
                auto Hw = [&](int i, int c) -> cd & { return J->Hraw[(size_t)i + (size_t)c * (m + 1)]; };
                const cx<double> *sv = mb + kMailHacc, *tv = sv + (jj + 1);
                auto S_ = [&](int i) { return cd(sv[i].x, sv[i].y); };
                auto T_ = [&](int i) { return cd(tv[i].x, tv[i].y); };
                const double rho = sc[kSlotRho];
                const cd e(sc[kSlotE], sc[kSlotE + 1]);
                bool end_cycle = false;
                double est = INFINITY;
                if (jj >= 1) {
                    for (int i = 0; i < jj; ++i) Hw(i, jj - 1) = J->cprov[(size_t)i] + S_(i);
                    Hw(jj, jj - 1) = rho;
                    for (int i = 0; i <= jj; ++i) Hr(i, jj - 1) = Hw(i, jj - 1);
                    const int c = jj - 1;
                    for (int i = 0; i < c; ++i) {
                        const cd h1 = Hr(i, c), h2 = Hr(i + 1, c);
                        Hr(i, c) = J->cs[i] * h1 + J->sn[i] * h2;
                        Hr(i + 1, c) = -std::conj(J->sn[i]) * h1 + J->cs[i] * h2;
                    }
                    double gc;
                    cd gs;
                    complex_givens(Hr(c, c), Hr(c + 1, c), gc, gs);
                    J->cs[c] = gc;
                    J->sn[c] = gs;
                    const cd h1 = Hr(c, c), h2 = Hr(c + 1, c);
                    Hr(c, c) = gc * h1 + gs * h2;
                    Hr(c + 1, c) = 0;
                    const cd g1 = J->g[c], g2 = J->g[c + 1];
                    J->g[c] = gc * g1 + gs * g2;
                    J->g[c + 1] = -std::conj(gs) * g1 + gc * g2;
                    J->iterations++;
                    est = std::abs(J->g[c + 1]);
                    const double cyc = std::max(thr(), J->o.inner_rtol * J->beta0);
                    end_cycle = est <= cyc || jj >= m || J->iterations >= J->o.maxiter || rho == 0.0 ||
                                !std::isfinite(est);
                } else if (!(rho > 0) || !std::isfinite(rho)) {
                    J->error = "zero or non-finite start vector in a GMRES cycle";
                    J->phase = PH_FINAL;
                    J->candidate = false;
                    break;
                }
                if (!end_cycle) {
                    for (int i = 0; i < jj; ++i) {
                        cd acc = T_(i);
                        for (int c = 0; c < jj; ++c) acc -= Hw(i, c) * S_(c);
                        J->cprov[(size_t)i] = acc / rho;
                    }
                    J->cprov[(size_t)jj] = e - (jj >= 1 ? S_(jj - 1) : cd(0, 0));
                    J->j = jj + 1;
                } else {
                    const int kk = jj;
                    J->y.assign((size_t)kk, cd(0, 0));
                    bool singular = false;
                    for (int i = kk - 1; i >= 0; --i) {
                        cd acc = J->g[i];
                        for (int q = i + 1; q < kk; ++q) acc -= Hr(i, q) * J->y[q];
                        if (Hr(i, i) == cd(0, 0)) {
                            singular = true;
                            break;
                        }
                        J->y[i] = acc / Hr(i, i);
                    }
                    if (singular) {
                        J->error = "singular GMRES Hessenberg";
                        J->phase = PH_FINAL;
                        J->candidate = false;
                    } else {
                        J->kend = kk;
                        J->phase = PH_END;
                    }
                }
                if (E->trace && J->iterations % 50 == 0 && jj >= 1)
                    std::fprintf(stderr, "[pimi batch] job %lld iteration %lld estimate %.3e |x| %.3e\n",
                                 (long long)J->id, (long long)J->iterations, est, J->xnorm);
                break;
            }
            case PX_OBSERVE: {
                const cx<double> *ov = mb + kMailHacc + (J->nobs + 1);
                for (int32_t q = 0; q < J->nobs; ++q) J->obs_out[(size_t)(J->sample * J->nobs + q)] = ov[q];
                J->xnorm = sc[S_RNORM];
                J->sample++;
                if (!dyn_next_interval(J, done)) {
                    J->error = "adaptive Krylov exponential action did not reach the requested time";
                    done = true;
                } else if (done) {
                    ok = true;
                }
                break;
            }
            case PX_ARNOLDI:
                // the staged steps in order, until the factorization is
                // complete (a breakdown ends it early: later steps' values
                // are never read) or the job fails
                for (int i = 0; i < J->arn_steps && J->phase == PX_ARNOLDI && !done; ++i) {
                    const cx<double> *mi = mb + (size_t)i * (size_t)J->mail_stride;
                    dyn_arnoldi(J, mi, (const double *)mi, done, ok);
                }
                break;
            case PX_PROJECT: {
                int d = -1;
                try {
                    d = J->X.decide(J->err, sc[S_RNORM], [&]() { return J->xnorm; }, J->F, J->eo);
                } catch (const Error &er) {
                    J->error = er.what();
                    done = true;
                    break;
                }
                if (d > 0) {
                    std::swap(J->x, J->r);
                    J->xnorm = sc[S_RNORM];
                    J->accepted++;
                    J->error_sum += J->err;
                    J->have = false;
                } else {
                    J->rejected++;
                }
                bool go;
                if (d > 0 && J->X.done()) {
                    J->phase = PX_OBSERVE;
                    go = true;
                } else {
                    go = d >= 0 && dyn_attempt(J);
                }
                if (!go) {
                    J->error = "adaptive Krylov exponential action did not reach the requested time";
                    done = true;
                }
                break;
            }
            default: {  // PH_FINAL
                J->applications++;
                const double lx = sc[S_LXNORM], xn = sc[S_XNORM];
                const cd tr(mb[kMailTrace].x, mb[kMailTrace].y);
                const double nres = lx / J->scale;
                ok = J->candidate && J->error.empty() && nres <= J->o.atol + J->o.rtol * xn && tr != cd(0, 0);
                if (!ok && J->candidate && J->error.empty() && J->weighted && J->tight > 1e-4 &&
                    J->iterations < J->o.maxiter && J->cycles < J->o.max_refinements) {
                    // the weighted residual met its threshold but |L x| / (s |x|) did not
                    // follow: continue the refinement with a tighter threshold
                    J->tight *= 0.1;
                    J->phase = PH_ARNOLDI;
                    J->begin_cycle = true;
                    J->j = 0;
                    J->beta0 = J->beta;
                    std::fill(J->H.begin(), J->H.end(), cd(0, 0));
                    std::fill(J->Hraw.begin(), J->Hraw.end(), cd(0, 0));
                    std::fill(J->g.begin(), J->g.end(), cd(0, 0));
                    J->g[0] = J->beta;
                    J->result.clear();
                    break;
                }
                if (tr != cd(0, 0))
                    for (auto &zz : J->result) {
                        const cd v = cd(zz.x, zz.y) / tr;
                        zz = cx<double>{v.real(), v.imag()};
                    }
                cd tr2(0, 0);
                const pimi_plan *P = J->plan;
                for (size_t q = 0; q < P->tidx.size(); ++q)
                    tr2 += P->tval[q] * cd(J->result[P->tidx[q]].x, J->result[P->tidx[q]].y);
                pimi_solve_info &I = J->info;
                std::memset(&I, 0, sizeof I);
                I.converged = ok ? 1 : 0;
                I.iterations = J->iterations;
                I.restarts = J->cycles;
                I.refinements = J->cycles;
                I.applications = J->applications;
                I.residual = std::abs(tr) > 0 ? lx / std::abs(tr) : lx;
                I.normalized_residual = std::abs(tr) > 0 ? nres / std::abs(tr) : nres;
                I.trace_error = std::abs(tr2 - 1.0);
                I.operator_scale = J->scale;
                I.linear_residual = J->beta;
                I.relative_residual = xn > 0 ? nres / xn : INFINITY;
                I.weighted_relative_residual = NAN;  // not formed by the engine
                I.seconds =
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - J->t_admit).count();
                if (!ok && J->error.empty()) {
                    char buf[256];
                    std::snprintf(buf, sizeof buf,
                                  "batched stationary solve did not converge in %lld iterations (%lld cycles); "
                                  "relative residual %.3e",
                                  (long long)J->iterations, (long long)J->cycles, I.relative_residual);
                    J->error = buf;
                }
                done = true;
                break;
            }
        }
        if (done) finish(E, J, ok);
        else still.push_back(J);
    }
    L.active.swap(still);
}

// Complete every lane's tick in flight (before anything reads engine-wide
// state that the ticks update).
static void drain(pimi_batch *E) {
    for (auto &L : E->lanes) {
        if (E->prec == PIMI_C32) complete_tick<float>(E, L);
        else complete_tick<double>(E, L);
    }
}

}  // namespace pimi

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace pimi;

extern "C" {

pimi_status pimi_batch_create(pimi_context *ctx, const pimi_batch_opts *opts, pimi_batch **out) {
    return guarded([&] {
        if (!ctx || !opts || !out) throw Error(PIMI_EINVAL, "null argument");
        if (opts->precision != PIMI_C64 && opts->precision != PIMI_C32) throw Error(PIMI_EINVAL, "unknown precision");
        if (opts->max_jobs < 0 || opts->memory_bytes < 0) throw Error(PIMI_EINVAL, "negative engine limits");
        HIPCHK(hipSetDevice(ctx->device));
        std::unique_ptr<pimi_batch> E(new pimi_batch());
        E->ctx = ctx;
        E->prec = opts->precision;
        E->max_jobs = opts->max_jobs > 0 ? opts->max_jobs : 1024;
        size_t bytes = (size_t)opts->memory_bytes;
        if (bytes == 0) {
            size_t fr = 0, tot = 0;
            HIPCHK(hipMemGetInfo(&fr, &tot));
            bytes = fr / 4 * 3;
        }
        const char *tr = std::getenv("PIMI_TRACE");
        E->trace = tr && tr[0] == '1';
        E->dia_w = E->prec == PIMI_C32 ? 1 : 0;  // measured best at 512 qubit models: 8 columns FP32, 4 FP64
        if (const char *tb = std::getenv("PIMI_DIA_TB_BATCH"))
            for (int w = 0; w < kNumDiaWidths; ++w)
                if (kDiaWidths[w] == std::atoi(tb)) E->dia_w = w;
        E->timing = (opts->flags & PIMI_BATCH_TIMING) != 0;
        const char *ng = std::getenv("PIMI_NO_GRAPH");
        E->use_graphs = !(ng && ng[0] == '1') && !E->timing;
        HIPCHK(hipStreamCreate(&E->stream));
        for (auto &L : E->lanes) HIPCHK(hipStreamCreate(&L.stream));
        E->arena.init(align_up(bytes, 256));
        *out = E.release();
    });
}

void pimi_batch_destroy(pimi_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    delete b;
}

pimi_status pimi_batch_job_bytes(const pimi_plan *plan, int32_t precision, int32_t krylovdim, int64_t *bytes) {
    return guarded([&] {
        if (!plan || !bytes) throw Error(PIMI_EINVAL, "null argument");
        if (precision != PIMI_C64 && precision != PIMI_C32) throw Error(PIMI_EINVAL, "unknown precision");
        if (krylovdim < 1 || krylovdim > kMaxKrylov) throw Error(PIMI_EINVAL, "krylovdim must lie in 1..4000");
        const int m = (int)std::min<int64_t>(krylovdim, plan->n);  // min(krylovdim, n) basis columns
        *bytes = (int64_t)job_bytes(plan->n, (int64_t)plan->tidx.size(), m, precision);
    });
}

pimi_status pimi_batch_submit(pimi_batch *E, pimi_plan *plan, const pimi_gmres_opts *o, int64_t *job_id) {
    return guarded([&] {
        if (!E || !plan || !o || !job_id) throw Error(PIMI_EINVAL, "null argument");
        BusyGuard guard(E->busy);
        if (!plan->ctx || plan->ctx->device != E->ctx->device)
            throw Error(PIMI_EINVAL, "the plan must live on the engine's device");
        if (o->krylovdim < 1 || o->krylovdim > kMaxKrylov) throw Error(PIMI_EINVAL, "krylovdim must lie in 1..4000");
        if (o->preconditioner != PIMI_PRECOND_SECTOR)
            throw Error(PIMI_EINVAL, "batched solves use the sector preconditioner");
        if (o->criterion != PIMI_CRITERION_BACKWARD)
            throw Error(PIMI_EINVAL, "batched solves use the scale-invariant (backward) criterion");
        if (!(o->atol >= 0) || !(o->rtol >= 0)) throw Error(PIMI_EINVAL, "atol and rtol must be nonnegative");
        if (!(o->inner_rtol > 0) || !(o->inner_rtol < 1)) throw Error(PIMI_EINVAL, "inner_rtol must lie in (0, 1)");
        if (o->maxiter < 1 || o->max_refinements < 1) throw Error(PIMI_EINVAL, "maxiter and max_refinements must be >= 1");
        if (!(o->operator_scale >= 0) || !std::isfinite(o->operator_scale))
            throw Error(PIMI_EINVAL, "operator_scale must be finite and nonnegative");
        if (o->weighting != PIMI_WEIGHT_NONE && o->weighting != PIMI_WEIGHT_SECTOR)
            throw Error(PIMI_EINVAL, "unknown weighting");
        HIPCHK(hipSetDevice(E->ctx->device));
        std::unique_ptr<BJob> J(new BJob());
        J->plan = plan;
        J->o = *o;
        J->n = plan->n;
        J->nt = (int64_t)plan->tidx.size();
        J->kdim = (int)std::min<int64_t>(o->krylovdim, plan->n);  // as KrylovWorkspace (src/krylov.jl:74-82)
        J->nblk = blocks_for(J->n);
        J->bytes = job_bytes(J->n, J->nt, J->kdim, E->prec);
        {
            J->table_bytes64 = (double)plan->main.table_bytes(16);
            J->table_bytes32 = (double)plan->main.table_bytes(8);
        }
        if (J->bytes > E->arena.size)
            throw Error(PIMI_ENOMEM, "the job's device state exceeds the engine's memory reservation");
        ensure_plan_preconditioner(plan);
        {
            double m3 = 0;
            for (auto m : plan->host.dims) m3 += (double)m * m * m;
            // four m x m products per sector: 8 flops per complex fma, 4 with a real factor U
            J->gemm_flops = 4 * (plan->u_real ? 4.0 : 8.0) * m3;
        }
        J->weighted = o->weighting == PIMI_WEIGHT_SECTOR;
        if (J->weighted) ensure_plan_weighting(plan);
        double tt = 0;
        for (double t : plan->tval) tt += t * t;
        J->vv.resize(plan->tval.size());
        for (size_t k = 0; k < J->vv.size(); ++k) J->vv[k] = plan->tval[k] / tt;
        if (J->weighted) {  // v -> S v (trace support ordered by sector, m entries each)
            size_t k = 0;
            for (int q = 0; q < plan->ns; ++q)
                for (int64_t a = 0; a < plan->host.dims[(size_t)q]; ++a, ++k) J->vv[k] *= plan->sigma[(size_t)q];
        }
        const int m = J->kdim;
        J->H.assign((size_t)(m + 1) * m, cd(0, 0));
        J->Hraw.assign((size_t)(m + 1) * m, cd(0, 0));
        J->cprov.assign((size_t)m + 2, cd(0, 0));
        J->g.assign((size_t)m + 1, cd(0, 0));
        J->sn.assign((size_t)m, cd(0, 0));
        J->cs.assign((size_t)m, 0.0);
        J->t_submit = std::chrono::steady_clock::now();
        J->id = E->next_id++;
        *job_id = J->id;
        E->queue.push_back(J.get());
        E->jobs[J->id] = std::move(J);
    });
}

pimi_status pimi_batch_submit_expv(pimi_batch *E, pimi_plan *plan, const double *rho0, const double *ts,
                                   int64_t nts, const pimi_expv_opts *o, int32_t nobs, const double *obs,
                                   int64_t *job_id) {
    return guarded([&] {
        if (!E || !plan || !rho0 || !ts || !o || !job_id || (nobs > 0 && !obs)) throw Error(PIMI_EINVAL, "null argument");
        BusyGuard guard(E->busy);
        if (!plan->ctx || plan->ctx->device != E->ctx->device)
            throw Error(PIMI_EINVAL, "the plan must live on the engine's device");
        if (E->prec != PIMI_C64) throw Error(PIMI_EINVAL, "dynamics jobs need an FP64 engine");
        if (o->krylovdim < 1 || o->krylovdim > kMaxKrylov) throw Error(PIMI_EINVAL, "krylovdim must lie in 1..4000");
        if (!(o->safety > 0 && o->safety < 1)) throw Error(PIMI_EINVAL, "safety must lie strictly between zero and one");
        if (o->max_steps <= 0) throw Error(PIMI_EINVAL, "max_steps must be positive");
        if (!(o->atol >= 0) || !(o->rtol >= 0)) throw Error(PIMI_EINVAL, "atol and rtol must be nonnegative");
        if (nts < 1) throw Error(PIMI_EINVAL, "at least one output time is required");
        if (nobs < 0) throw Error(PIMI_EINVAL, "nobs must be nonnegative");
        for (int64_t k = 0; k < nts; ++k)
            if (!std::isfinite(ts[k])) throw Error(PIMI_EINVAL, "output times must be finite");
        for (int64_t k = 1; k < nts; ++k)
            if (!(ts[k] >= ts[k - 1])) throw Error(PIMI_EINVAL, "Krylov exponential output times must be nondecreasing");
        HIPCHK(hipSetDevice(E->ctx->device));
        std::unique_ptr<BJob> J(new BJob());
        J->kind = JK_DYNAMICS;
        J->plan = plan;
        J->eo = *o;
        J->n = plan->n;
        J->nt = 0;
        J->kdim = (int)std::min<int64_t>(o->krylovdim, plan->n);  // as KrylovWorkspace (src/krylov.jl:74-82)
        J->nblk = blocks_for(J->n);
        J->nobs = nobs;
        J->bytes = dyn_job_bytes(J->n, J->kdim, nobs);
        J->table_bytes64 = (double)plan->main.table_bytes(16);
        J->table_bytes32 = (double)plan->main.table_bytes(8);
        if (J->bytes > E->arena.size)
            throw Error(PIMI_ENOMEM, "the job's device state exceeds the engine's memory reservation");
        J->ts.assign(ts, ts + nts);
        J->rho0.resize((size_t)J->n);
        std::memcpy(J->rho0.data(), rho0, sizeof(cx<double>) * (size_t)J->n);
        J->obs_host.resize((size_t)J->n * (size_t)nobs);
        if (nobs > 0) std::memcpy(J->obs_host.data(), obs, sizeof(cx<double>) * J->obs_host.size());
        J->obs_out.assign((size_t)nts * (size_t)nobs, cx<double>{0, 0});
        J->t_submit = std::chrono::steady_clock::now();
        J->id = E->next_id++;
        *job_id = J->id;
        E->queue.push_back(J.get());
        E->jobs[J->id] = std::move(J);
    });
}

pimi_status pimi_batch_run(pimi_batch *E, int32_t max_ticks, int32_t stop_when_finished, int64_t *finished) {
    return guarded([&] {
        if (!E) throw Error(PIMI_EINVAL, "null argument");
        BusyGuard guard(E->busy);
        HIPCHK(hipSetDevice(E->ctx->device));
        // One round: each lane reads back its tick in flight (while the other
        // lane's tick runs) and launches its next.  A lane's tick may still
        // be in flight when this returns; the next run (or stats, or destroy)
        // completes it.  Jobs appear in the finished list only once their
        // last tick is complete, so fetch never waits on the device.
        const size_t before = E->finished.size();
        for (int32_t t = 0; t < max_ticks; ++t) {
            if (E->nactive() == 0 && E->queue.empty() && !E->any_inflight()) break;
            const size_t f0 = E->finished.size();
            for (auto &L : E->lanes) {
                if (E->prec == PIMI_C32) {
                    complete_tick<float>(E, L);
                    launch_tick<float>(E, L);
                } else {
                    complete_tick<double>(E, L);
                    launch_tick<double>(E, L);
                }
            }
            if (stop_when_finished && E->finished.size() > f0) break;
        }
        if (finished) *finished = (int64_t)(E->finished.size() - before);
    });
}

pimi_status pimi_batch_counts(pimi_batch *E, int64_t *counts) {
    return guarded([&] {
        if (!E || !counts) throw Error(PIMI_EINVAL, "null argument");
        counts[0] = (int64_t)E->queue.size();
        counts[1] = (int64_t)E->nactive();
        counts[2] = (int64_t)E->finished.size();
        counts[3] = E->ticks;
        counts[4] = (int64_t)E->arena.used;
        counts[5] = (int64_t)E->arena.size;
    });
}

pimi_status pimi_batch_graph_counts(pimi_batch *E, int64_t *counts) {
    return guarded([&] {
        if (!E || !counts) throw Error(PIMI_EINVAL, "null argument");
        counts[0] = counts[1] = counts[2] = 0;
        for (const auto &L : E->lanes) {
            counts[0] += L.graph_ticks;
            counts[1] += L.graph_instantiations;
            counts[2] += L.graph_updates;
        }
    });
}

pimi_status pimi_batch_stats(pimi_batch *E, double *out, int32_t reset) {
    return guarded([&] {
        if (!E || !out) throw Error(PIMI_EINVAL, "null argument");
        BusyGuard guard(E->busy);
        HIPCHK(hipSetDevice(E->ctx->device));
        drain(E);  // statistics of every launched tick
        for (int c = 0; c < 8; ++c) {
            out[c] = E->stat_ms[c];
            out[8 + c] = E->stat_bytes[c];
            out[16 + c] = E->stat_flops[c];
        }
        if (reset)
            for (int c = 0; c < 8; ++c) E->stat_ms[c] = E->stat_bytes[c] = E->stat_flops[c] = 0;
    });
}

pimi_status pimi_batch_next_finished(pimi_batch *E, int64_t *job_id) {
    return guarded([&] {
        if (!E || !job_id) throw Error(PIMI_EINVAL, "null argument");
        BusyGuard guard(E->busy);
        if (E->finished.empty()) {
            *job_id = -1;
            return;
        }
        *job_id = E->finished.front();
        E->finished.pop_front();
    });
}

pimi_status pimi_batch_fetch_expv(pimi_batch *E, int64_t job_id, double *obs_out, double *rho_final,
                                  pimi_solve_info *info) {
    pimi_status st = PIMI_OK;
    std::string msg;
    pimi_status g = guarded([&] {
        if (!E || !info) throw Error(PIMI_EINVAL, "null argument");
        BusyGuard guard(E->busy);
        auto it = E->jobs.find(job_id);
        if (it == E->jobs.end()) throw Error(PIMI_EINVAL, "unknown batch job");
        BJob *J = it->second.get();
        if (J->kind != JK_DYNAMICS) throw Error(PIMI_EINVAL, "not a dynamics job");
        if (J->state != JS_DONE && J->state != JS_FAILED) throw Error(PIMI_EBUSY, "the batch job has not finished");
        if (J->nobs > 0 && !obs_out) throw Error(PIMI_EINVAL, "null argument");
        if (obs_out) std::memcpy(obs_out, J->obs_out.data(), sizeof(cx<double>) * J->obs_out.size());
        if (rho_final) std::memcpy(rho_final, J->result.data(), sizeof(cx<double>) * (size_t)J->n);
        *info = J->info;
        if (J->state == JS_FAILED) {
            st = PIMI_ENOTCONVERGED;
            msg = J->error.empty() ? "adaptive Krylov exponential action did not reach the requested time" : J->error;
        }
        for (auto q = E->finished.begin(); q != E->finished.end(); ++q)
            if (*q == job_id) {
                E->finished.erase(q);
                break;
            }
        E->jobs.erase(it);
    });
    if (g != PIMI_OK) return g;
    if (st != PIMI_OK) set_error(msg);
    return st;
}

pimi_status pimi_batch_fetch(pimi_batch *E, int64_t job_id, double *rho_out, pimi_solve_info *info) {
    pimi_status st = PIMI_OK;
    std::string msg;
    pimi_status g = guarded([&] {
        if (!E || !rho_out || !info) throw Error(PIMI_EINVAL, "null argument");
        BusyGuard guard(E->busy);
        auto it = E->jobs.find(job_id);
        if (it == E->jobs.end()) throw Error(PIMI_EINVAL, "unknown batch job");
        BJob *J = it->second.get();
        if (J->state != JS_DONE && J->state != JS_FAILED) throw Error(PIMI_EBUSY, "the batch job has not finished");
        std::memcpy(rho_out, J->result.data(), sizeof(cx<double>) * (size_t)J->n);
        *info = J->info;
        if (J->state == JS_FAILED) {
            st = PIMI_ENOTCONVERGED;
            msg = J->error.empty() ? "batched stationary solve did not converge" : J->error;
        }
        for (auto q = E->finished.begin(); q != E->finished.end(); ++q)
            if (*q == job_id) {
                E->finished.erase(q);
                break;
            }
        E->jobs.erase(it);
    });
    if (g != PIMI_OK) return g;
    if (st != PIMI_OK) set_error(msg);
    return st;
}

pimi_status pimi_batch_autotune_apply(pimi_batch *E, pimi_plan *const *plans, int32_t nplans, int32_t reps,
                                      int32_t *width_out, int32_t *widths_out) {
    return guarded([&] {
        if (!E || !plans || nplans < 1 || reps < 1 || !width_out) throw Error(PIMI_EINVAL, "invalid argument");
        BusyGuard guard(E->busy);
        HIPCHK(hipSetDevice(E->ctx->device));
        const size_t es = E->prec == PIMI_C32 ? sizeof(cx<float>) : sizeof(cx<double>);
        std::vector<DevBuf<char>> xs((size_t)nplans), ys((size_t)nplans);
        std::vector<pimi_plan *> banded;
        for (int32_t k = 0; k < nplans; ++k) {
            pimi_plan *P = plans[k];
            if (!P || !P->ctx || P->ctx->device != E->ctx->device)
                throw Error(PIMI_EINVAL, "every plan must live on the engine's device");
            if (!P->main.dia_ok) continue;
            xs[banded.size()].alloc(es * (size_t)P->n);
            ys[banded.size()].alloc(es * (size_t)P->n);
            HIPCHK(hipMemsetAsync(xs[banded.size()].p, 0x3f, es * (size_t)P->n, E->stream));
            banded.push_back(P);
        }
        auto report = [&] {
            *width_out = kDiaWidths[E->dia_w];
            if (widths_out)
                for (int32_t k = 0; k < nplans; ++k)
                    widths_out[k] = plans[k]->main.dia_ok ? kDiaWidths[E->width_for(plans[k])] : 0;
        };
        if (banded.empty()) {  // nothing banded: the width is not used
            report();
            return;
        }
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        DevBuf<BDia> dvd;
        dvd.alloc(banded.size());
        // the fastest width for one launch over the plans idx (ties: the first)
        auto fastest = [&](const std::vector<size_t> &idx, int dflt) {
            int best = dflt;
            float best_ms = INFINITY;
            for (int w = 0; w < kNumDiaWidths; ++w) {
                std::vector<BDia> vd;
                std::vector<const pimi_plan *> ps;
                int64_t totd = 0;
                for (size_t k : idx) {
                    vd.push_back(E->prec == PIMI_C32
                                     ? dia_desc<float>(totd, banded[k], xs[k].p, ys[k].p, false, 1.0, w)
                                     : dia_desc<double>(totd, banded[k], xs[k].p, ys[k].p, false, 1.0, w));
                    ps.push_back(banded[k]);
                    totd += banded[k]->ndtiles[w];
                }
                HIPCHK(hipMemcpyAsync(dvd.p, vd.data(), sizeof(BDia) * vd.size(), hipMemcpyHostToDevice,
                                      E->stream));
                const int nvd = (int)vd.size();
                const DevStencil *M = common_programs(ps.begin(), ps.end());
                auto launch = [&] {
                    if (E->prec == PIMI_C32) launch_kb_dia<float>(E->stream, dvd.p, nvd, totd, w, M);
                    else launch_kb_dia<double>(E->stream, dvd.p, nvd, totd, w, M);
                };
                launch();  // warm
                HIPCHK(hipEventRecord(e0, E->stream));
                for (int32_t r = 0; r < reps; ++r) launch();
                HIPCHK(hipEventRecord(e1, E->stream));
                HIPCHK(hipGetLastError());
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best_ms) {
                    best_ms = ms;
                    best = w;
                }
            }
            return best;
        };
        // one launch over every plan: the width of sizes not tuned here
        std::vector<size_t> all(banded.size());
        for (size_t k = 0; k < all.size(); ++k) all[k] = k;
        E->dia_w = fastest(all, E->dia_w);
        // then each plan size on its own
        std::map<std::pair<int64_t, int>, std::vector<size_t>> sizes;
        for (size_t k = 0; k < banded.size(); ++k) sizes[{banded[k]->n, banded[k]->ns}].push_back(k);
        for (const auto &kv : sizes) E->dia_w_size[kv.first] = fastest(kv.second, E->dia_w);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        report();
    });
}

// One launch applying L to one vector of every plan (engine precision):
// descriptors for the banded and the ELL kernels, uploaded once; `launch`
// enqueues the apply on the engine stream.  xs[k], ys[k]: device vectors of
// plan k.  Shared by the multi-model apply and its throughput probe.
struct MultiApply {
    DevBuf<BApply> dv;
    DevBuf<BDia> dvd;
    const DevStencil *programs = nullptr;
    int nv = 0, nvd = 0;
    int64_t tot = 0, totd = 0, per_xcd = 0, bytes = 0;

    void prepare(pimi_batch *E, pimi_plan *const *plans, int32_t nplans, void *const *xs, void *const *ys) {
        const size_t es = E->prec == PIMI_C32 ? sizeof(cx<float>) : sizeof(cx<double>);
        std::vector<BApply> v;
        std::vector<BDia> vd;
        programs = common_programs(plans, plans + nplans);
        for (int32_t k = 0; k < nplans; ++k) {
            pimi_plan *P = plans[k];
            bytes += 2 * P->n * (int64_t)es + P->main.table_bytes((int)es);
            if (P->main.dia_ok) {
                vd.push_back(E->prec == PIMI_C32 ? dia_desc<float>(totd, P, xs[k], ys[k], false, 1.0, E->dia_w)
                                                 : dia_desc<double>(totd, P, xs[k], ys[k], false, 1.0, E->dia_w));
                totd += P->ndtiles[E->dia_w];
                continue;
            }
            const void *val = E->prec == PIMI_C32 ? (const void *)P->main.val32.p : (const void *)P->main.val64.p;
            v.push_back(BApply{tot, xs[k], ys[k], P->offs.p, P->dims.p, P->tile_sector.p, P->atiles4.p,
                               P->main.entry_base.p, P->main.entries.p, P->main.col.p, val, P->n, 1.0});
            tot += P->natiles4;
        }
        if (!v.empty()) {
            dv.alloc(v.size());
            HIPCHK(hipMemcpyAsync(dv.p, v.data(), sizeof(BApply) * v.size(), hipMemcpyHostToDevice, E->stream));
        }
        if (!vd.empty()) {
            dvd.alloc(vd.size());
            HIPCHK(hipMemcpyAsync(dvd.p, vd.data(), sizeof(BDia) * vd.size(), hipMemcpyHostToDevice, E->stream));
        }
        HIPCHK(hipStreamSynchronize(E->stream));  // the host descriptor arrays go out of scope
        per_xcd = (tot + kNumXcd - 1) / kNumXcd;
        nv = (int)v.size();
        nvd = (int)vd.size();
    }
    void launch(pimi_batch *E) const {
        if (nvd) {
            if (E->prec == PIMI_C32) launch_kb_dia<float>(E->stream, dvd.p, nvd, totd, E->dia_w, programs);
            else launch_kb_dia<double>(E->stream, dvd.p, nvd, totd, E->dia_w, programs);
        }
        if (nv) {
            if (E->prec == PIMI_C32)
                hipLaunchKernelGGL(kb_apply<float>, dim3((unsigned)(per_xcd * kNumXcd)), dim3(kApplyWave), 0,
                                   E->stream, dv.p, nv, tot, per_xcd);
            else
                hipLaunchKernelGGL(kb_apply<double>, dim3((unsigned)(per_xcd * kNumXcd)), dim3(kApplyWave), 0,
                                   E->stream, dv.p, nv, tot, per_xcd);
        }
        HIPCHK(hipGetLastError());
    }
};

static void check_multi_plans(pimi_batch *E, pimi_plan *const *plans, int32_t nplans) {
    for (int32_t k = 0; k < nplans; ++k) {
        pimi_plan *P = plans[k];
        if (!P || !P->ctx || P->ctx->device != E->ctx->device)
            throw Error(PIMI_EINVAL, "every plan must live on the engine's device");
    }
}

pimi_status pimi_batch_apply(pimi_batch *E, pimi_plan *const *plans, int32_t nplans, const double *x, double *y) {
    return guarded([&] {
        if (!E || !plans || nplans < 1 || !x || !y) throw Error(PIMI_EINVAL, "invalid argument");
        BusyGuard guard(E->busy);
        HIPCHK(hipSetDevice(E->ctx->device));
        check_multi_plans(E, plans, nplans);
        const bool f32 = E->prec == PIMI_C32;
        const size_t es = f32 ? sizeof(cx<float>) : sizeof(cx<double>);
        int64_t total = 0;
        std::vector<int64_t> off((size_t)nplans + 1, 0);
        for (int32_t k = 0; k < nplans; ++k) off[(size_t)k + 1] = off[(size_t)k] + plans[k]->n;
        total = off[(size_t)nplans];
        const double *ylo = y, *yhi = y + 2 * total, *xlo = x, *xhi = x + 2 * total;
        if (!(yhi <= xlo || xhi <= ylo)) throw Error(PIMI_EINVAL, "x and y must not overlap");
        DevBuf<char> dx, dy;
        dx.alloc(es * (size_t)total);
        dy.alloc(es * (size_t)total);
        std::vector<float> xf;
        if (f32) {
            xf.resize((size_t)(2 * total));
            for (int64_t i = 0; i < 2 * total; ++i) xf[(size_t)i] = (float)x[i];
            HIPCHK(hipMemcpyAsync(dx.p, xf.data(), es * (size_t)total, hipMemcpyHostToDevice, E->stream));
        } else {
            HIPCHK(hipMemcpyAsync(dx.p, x, es * (size_t)total, hipMemcpyHostToDevice, E->stream));
        }
        std::vector<void *> xs((size_t)nplans), ys((size_t)nplans);
        for (int32_t k = 0; k < nplans; ++k) {
            xs[(size_t)k] = dx.p + es * (size_t)off[(size_t)k];
            ys[(size_t)k] = dy.p + es * (size_t)off[(size_t)k];
        }
        MultiApply M;
        M.prepare(E, plans, nplans, xs.data(), ys.data());
        M.launch(E);
        if (f32) {
            std::vector<float> yf((size_t)(2 * total));
            HIPCHK(hipMemcpyAsync(yf.data(), dy.p, es * (size_t)total, hipMemcpyDeviceToHost, E->stream));
            HIPCHK(hipStreamSynchronize(E->stream));
            for (int64_t i = 0; i < 2 * total; ++i) y[i] = (double)yf[(size_t)i];
        } else {
            HIPCHK(hipMemcpyAsync(y, dy.p, es * (size_t)total, hipMemcpyDeviceToHost, E->stream));
            HIPCHK(hipStreamSynchronize(E->stream));
        }
    });
}

pimi_status pimi_batch_apply_bench(pimi_batch *E, pimi_plan *const *plans, int32_t nplans, int32_t reps,
                                   double *ms_avg, int64_t *algorithmic_bytes) {
    return guarded([&] {
        if (!E || !plans || nplans < 1 || reps < 1 || !ms_avg || !algorithmic_bytes)
            throw Error(PIMI_EINVAL, "invalid argument");
        BusyGuard guard(E->busy);
        HIPCHK(hipSetDevice(E->ctx->device));
        check_multi_plans(E, plans, nplans);
        const size_t es = E->prec == PIMI_C32 ? sizeof(cx<float>) : sizeof(cx<double>);
        std::vector<DevBuf<char>> xs((size_t)nplans), ys((size_t)nplans);
        std::vector<void *> xp((size_t)nplans), yp((size_t)nplans);
        for (int32_t k = 0; k < nplans; ++k) {
            pimi_plan *P = plans[k];
            xs[k].alloc(es * (size_t)P->n);
            ys[k].alloc(es * (size_t)P->n);
            HIPCHK(hipMemsetAsync(xs[k].p, 0x3f, es * (size_t)P->n, E->stream));
            xp[(size_t)k] = xs[k].p;
            yp[(size_t)k] = ys[k].p;
        }
        MultiApply M;
        M.prepare(E, plans, nplans, xp.data(), yp.data());
        M.launch(E);  // warm-up
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, E->stream));
        for (int32_t r = 0; r < reps; ++r) M.launch(E);
        HIPCHK(hipEventRecord(e1, E->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_avg = ms / reps;
        *algorithmic_bytes = M.bytes;
    });
}

}  // extern "C"
