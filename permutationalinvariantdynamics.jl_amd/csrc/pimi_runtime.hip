// pimi runtime: device plan, workspaces, Krylov drivers and the C ABI.
//
// One workspace = one HIP stream + one Krylov basis, owned by one caller at a
// time (a concurrent second caller gets PIMI_EBUSY, as the reference's
// threaded workspace rejects concurrent use, src/threaded_apply.jl:575-576).
// The Liouvillian apply never uses atomics: every output coordinate is
// written by exactly one lane.
//
// Solver control follows the reference:
//   restarted GMRES + stationary-state wrapper  src/krylov.jl:710-891
//   adaptive Krylov exponential action           src/krylov_extensions.jl:1346-1568
//   solve_dynamics interval loop + observables   src/highlevel.jl:489-525,614-733
// Orthogonalization is classical Gram--Schmidt with one full
// reorthogonalization (CGS2: two GEMV-shaped passes over the basis) instead of
// the reference's modified Gram--Schmidt with two passes (MGS2); both give a
// basis orthogonal to working precision, and the solves agree to the solver
// tolerance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#include <cfloat>

#include "pimi_internal.hpp"
#include "pimi_expv.hpp"

namespace pimi {

constexpr int kDotBlocks = 240;  // partial-sum blocks per column (multiple of 8 XCDs)
constexpr int kMaxObservables = 100;

// ---------------------------------------------------------------------------
// Krylov kernels
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_dots_partial(const cx<R> *__restrict__ V, int64_t ldv,
                                                            const cx<R> *__restrict__ w, int64_t n,
                                                            cx<double> *__restrict__ partial) {
    __shared__ double sh[2 * (kRedBlock / 64)];
    const int c = blockIdx.y;
    const cx<R> *v = V + (int64_t)c * ldv;
    double re = 0, im = 0;
    for (int64_t k = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRedBlock) {
        cx<R> a = v[k], b = w[k];
        re += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
        im += (double)a.x * (double)b.y - (double)a.y * (double)b.x;
    }
    double sums[2] = {re, im};
    block_sums(sums, sh);
    if (threadIdx.x == 0) partial[(int64_t)c * gridDim.x + blockIdx.x] = cx<double>{sums[0], sums[1]};
}

// h[c] = sum of the partials of column c; if hacc != null, hacc[c] += h[c].
__global__ __launch_bounds__(kRedBlock) void k_reduce_cols(const cx<double> *__restrict__ partial, int nb,
                                                           cx<double> *__restrict__ h, cx<double> *hacc) {
    __shared__ double sh[kRedBlock / 64];
    const int c = blockIdx.x;
    double re = 0, im = 0;
    for (int k = threadIdx.x; k < nb; k += kRedBlock) {
        re += partial[(int64_t)c * nb + k].x;
        im += partial[(int64_t)c * nb + k].y;
    }
    re = block_sum(re, sh);
    __syncthreads();
    im = block_sum(im, sh);
    if (threadIdx.x == 0) {
        h[c] = cx<double>{re, im};
        if (hacc) hacc[c] = cx<double>{hacc[c].x + re, hacc[c].y + im};
    }
}

// w -= V[:, 0:ncol] h

template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_gs_update(cx<R> *__restrict__ w, const cx<R> *__restrict__ V,
                                                         int64_t ldv, int ncol, const cx<double> *__restrict__ h,
                                                         int64_t n) {
    __shared__ cx<double> hs[kColChunk];
    for (int64_t k0 = (int64_t)blockIdx.x * kRedBlock; k0 < n; k0 += (int64_t)gridDim.x * kRedBlock) {
        const int64_t k = k0 + threadIdx.x;
        double re = 0, im = 0;
        if (k < n) { re = w[k].x; im = w[k].y; }
        for (int c0 = 0; c0 < ncol; c0 += kColChunk) {
            const int cn = min(kColChunk, ncol - c0);
            __syncthreads();
            for (int c = threadIdx.x; c < cn; c += kRedBlock) hs[c] = h[c0 + c];
            __syncthreads();
            if (k < n)
                for (int c = 0; c < cn; ++c) {
                    cx<R> v = V[k + (int64_t)(c0 + c) * ldv];
                    re -= (double)v.x * hs[c].x - (double)v.y * hs[c].y;
                    im -= (double)v.x * hs[c].y + (double)v.y * hs[c].x;
                }
        }
        if (k < n) w[k] = cx<R>{(R)re, (R)im};
    }
}

// ---------------------------------------------------------------------------
// Reductions finished by their consumer.  A producer kernel leaves per-
// workgroup partial sums; the next kernel, which needs the total anyway, has
// every workgroup add them up itself in one fixed order (lane-strided, then
// across the wave), so all workgroups hold bitwise the same total and one
// launch per reduction disappears without atomics or device-wide fences.
// ---------------------------------------------------------------------------
constexpr int kFusedDotCols = 16;  // widest basis whose dot partials a consumer reduces
constexpr int kNormBlocks = 512;   // grid of kernels that leave norm partials

// sum over q < G of p[q * stride] by one wave
__device__ __forceinline__ double wave_total(const double *p, int G, int stride = 1) {
    const int lane = threadIdx.x & 63;
    double t = 0;
    for (int q = lane; q < G; q += 64) t += p[(int64_t)q * stride];
    return wave_sum(t);
}

// w -= V[:, 0:ncol] h with h the reduction of the dot partials
// dpart[c * G + q] (q < G) of k_dots_partial (ncol <= kFusedDotCols);
// workgroup 0 adds h into hacc.  With npart != null, each workgroup leaves
// its share of ||w||^2 (of the stored w) in npart[blockIdx.x].
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_gs_update_r(cx<R> *__restrict__ w, const cx<R> *__restrict__ V,
                                                           int64_t ldv, int ncol, const cx<double> *__restrict__ dpart,
                                                           int G, cx<double> *hacc, int64_t n,
                                                           double *__restrict__ npart) {
    __shared__ cx<double> hs[kFusedDotCols];
    __shared__ double sh[kRedBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int c = wid; c < ncol; c += kRedBlock / 64) {
        const double re = wave_total(&dpart[(int64_t)c * G].x, G, 2);
        const double im = wave_total(&dpart[(int64_t)c * G].y, G, 2);
        if (lane == 0) hs[c] = cx<double>{re, im};
    }
    __syncthreads();
    if (blockIdx.x == 0 && hacc && (int)threadIdx.x < ncol)
        hacc[threadIdx.x] = cx<double>{hacc[threadIdx.x].x + hs[threadIdx.x].x, hacc[threadIdx.x].y + hs[threadIdx.x].y};
    double s = 0;
    for (int64_t k = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRedBlock) {
        double re = w[k].x, im = w[k].y;
        for (int c = 0; c < ncol; ++c) {
            const cx<R> v = V[k + (int64_t)c * ldv];
            re -= (double)v.x * hs[c].x - (double)v.y * hs[c].y;
            im -= (double)v.x * hs[c].y + (double)v.y * hs[c].x;
        }
        const cx<R> o = {(R)re, (R)im};
        w[k] = o;
        s += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
    }
    if (npart) {
        s = block_sum(s, sh);
        if (threadIdx.x == 0) npart[blockIdx.x] = s;
    }
}

// One classical Gram-Schmidt pass with the normalization folded in.  The dot
// partials dpart[c * G + q] (k_dots_partial) cover columns c = 0..ncol, column
// ncol being w itself, so they hold h = V[:, 0:ncol]^H w and ||w||^2.  Every
// workgroup reduces them in one fixed order and forms beta = sqrt(||w||^2 -
// |h|^2) (w - V h and V h are orthogonal parts of w); v = (w - V h) / beta is
// written over w (0 when beta is 0: an exact breakdown).  Workgroup 0 writes h
// to hcol and beta to *nrm.  write_v = 0 runs the reductions only
// (the last step of an inner solve, whose vector is not used).
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_gs_update_scale_r(cx<R> *w, const cx<R> *__restrict__ V, int64_t ldv,
                                                                 int ncol, const cx<double> *__restrict__ dpart, int G,
                                                                 cx<double> *hcol, double *nrm, int64_t n,
                                                                 int write_v) {
    __shared__ cx<double> hs[kFusedDotCols + 1];
    __shared__ double beta_s;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int c = wid; c <= ncol; c += kRedBlock / 64) {
        const double re = wave_total(&dpart[(int64_t)c * G].x, G, 2);
        const double im = wave_total(&dpart[(int64_t)c * G].y, G, 2);
        if (lane == 0) hs[c] = cx<double>{re, im};
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double hh = 0;
        for (int c = 0; c < ncol; ++c) hh += hs[c].x * hs[c].x + hs[c].y * hs[c].y;
        const double b2 = hs[ncol].x - hh;
        beta_s = b2 > 0 ? sqrt(b2) : 0.0;
        if (blockIdx.x == 0) *nrm = beta_s;
    }
    if (blockIdx.x == 0 && (int)threadIdx.x < ncol) hcol[threadIdx.x] = hs[threadIdx.x];
    __syncthreads();
    if (!write_v) return;
    const double beta = beta_s;
    for (int64_t k = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRedBlock) {
        double re = w[k].x, im = w[k].y;
        for (int c = 0; c < ncol; ++c) {
            const cx<R> v = V[k + (int64_t)c * ldv];
            re -= (double)v.x * hs[c].x - (double)v.y * hs[c].y;
            im -= (double)v.x * hs[c].y + (double)v.y * hs[c].x;
        }
        w[k] = beta == 0.0 ? cx<R>{0, 0} : cx<R>{(R)(re / beta), (R)(im / beta)};
    }
}

// dst = src / ||src|| with ||src||^2 the reduction of npart[0:G]; workgroup 0
// writes ||src|| to *nrm.  A zero norm (exact breakdown) gives dst = 0.
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_scale_norm_r(cx<R> *dst, const cx<R> *src, int64_t n,
                                                            const double *__restrict__ npart, int G,
                                                            double *__restrict__ nrm) {
    __shared__ double d;
    if (threadIdx.x < 64) {
        const double t = wave_total(npart, G);
        if (threadIdx.x == 0) {
            d = sqrt(t);
            if (blockIdx.x == 0) *nrm = d;
        }
    }
    __syncthreads();
    const double dd = d;
    for (int64_t k = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRedBlock)
        dst[k] = dd == 0.0 ? cx<R>{0, 0} : cx<R>{(R)(src[k].x / dd), (R)(src[k].y / dd)};
}

// x += V[:, 0:ncol] c   (or, with beta_scale, x = beta * V c when overwrite)
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_gemv(cx<R> *__restrict__ x, const cx<R> *__restrict__ V, int64_t ldv,
                                                    int ncol, const cx<double> *__restrict__ c, int64_t n,
                                                    double beta, int overwrite) {
    __shared__ cx<double> cs[kColChunk];
    for (int64_t k0 = (int64_t)blockIdx.x * kRedBlock; k0 < n; k0 += (int64_t)gridDim.x * kRedBlock) {
        const int64_t k = k0 + threadIdx.x;
        double re = 0, im = 0;
        for (int j0 = 0; j0 < ncol; j0 += kColChunk) {
            const int jn = min(kColChunk, ncol - j0);
            __syncthreads();
            for (int j = threadIdx.x; j < jn; j += kRedBlock) cs[j] = c[j0 + j];
            __syncthreads();
            if (k < n)
                for (int j = 0; j < jn; ++j) {
                    cx<R> v = V[k + (int64_t)(j0 + j) * ldv];
                    re += (double)v.x * cs[j].x - (double)v.y * cs[j].y;
                    im += (double)v.x * cs[j].y + (double)v.y * cs[j].x;
                }
        }
        if (k < n) {
            if (overwrite) x[k] = cx<R>{(R)(beta * re), (R)(beta * im)};
            else x[k] = cx<R>{(R)(x[k].x + re), (R)(x[k].y + im)};
        }
    }
}

// partial sums of |w|^2 after scaling by 1/amax is not needed for unit-scale
// Krylov vectors; vectors here stay far from the double range limits.
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_norm_partial(const cx<R> *__restrict__ w, int64_t n,
                                                            double *__restrict__ partial) {
    __shared__ double sh[kRedBlock / 64];
    double s = 0;
    for (int64_t k = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRedBlock) {
        double a = w[k].x, b = w[k].y;
        s += a * a + b * b;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedBlock) void k_reduce_norm(const double *__restrict__ partial, int nb,
                                                           double *__restrict__ out) {
    __shared__ double sh[kRedBlock / 64];
    double s = 0;
    for (int k = threadIdx.x; k < nb; k += kRedBlock) s += partial[k];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = sqrt(s);
}

// dst = src / *denom; a zero denominator (exact breakdown) gives dst = 0
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_scale_div(cx<R> *dst, const cx<R> *src, int64_t n,
                                                         const double *__restrict__ denom) {
    const double d = *denom;
    for (int64_t k = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRedBlock)
        dst[k] = d == 0.0 ? cx<R>{0, 0} : cx<R>{(R)(src[k].x / d), (R)(src[k].y / d)};
}

// r = b - w
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_sub(cx<R> *__restrict__ r, const cx<R> *__restrict__ b,
                                                   const cx<R> *__restrict__ w, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRedBlock)
        r[k] = cx<R>{b[k].x - w[k].x, b[k].y - w[k].y};
}

// y += a x
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_axpy(cx<R> *__restrict__ y, const cx<R> *__restrict__ x, int64_t n,
                                                    double are, double aim) {
    for (int64_t k = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRedBlock) {
        const cx<R> v = x[k];
        y[k] = cx<R>{(R)(y[k].x + are * v.x - aim * v.y), (R)(y[k].y + are * v.y + aim * v.x)};
    }
}

// Fixed-step RK4 stage updates (src/evolution.jl:75-118): with k1 the latest
// derivative, k2 the weighted stage accumulator and tmp the next stage state,
//   first: k2 = k1;         tmp = x + c k1
//   mid:   k2 = k2 + 2 k1;  tmp = x + c k1
//   last:  k2 = k2 + k1;    x = x + c k2
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_rk4_stage(int stage, cx<R> *__restrict__ x, cx<R> *__restrict__ tmp,
                                                         cx<R> *__restrict__ k2, const cx<R> *__restrict__ k1,
                                                         int64_t n, double c) {
    for (int64_t k = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRedBlock) {
        const cx<R> a = k1[k];
        if (stage == 0) {
            k2[k] = a;
            tmp[k] = cx<R>{(R)(x[k].x + c * a.x), (R)(x[k].y + c * a.y)};
        } else if (stage == 1) {
            const cx<R> s = k2[k];
            k2[k] = cx<R>{s.x + 2 * a.x, s.y + 2 * a.y};
            tmp[k] = cx<R>{(R)(x[k].x + c * a.x), (R)(x[k].y + c * a.y)};
        } else {
            const cx<R> s = {k2[k].x + a.x, k2[k].y + a.y};
            x[k] = cx<R>{(R)(x[k].x + c * s.x), (R)(x[k].y + c * s.y)};
        }
    }
}

// Trace part of the bordered stationary operator: y += (t . z) v, with t, v
// supported on the block diagonals (src/krylov.jl:855-860).
template <typename R>
__global__ __launch_bounds__(1024) void k_trace_axpy(cx<R> *__restrict__ y, const cx<R> *__restrict__ z,
                                                     const int64_t *__restrict__ idx, const double *__restrict__ t,
                                                     const double *__restrict__ v, int64_t nt) {
    // one workgroup, latency-bound: each thread issues kTraceUnroll index
    // loads, then their gathers, before it waits (nt = sum of the sector
    // sizes, 3721 at N=120: one round)
    constexpr int kTraceUnroll = 4;
    constexpr int64_t kSpan = 1024 * kTraceUnroll;
    __shared__ double sh[1024 / 64];
    double re = 0, im = 0;
    for (int64_t k0 = 0; k0 < nt; k0 += kSpan) {
        int64_t ix[kTraceUnroll];
        double tv[kTraceUnroll];
        cx<R> zz[kTraceUnroll];
#pragma unroll
        for (int u = 0; u < kTraceUnroll; ++u) {
            const int64_t k = k0 + u * 1024 + threadIdx.x;
            ix[u] = k < nt ? idx[k] : -1;
            tv[u] = k < nt ? t[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kTraceUnroll; ++u) zz[u] = ix[u] >= 0 ? z[ix[u]] : cx<R>{0, 0};
#pragma unroll
        for (int u = 0; u < kTraceUnroll; ++u) {
            re += tv[u] * (double)zz[u].x;
            im += tv[u] * (double)zz[u].y;
        }
    }
    re = block_sum(re, sh);
    __syncthreads();
    im = block_sum(im, sh);
    for (int64_t k0 = 0; k0 < nt; k0 += kSpan) {
        int64_t ix[kTraceUnroll];
        double vv[kTraceUnroll];
        cx<R> yy[kTraceUnroll];
#pragma unroll
        for (int u = 0; u < kTraceUnroll; ++u) {
            const int64_t k = k0 + u * 1024 + threadIdx.x;
            ix[u] = k < nt ? idx[k] : -1;
            vv[u] = k < nt ? v[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kTraceUnroll; ++u) yy[u] = ix[u] >= 0 ? y[ix[u]] : cx<R>{0, 0};
#pragma unroll
        for (int u = 0; u < kTraceUnroll; ++u)
            if (ix[u] >= 0) y[ix[u]] = cx<R>{(R)(yy[u].x + re * vv[u]), (R)(yy[u].y + im * vv[u])};
    }
}

// zero a buffer of complex doubles (a kernel, so that it can live inside a
// captured graph, where memset nodes are avoided)
__global__ void k_zero_c64(cx<double> *__restrict__ p, int64_t count) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x)
        p[k] = cx<double>{0, 0};
}

// min |beta e1 - H y| for the inner solver's (k+1) x k upper Hessenberg H,
// H(i, j) at Hin[i + j (k+1)] for i <= j, subdiagonal in nrm[0..k-1], beta in
// nrm[k].  One lane; k <= 32.  A zero subdiagonal ends the space early.
__global__ void k_inner_lsq(const cx<double> *__restrict__ Hin, const double *__restrict__ nrm, int k,
                            cx<double> *__restrict__ y) {
    // One thread solves the small problem; its arrays live in LDS (dynamically
    // indexed private arrays would go to scratch memory, ~100x the latency).
    constexpr int KMAX = 32;
    __shared__ double Rr[(KMAX + 1) * KMAX / 2 + KMAX], Ri[(KMAX + 1) * KMAX / 2 + KMAX];  // packed upper triangle
    __shared__ double gr[KMAX + 1], gi[KMAX + 1];
    __shared__ double csv[KMAX], snr[KMAX], sni[KMAX];
    __shared__ cx<double> ys[KMAX];
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    auto idx = [](int i, int j) { return j * (j + 1) / 2 + i; };  // i <= j
    int kk = k;
    for (int j = 0; j < k; ++j)
        if (nrm[j] == 0.0) { kk = j + 1; break; }
    for (int i = 0; i <= kk; ++i) { gr[i] = 0; gi[i] = 0; }
    gr[0] = nrm[k];
    // rotate column by column
    for (int j = 0; j < kk; ++j) {
        // load column j (rows 0..j) and subdiagonal
        for (int i = 0; i <= j; ++i) { Rr[idx(i, j)] = Hin[i + (int64_t)j * (k + 1)].x; Ri[idx(i, j)] = Hin[i + (int64_t)j * (k + 1)].y; }
        double sub_r = nrm[j], sub_i = 0.0;
        for (int i = 0; i < j; ++i) {  // previous rotations
            double h1r = Rr[idx(i, j)], h1i = Ri[idx(i, j)];
            double h2r = Rr[idx(i + 1, j)], h2i = Ri[idx(i + 1, j)];
            double c = csv[i], sr = snr[i], si = sni[i];
            // h1' = c h1 + s h2 ; h2' = -conj(s) h1 + c h2
            Rr[idx(i, j)] = c * h1r + (sr * h2r - si * h2i);
            Ri[idx(i, j)] = c * h1i + (sr * h2i + si * h2r);
            Rr[idx(i + 1, j)] = -(sr * h1r + si * h1i) + c * h2r;
            Ri[idx(i + 1, j)] = -(sr * h1i - si * h1r) + c * h2i;
        }
        // new rotation zeroing the subdiagonal
        double ar = Rr[idx(j, j)], ai = Ri[idx(j, j)];
        double absa = sqrt(ar * ar + ai * ai), absb = sqrt(sub_r * sub_r + sub_i * sub_i);
        double c, sr, si;
        if (absb == 0) { c = 1; sr = 0; si = 0; }
        else if (absa == 0) { c = 0; sr = 1; si = 0; }
        else {
            double scale = absa + absb;
            double nrmab = scale * sqrt((absa / scale) * (absa / scale) + (absb / scale) * (absb / scale));
            double alr = ar / absa, ali = ai / absa;
            c = absa / nrmab;
            // s = alpha * conj(b) / nrmab
            sr = (alr * sub_r + ali * sub_i) / nrmab;
            si = (ali * sub_r - alr * sub_i) / nrmab;
        }
        csv[j] = c; snr[j] = sr; sni[j] = si;
        Rr[idx(j, j)] = c * ar + (sr * sub_r - si * sub_i);
        Ri[idx(j, j)] = c * ai + (sr * sub_i + si * sub_r);
        double g1r = gr[j], g1i = gi[j];
        gr[j] = c * g1r;
        gi[j] = c * g1i;
        gr[j + 1] = -(sr * g1r + si * g1i);
        gi[j + 1] = -(sr * g1i - si * g1r);
    }
    for (int i = kk - 1; i >= 0; --i) {
        double accr = gr[i], acci = gi[i];
        for (int jj = i + 1; jj < kk; ++jj) {
            double hr = Rr[idx(i, jj)], hi = Ri[idx(i, jj)];
            accr -= hr * ys[jj].x - hi * ys[jj].y;
            acci -= hr * ys[jj].y + hi * ys[jj].x;
        }
        double dr = Rr[idx(i, i)], di = Ri[idx(i, i)];
        double den = dr * dr + di * di;
        ys[i] = den == 0 ? cx<double>{0, 0}
                         : cx<double>{(accr * dr + acci * di) / den, (acci * dr - accr * di) / den};
    }
    for (int i = 0; i < k; ++i) y[i] = i < kk ? ys[i] : cx<double>{0, 0};
}

// min |beta e1 - H y| for a (k+1) x k upper Hessenberg H (k <= K <= 8) by
// complex Givens rotations, on one wave: lane c < k holds column c of H, lane
// k the right-hand side, each in registers indexed by row (K is a compile-time
// bound and the row loops are unrolled, so no index is dynamic).  Rotation j
// is formed by lane j and broadcast; every column to its right and the
// right-hand side apply it to their rows j, j+1 (the same arithmetic, in the
// same order per column, as applying rotations 0..c-1 to column c when it is
// reached, as k_inner_lsq does).  Back substitution: y_i from lane i, the sum
// of R(i, l) y_l over l > i reduced across lanes 0..7.  Hc(i, j) gives H(i, j)
// for i <= j, sub(j) the real subdiagonal H(j+1, j); a zero subdiagonal ends
// the space early.  Every lane of the wave must call it; lanes < k write ys.
template <int K, typename HF, typename SF>
__device__ __forceinline__ void hess_lsq_wave(int k, double beta, HF Hc, SF sub, cx<double> *ys) {
    static_assert(K <= 8, "the back substitution reduces across 8 lanes");
    const int lane = threadIdx.x & 63;
    double hr[K + 1], hi[K + 1];
#pragma unroll
    for (int r = 0; r <= K; ++r) { hr[r] = 0; hi[r] = 0; }
    int kk = k;
    for (int j = k - 1; j >= 0; --j)
        if (sub(j) == 0.0) kk = j + 1;
    if (lane < k) {
#pragma unroll
        for (int r = 0; r <= K; ++r) {
            if (r <= lane) {
                const cx<double> h = Hc(r, lane);
                hr[r] = h.x;
                hi[r] = h.y;
            } else if (r == lane + 1) {
                hr[r] = sub(lane);
            }
        }
    } else if (lane == k) {
        hr[0] = beta;
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j >= kk) break;
        double c = 0, sr = 0, si = 0;
        if (lane == j) {
            const double ar = hr[j], ai = hi[j], sub_r = hr[j + 1], sub_i = 0.0;
            const double absa = sqrt(ar * ar + ai * ai), absb = sqrt(sub_r * sub_r + sub_i * sub_i);
            if (absb == 0) { c = 1; sr = 0; si = 0; }
            else if (absa == 0) { c = 0; sr = 1; si = 0; }
            else {
                const double scale = absa + absb;
                const double nrmab = scale * sqrt((absa / scale) * (absa / scale) + (absb / scale) * (absb / scale));
                const double alr = ar / absa, ali = ai / absa;
                c = absa / nrmab;
                sr = (alr * sub_r + ali * sub_i) / nrmab;
                si = (ali * sub_r - alr * sub_i) / nrmab;
            }
            hr[j] = c * ar + (sr * sub_r - si * sub_i);
            hi[j] = c * ai + (sr * sub_i + si * sub_r);
            hr[j + 1] = 0;
            hi[j + 1] = 0;
        }
        c = __shfl(c, j);
        sr = __shfl(sr, j);
        si = __shfl(si, j);
        if (lane > j && lane <= k) {
            const double h1r = hr[j], h1i = hi[j], h2r = hr[j + 1], h2i = hi[j + 1];
            hr[j] = c * h1r + (sr * h2r - si * h2i);
            hi[j] = c * h1i + (sr * h2i + si * h2r);
            hr[j + 1] = -(sr * h1r + si * h1i) + c * h2r;
            hi[j + 1] = -(sr * h1i - si * h1r) + c * h2i;
        }
    }
    double yr = 0, yi = 0;
#pragma unroll
    for (int i = K - 1; i >= 0; --i) {
        if (i >= kk) continue;
        double pr = 0, pim = 0;
        if (lane > i && lane < kk) {
            pr = hr[i] * yr - hi[i] * yi;
            pim = hr[i] * yi + hi[i] * yr;
        }
#pragma unroll
        for (int m = 1; m < 8; m <<= 1) {
            pr += __shfl_xor(pr, m);
            pim += __shfl_xor(pim, m);
        }
        const double gir = __shfl(hr[i], k), gii = __shfl(hi[i], k);
        if (lane == i) {
            const double accr = gir - pr, acci = gii - pim;
            const double dr = hr[i], di = hi[i];
            const double den = dr * dr + di * di;
            if (den != 0) {
                yr = (accr * dr + acci * di) / den;
                yi = (acci * dr - accr * di) / den;
            }
        }
    }
    if (lane < k) ys[lane] = cx<double>{yr, yi};
}

// The end of an inner solve of k <= kLsqRegs steps in one kernel: every
// workgroup reduces the last step's dot partials (columns 0..k, column k the
// candidate w itself: the Hessenberg column and ||w||^2, as
// k_gs_update_scale_r does), solves the least squares (wave 0, registers),
// and writes its share of z = Z y in FP64.  Replaces the last step's
// normalization, the least-squares launch and the combination launch.
constexpr int kLsqRegs = 8;
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_inner_finish(const cx<double> *__restrict__ dpart, int G,
                                                            const cx<double> *__restrict__ Hin,
                                                            const double *__restrict__ nrm, int k,
                                                            const cx<R> *__restrict__ Z, int64_t n,
                                                            cx<double> *__restrict__ zout) {
    __shared__ cx<double> hl[kLsqRegs + 1];
    __shared__ cx<double> ys[kLsqRegs];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int c = wid; c <= k; c += kRedBlock / 64) {
        const double re = wave_total(&dpart[(int64_t)c * G].x, G, 2);
        const double im = wave_total(&dpart[(int64_t)c * G].y, G, 2);
        if (lane == 0) hl[c] = cx<double>{re, im};
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // wave 0
        double hh = 0;
        for (int c = 0; c < k; ++c) hh += hl[c].x * hl[c].x + hl[c].y * hl[c].y;
        const double b2 = hl[k].x - hh;
        const double beta_last = b2 > 0 ? sqrt(b2) : 0.0;
        hess_lsq_wave<kLsqRegs>(
            k, nrm[k],
            [Hin, k](int i, int j) { return j < k - 1 ? Hin[i + (int64_t)j * (k + 1)] : hl[i]; },
            [nrm, k, beta_last](int j) { return j < k - 1 ? nrm[j] : beta_last; }, ys);
    }
    __syncthreads();
    cx<double> y[kLsqRegs];
#pragma unroll
    for (int u = 0; u < kLsqRegs; ++u) y[u] = ys[u];
    for (int64_t q = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kRedBlock) {
        cx<R> v[kLsqRegs];
#pragma unroll
        for (int u = 0; u < kLsqRegs; ++u) v[u] = u < k ? Z[q + (int64_t)u * n] : cx<R>{0, 0};
        double re = 0, im = 0;
#pragma unroll
        for (int u = 0; u < kLsqRegs; ++u) {
            if (u >= k) break;
            re += (double)v[u].x * y[u].x - (double)v[u].y * y[u].y;
            im += (double)v[u].x * y[u].y + (double)v[u].y * y[u].x;
        }
        zout[q] = cx<double>{re, im};
    }
}

// scatter a sparse vector into a zeroed dense one
template <typename R>
__global__ void k_scatter(cx<R> *__restrict__ y, const int64_t *__restrict__ idx, const double *__restrict__ v,
                          int64_t nt) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nt) y[idx[k]] = cx<R>{(R)v[k], 0};
}

template <typename Rd, typename Rs>
__global__ void k_convert(cx<Rd> *__restrict__ dst, const cx<Rs> *__restrict__ src, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        dst[k] = cx<Rd>{(Rd)src[k].x, (Rd)src[k].y};
}

// ---------------------------------------------------------------------------
// host objects
// ---------------------------------------------------------------------------
}  // namespace pimi

struct pimi_workspace {
    pimi_plan *plan = nullptr;
    int prec = PIMI_C64;
    int m = 30;
    hipStream_t stream = nullptr;
    std::atomic<int> busy{0};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Arnoldi lookahead: two pinned slots of (m + 2) entries, each holding one
    // step's Gram-Schmidt column and, after it, the new basis vector's norm;
    // an event per slot marks its copies done
    pimi::cx<double> *hlook = nullptr;
    hipEvent_t look_ev[2] = {nullptr, nullptr};
    double last_apply_ms = 0.0;
    // double-precision Krylov storage
    pimi::DevBuf<pimi::cx<double>> V, w, r, x, b, trial, cur, tmp, obs;
    pimi::DevBuf<pimi::cx<float>> V32, w32, x32, y32;
    pimi::DevBuf<pimi::cx<double>> partial, hvec, hacc, coef;
    // Krylov exponential Arnoldi (delayed reorthogonalization): reduction
    // partials, per-step mail regions, update coefficients, per-step
    // descriptors and the pinned read-back of the regions
    pimi::DevBuf<pimi::cx<double>> ex_partial, ex_mail, ex_h;
    pimi::DevBuf<char> ex_desc;
    pimi::cx<double> *ex_host = nullptr;
    int64_t ex_stride = 0;
    bool ex_ready = false;
    pimi::DevBuf<double> npartial, nscal, d_vval;
    pimi::DevBuf<pimi::cx<double>> zbuf, pt1, pt2;  // preconditioner work vectors
    // flexible GMRES: outer search directions Z and the inner solver's vectors
    pimi::DevBuf<pimi::cx<double>> Zo, Hin;
    // flexible GMRES outer loop, delayed reorthogonalization: reduction
    // partials, coefficient mailbox, update coefficients, scalars, per-step
    // descriptors, and two pinned read-back slots
    pimi::DevBuf<pimi::cx<double>> dc_partial, dc_mail, dc_h;
    pimi::DevBuf<double> dc_slot;
    pimi::DevBuf<char> dc_desc;
    pimi::cx<double> *dc_look = nullptr;
    bool dc_ready = false;
    pimi::DevBuf<char> inner_mem;  // (kin+1) basis + 4 work vectors in the inner precision
    pimi::DevBuf<double> inner_norms;
    pimi::DevBuf<pimi::cx<double>> inner_io, inner_y;  // fixed input/output vectors and LSQ solution
    int inner_k = 0;
    hipGraphExec_t inner_graph = nullptr;  // captured inner solve (one per (kin, precision))
    int inner_graph_k = 0, inner_graph_prec = -1;
    double inner_graph_scale = 0.0;
    int inner_graph_width = -1;  // banded tile width index the graph was captured with
    bool use_graphs = false;  // PIMI_INNER_GRAPH=1: the inner solve as a captured HIP graph
    // the inner solve's last step, least squares and combination in one launch
    // (PIMI_INNER_FINISH=0: three launches, k_inner_lsq on one lane)
    bool inner_finish = true;
    double *hpin = nullptr;  // pinned host staging (small results)
    bool trace_progress = false;  // PIMI_TRACE=1: solver progress lines on stderr
    bool weighted = false;        // current stationary solve iterates on u = S x
    // conditional-trajectory buffers (n x capacity each) and per-column scalars
    int64_t traj_cap = 0;
    pimi::DevBuf<pimi::cx<double>> tX, tY, tTmp, tK1, tK2, tPartial, tH, tZ, tLam;
};

namespace pimi {

// banded-apply tile width index: from an environment override (4, 8 or 16
// columns) or the default
static int dia_width_default(const char *env, int dflt) {
    const char *v = std::getenv(env);
    if (!v) return dflt;
    const int tb = std::atoi(v);
    for (int w = 0; w < kNumDiaWidths; ++w)
        if (kDiaWidths[w] == tb) return w;
    return dflt;
}

static void finalize_plan(pimi_plan *P) {
    const FusedPlan &F = P->host;
    build_stencil(F, P->st);
    P->n = F.n;
    P->ns = F.ns;
    std::vector<int64_t> offs(F.offs);
    offs.push_back(F.n);
    std::vector<int32_t> dims(F.ns);
    for (int s = 0; s < F.ns; ++s) dims[s] = (int32_t)F.dims[s];
    P->ntiles = (F.n + kApplyBlock - 1) / kApplyBlock;
    P->tiles_per_xcd = (P->ntiles + kNumXcd - 1) / kNumXcd;
    std::vector<int32_t> tsec((size_t)P->ntiles);
    int s = 0;
    for (int64_t t = 0; t < P->ntiles; ++t) {
        int64_t i = t * kApplyBlock;
        while (i >= offs[s + 1]) ++s;
        tsec[(size_t)t] = s;
    }
    // apply tiles: sector-major, then column groups, then 64-row chunks
    auto make_tiles = [&](int cols, int rows) {
        std::vector<ApplyTile> t;
        for (int q = 0; q < F.ns; ++q) {
            const int32_t m = (int32_t)F.dims[q];
            for (int32_t b0 = 0; b0 < m; b0 += cols)
                for (int32_t a0 = 0; a0 < m; a0 += rows)
                    t.push_back(ApplyTile{q, b0, a0, std::min<int32_t>(cols, m - b0)});
        }
        return t;
    };
    std::vector<ApplyTile> at = make_tiles(1, kApplyWave), at4 = make_tiles(kApplyCols, kApplyWave);
    P->natiles = (int64_t)at.size();
    P->natiles4 = (int64_t)at4.size();
    P->atiles_per_xcd = (P->natiles + kNumXcd - 1) / kNumXcd;
    P->main.build(P->st, F.dims, F.offs);
    // trace functional (src/source_protocol.jl:15-34)
    P->tidx.clear();
    P->tval.clear();
    for (int q = 0; q < F.ns; ++q) {
        int64_t m = F.dims[q];
        double sc = F.geom.sqrt_f.empty() ? 1.0 : F.geom.sqrt_f[q];
        for (int64_t a = 0; a < m; ++a) {
            P->tidx.push_back(F.offs[q] + a + a * m);
            P->tval.push_back(sc);
        }
    }
    if (!P->ctx) return;  // host-only plan: lowering, export, trace and observables only
    HIPCHK(hipSetDevice(P->ctx->device));
    P->offs.upload(offs.data(), offs.size());
    P->dims.upload(dims.data(), dims.size());
    P->tile_sector.upload(tsec.data(), tsec.size());
    P->atiles.upload(at.data(), at.size());
    P->atiles4.upload(at4.data(), at4.size());
    for (int w = 0; w < kNumDiaWidths; ++w) {
        std::vector<ApplyTile> dt = make_tiles(kDiaWidths[w], kDiaRows);
        P->ndtiles[w] = (int64_t)dt.size();
        P->dtiles[w].upload(dt.data(), dt.size());
    }
    P->dia_w_single[0] = P->dia_w_single[1] = dia_width_default("PIMI_DIA_TB_SINGLE", 0);
    P->dia_w_multi[0] = P->dia_w_multi[1] = dia_width_default("PIMI_DIA_TB_MULTI", 1);
    P->main.push();
    dia_setup_programs(P->main, P->ctx->device);
    P->d_tidx.upload(P->tidx.data(), P->tidx.size());
    P->d_tval.upload(P->tval.data(), P->tval.size());
}

// ---------------------------------------------------------------------------
// device operations on a workspace
// ---------------------------------------------------------------------------
static void ensure_diag_stencil(pimi_plan *P) {
    std::lock_guard<std::recursive_mutex> lk(P->lazy_mu);
    if (P->diag_ready) return;
    Stencil st;
    build_stencil(P->host, st, true);
    P->diag.upload(st, P->host.dims, P->host.offs);
    dia_setup_programs(P->diag, P->ctx->device);
    P->diag_ready = true;
}

static void ensure_adjoint_stencil(pimi_plan *P) {
    std::lock_guard<std::recursive_mutex> lk(P->lazy_mu);
    if (P->adj_ready) return;
    FusedPlan A;
    adjoint_plan(P->host, A);
    Stencil st;
    build_stencil(A, st);
    P->adj.upload(st, P->host.dims, P->host.offs);
    dia_setup_programs(P->adj, P->ctx->device);
    P->adj_ready = true;
}

enum class ApplyOp { Full, Diag, Adjoint };

// `weighted`: the sector-weighted operator S L S^{-1} (full stencil only; the
// sector-diagonal part D is unchanged by a sector-constant S)
// y = out_scale * (stencil) x for ncols vectors through the banded kernel,
// tile width kDiaWidths[w]
template <typename R>
static void launch_dia(hipStream_t stream, const pimi_plan *P, const DevStencil &S, bool weighted, const cx<R> *x,
                       cx<R> *y, double out_scale, int64_t ncols, int w) {
    const cx<R> *coef;
    if constexpr (sizeof(R) == 8) coef = weighted ? S.coefw64.p : S.coef64.p;
    else coef = weighted ? S.coefw32.p : S.coef32.p;
    // one tile per workgroup, its groups split over kDiaSplit waves
    const int64_t nt = P->ndtiles[w];
    const int64_t per_xcd = (nt + kNumXcd - 1) / kNumXcd;
    dim3 grid((unsigned)(per_xcd * kNumXcd), (unsigned)std::min<int64_t>(ncols, 65535));
    dim3 block(kApplyWave * kDiaSplit);
    if (launch_dia_program<R>(stream, S, w, x, y, P->n, P->dtiles[w].p, nt, per_xcd, P->offs.p, P->dims.p, coef,
                              (R)out_scale, ncols, grid, block))
        return;
#define PIMI_DIA_LAUNCH(TB)                                                                                   \
    hipLaunchKernelGGL((k_apply_dia_split<R, TB>), grid, block, 0, stream, x, y, P->n, P->dtiles[w].p, nt,     \
                       per_xcd, P->offs.p, P->dims.p, S.tables(), coef, (R)out_scale, ncols)
    switch (kDiaWidths[w]) {
        case 4: PIMI_DIA_LAUNCH(4); break;
        case 8: PIMI_DIA_LAUNCH(8); break;
        default: PIMI_DIA_LAUNCH(16); break;
    }
#undef PIMI_DIA_LAUNCH
    HIPCHK(hipGetLastError());
}

template <typename R>
static void launch_apply(pimi_workspace *W, const cx<R> *x, cx<R> *y, double out_scale, int64_t ncols = 1,
                         bool diag = false, bool weighted = false, bool adjoint = false) {
    pimi_plan *P = W->plan;
    const ApplyOp op = adjoint ? ApplyOp::Adjoint : (diag ? ApplyOp::Diag : ApplyOp::Full);
    const DevStencil &S = op == ApplyOp::Diag ? P->diag : (op == ApplyOp::Adjoint ? P->adj : P->main);
    if (S.dia_ok) {
        const int pi = sizeof(R) == 8 ? 0 : 1;
        launch_dia<R>(W->stream, P, S, op == ApplyOp::Full && weighted, x, y, out_scale, ncols,
                      ncols > 1 ? P->dia_w_multi[pi] : P->dia_w_single[pi]);
        return;
    }
    const ApplyEntry *entries = (op == ApplyOp::Full && weighted) ? S.entries_w.p : S.entries.p;
    const cx<R> *val;
    if constexpr (sizeof(R) == 8) val = S.val64.p;
    else val = S.val32.p;
    dim3 grid((unsigned)(P->atiles_per_xcd * kNumXcd), (unsigned)std::min<int64_t>(ncols, 65535));
    hipLaunchKernelGGL(k_apply<R>, grid, dim3(kApplyWave), 0, W->stream, x, y, P->n, P->atiles.p, P->natiles,
                       P->atiles_per_xcd, P->offs.p, P->dims.p, S.entry_base.p, entries, S.col.p, val, (R)out_scale,
                       ncols);
    HIPCHK(hipGetLastError());
}

template <typename R>
static void dev_norm(pimi_workspace *W, const cx<R> *v, int64_t n, double *dst) {
    hipLaunchKernelGGL(k_norm_partial<R>, dim3(kDotBlocks), dim3(kRedBlock), 0, W->stream, v, n, W->npartial.p);
    hipLaunchKernelGGL(k_reduce_norm, dim3(1), dim3(kRedBlock), 0, W->stream, W->npartial.p, kDotBlocks, dst);
}

// h[0:ncol] = V[:,0:ncol]^H w ; optionally hacc += h
template <typename R>
static void dev_dots(pimi_workspace *W, const cx<R> *V, int64_t ldv, int ncol, const cx<R> *w, int64_t n,
                     cx<double> *h, cx<double> *hacc) {
    hipLaunchKernelGGL(k_dots_partial<R>, dim3(kDotBlocks, ncol), dim3(kRedBlock), 0, W->stream, V, ldv, w, n,
                       W->partial.p);
    hipLaunchKernelGGL(k_reduce_cols, dim3(ncol), dim3(kRedBlock), 0, W->stream, W->partial.p, kDotBlocks, h, hacc);
}

// `passes` (two by default) passes of classical Gram-Schmidt of w against
// V[:, 0:ncol], the coefficients accumulated into hcol.  Narrow bases
// (ncol <= kFusedDotCols) take two launches per pass: the dot partials, then
// the update, which reduces them itself; after the last pass W->npartial
// holds the partials of ||w||^2 for dev_scale_normalize.  Wider bases reduce
// the dots in a launch of their own and the norm partials follow the last
// update.
template <typename R>
static void dev_cgs2(pimi_workspace *W, const cx<R> *V, int64_t ldv, int ncol, cx<R> *w, int64_t n,
                     cx<double> *hcol, int passes = 2) {
    for (int pass = 0; pass < passes; ++pass) {
        if (ncol <= kFusedDotCols) {
            hipLaunchKernelGGL(k_dots_partial<R>, dim3(kDotBlocks, ncol), dim3(kRedBlock), 0, W->stream, V, ldv, w,
                               n, W->partial.p);
            hipLaunchKernelGGL(k_gs_update_r<R>, dim3(kNormBlocks), dim3(kRedBlock), 0, W->stream, w, V, ldv, ncol,
                               W->partial.p, kDotBlocks, hcol, n, pass == passes - 1 ? W->npartial.p : nullptr);
        } else {
            dev_dots<R>(W, V, ldv, ncol, w, n, W->hvec.p, hcol);
            hipLaunchKernelGGL(k_gs_update<R>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, w, V, ldv, ncol,
                               W->hvec.p, n);
        }
    }
    if (ncol > kFusedDotCols)
        hipLaunchKernelGGL(k_norm_partial<R>, dim3(kNormBlocks), dim3(kRedBlock), 0, W->stream, w, n, W->npartial.p);
}

// dst = w / ||w|| and *nrm = ||w||, from the norm partials dev_cgs2 left
template <typename R>
static void dev_scale_normalize(pimi_workspace *W, cx<R> *dst, const cx<R> *w, int64_t n, double *nrm) {
    hipLaunchKernelGGL(k_scale_norm_r<R>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, dst, w, n,
                       W->npartial.p, kNormBlocks, nrm);
}

static void sync(pimi_workspace *W) { HIPCHK(hipStreamSynchronize(W->stream)); }

static void d2h(pimi_workspace *W, void *dst, const void *src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, W->stream));
    sync(W);
}
// Host-to-device copies here come from short-lived host buffers, so they
// complete before returning.
static void h2d(pimi_workspace *W, void *dst, const void *src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, W->stream));
    sync(W);
}

// ---------------------------------------------------------------------------
// deterministic probe vectors (splitmix64 + Box-Muller), seed 0x51ca1e
// ---------------------------------------------------------------------------
static uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double unif(uint64_t &s) { return ((splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

static void probe_vector(uint64_t &st, std::vector<cx<double>> &x) {
    for (auto &z : x) {
        double u = unif(st), v = unif(st);
        double r = std::sqrt(-2.0 * std::log(u)), a = 6.283185307179586 * v;
        z = cx<double>{r * std::cos(a) * 0.7071067811865476, r * std::sin(a) * 0.7071067811865476};
    }
}

// ---------------------------------------------------------------------------
// workspace buffers
// ---------------------------------------------------------------------------
static void ensure_workspace(pimi_workspace *W) {
    int64_t n = W->plan->n;
    int m = W->m;
    W->V.alloc((size_t)n * (m + 3));  // the Krylov exponential's Arnoldi takes up to m + 2 steps
    W->w.alloc(n);
    W->r.alloc(n);
    W->x.alloc(n);
    W->b.alloc(n);
    W->trial.alloc(n);
    W->cur.alloc(n);
    W->tmp.alloc(n);
    const int maxcols = std::max(m + 2, kMaxObservables);
    W->partial.alloc((size_t)kDotBlocks * maxcols);
    W->hvec.alloc(maxcols);
    W->hacc.alloc((size_t)(m + 2));
    W->coef.alloc(m + 2);
    W->npartial.alloc(std::max(kDotBlocks, kNormBlocks));
    W->nscal.alloc(8);
    if (W->prec == PIMI_C32) {
        W->V32.alloc((size_t)n * (m + 1));
        W->w32.alloc(n);
        W->x32.alloc(n);
        W->y32.alloc(n);
    }
}

// y = A z for the bordered stationary operator: L z / s + (t . z) v; with
// sector weighting (W->weighted) S A S^{-1}: t -> ones on the trace support,
// v -> sigma v (held in d_vval)
static void steady_apply(pimi_workspace *W, const cx<double> *z, cx<double> *y, double invscale) {
    pimi_plan *P = W->plan;
    launch_apply<double>(W, z, y, invscale, 1, false, W->weighted);
    hipLaunchKernelGGL(k_trace_axpy<double>, dim3(1), dim3(1024), 0, W->stream, y, z, P->d_tidx.p,
                       W->weighted ? P->d_ones.p : P->d_tval.p, W->d_vval.p, (int64_t)P->tidx.size());
    HIPCHK(hipGetLastError());
}

static double read_scalar(pimi_workspace *W, const double *dptr) {
    d2h(W, W->hpin, dptr, sizeof(double));
    return W->hpin[0];
}

// Arnoldi lookahead.  Step j's device work ends with the Gram-Schmidt
// coefficients in hacc[0:j+1] and the norm of the new basis vector in
// nscal[0]; both are copied to pinned slot j % 2 and the slot's event
// recorded.  Step j+1 needs only V[:, j+1], which step j leaves on the
// device, so the solvers enqueue step j+1 before waiting for slot j: the
// host's Givens update of column j overlaps the device's step j+1.  Every
// buffer the steps share is reused in stream order, after the slot's copies.
static void look_post(pimi_workspace *W, int j) {
    cx<double> *slot = W->hlook + (size_t)(j & 1) * ((size_t)W->m + 2);
    HIPCHK(hipMemcpyAsync(slot, W->hacc.p, sizeof(cx<double>) * (j + 1), hipMemcpyDeviceToHost, W->stream));
    HIPCHK(hipMemcpyAsync(slot + W->m + 1, W->nscal.p, sizeof(double), hipMemcpyDeviceToHost, W->stream));
    HIPCHK(hipEventRecord(W->look_ev[j & 1], W->stream));
}

// waits for step j's slot; returns its coefficients and the subdiagonal norm
static const cx<double> *look_wait(pimi_workspace *W, int j, double &hnext) {
    HIPCHK(hipEventSynchronize(W->look_ev[j & 1]));
    const cx<double> *slot = W->hlook + (size_t)(j & 1) * ((size_t)W->m + 2);
    hnext = *reinterpret_cast<const double *>(slot + W->m + 1);
    return slot;
}


void ensure_plan_preconditioner(pimi_plan *P) {
    std::lock_guard<std::recursive_mutex> lk(P->lazy_mu);
    if (!P->prec_ready) {
        std::vector<cd> U, inv;
        build_sector_preconditioner(P->host, U, inv);
        std::vector<cx<double>> u64(U.size()), i64(U.size());
        std::vector<cx<float>> u32(U.size()), i32(U.size());
        for (size_t k = 0; k < U.size(); ++k) {
            u64[k] = cx<double>{U[k].real(), U[k].imag()};
            i64[k] = cx<double>{inv[k].real(), inv[k].imag()};
            u32[k] = cx<float>{(float)U[k].real(), (float)U[k].imag()};
            i32[k] = cx<float>{(float)inv[k].real(), (float)inv[k].imag()};
        }
        std::vector<GemmTile> tiles;
        for (int s = 0; s < P->ns; ++s) {
            int nt = (int)((P->host.dims[s] + kGemmTile - 1) / kGemmTile);
            for (int tj = 0; tj < nt; ++tj)
                for (int ti = 0; ti < nt; ++ti) tiles.push_back(GemmTile{s, ti, tj, 0});
        }
        P->pU64.upload(u64.data(), u64.size());
        P->pInv64.upload(i64.data(), i64.size());
        P->pU32.upload(u32.data(), u32.size());
        P->pInv32.upload(i32.data(), i32.size());
        P->u_real = std::all_of(U.begin(), U.end(), [](const cd &z) { return z.imag() == 0.0; });
        if (P->u_real) {
            std::vector<double> r64(U.size());
            std::vector<float> r32(U.size());
            for (size_t k = 0; k < U.size(); ++k) {
                r64[k] = U[k].real();
                r32[k] = (float)U[k].real();
            }
            P->pUr64.upload(r64.data(), r64.size());
            P->pUr32.upload(r32.data(), r32.size());
        }
        P->gtiles.upload(tiles.data(), tiles.size());
        P->ngtiles = (int64_t)tiles.size();
        std::vector<GemmTile> mt;
        for (int s = 0; s < P->ns; ++s) {
            int nt = (int)((P->host.dims[s] + kMfmaTile - 1) / kMfmaTile);
            for (int tj = 0; tj < nt; ++tj)
                for (int ti = 0; ti < nt; ++ti) mt.push_back(GemmTile{s, ti, tj, 0});
        }
        P->mtiles.upload(mt.data(), mt.size());
        P->nmtiles = (int64_t)mt.size();
        // The single-plan launch keeps each sector on one XCD (workgroup w runs
        // on XCD w mod 8, each XCD has its own L2): sectors go to XCDs by
        // longest-processing-time on their m^3 cost, and XCD x's tiles take
        // workgroup ids x, x + 8, x + 16, ...; shorter lists are padded with
        // empty tiles (m = 0).  Each operand block then enters one L2
        // instead of all eight.
        {
            for (int s = 0; s < P->ns; ++s)  // the GEMM's buffer loads address a block with 32-bit byte offsets
                if (P->host.dims[s] > 11585)
                    throw Error(PIMI_EINVAL, "sector block too large for the preconditioner (m > 11585)");
            std::vector<int> order((size_t)P->ns);
            for (int s = 0; s < P->ns; ++s) order[(size_t)s] = s;
            auto cost = [&](int s) {
                const int64_t nt = (P->host.dims[s] + kMfmaTile - 1) / kMfmaTile;
                return nt * nt * nt;
            };
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
            std::vector<std::vector<MTile>> per(kNumXcd);
            int64_t load[kNumXcd] = {0};
            for (int s : order) {
                int x = 0;
                for (int k = 1; k < kNumXcd; ++k)
                    if (load[k] < load[x]) x = k;
                load[x] += cost(s);
                const int nt = (int)((P->host.dims[s] + kMfmaTile - 1) / kMfmaTile);
                for (int tj = 0; tj < nt; ++tj)
                    for (int ti = 0; ti < nt; ++ti)
                        per[(size_t)x].push_back(MTile{P->host.offs[(size_t)s], (int32_t)P->host.dims[s], ti, tj, 0});
            }
            size_t len = 0;
            for (auto &v : per) len = std::max(len, v.size());
            std::vector<MTile> xt(len * kNumXcd, MTile{0, 0, 0, 0, 0});
            for (int x = 0; x < kNumXcd; ++x)
                for (size_t k = 0; k < per[(size_t)x].size(); ++k) xt[k * kNumXcd + (size_t)x] = per[(size_t)x][k];
            P->xtiles.upload(xt.data(), xt.size());
            P->nxtiles = (int64_t)xt.size();
        }
        P->prec_ready = true;
    }
}

void ensure_plan_weighting(pimi_plan *P) {
    std::lock_guard<std::recursive_mutex> lk(P->lazy_mu);
    if (P->weight_ready) return;
    const FusedPlan &F = P->host;
    P->sigma.assign((size_t)F.ns, 1.0);
    if (!F.geom.sqrt_f.empty())
        for (int q = 0; q < F.ns; ++q) P->sigma[(size_t)q] = F.geom.sqrt_f[(size_t)q];
    std::vector<double> ones(P->tidx.size(), 1.0), inv((size_t)F.ns);
    for (int q = 0; q < F.ns; ++q) inv[(size_t)q] = 1.0 / P->sigma[(size_t)q];
    P->main.push_weighted(P->sigma);
    P->d_ones.upload(ones.data(), ones.size());
    P->d_inv_sigma.upload(inv.data(), inv.size());
    P->d_sigma.upload(P->sigma.data(), P->sigma.size());
    P->weight_ready = true;
}

static void ensure_preconditioner(pimi_workspace *W) {
    pimi_plan *P = W->plan;
    ensure_plan_preconditioner(P);
    if (!W->zbuf.p) {
        W->zbuf.alloc(P->n);
        W->pt1.alloc(P->n);
        W->pt2.alloc(P->n);
    }
}

// out = scale * U ((U^H r U) .* inv_den) U^H, sector by sector
template <typename R>
static void apply_precond_t(pimi_workspace *W, const cx<R> *r, cx<R> *out, double scale, cx<R> *t1, cx<R> *t2) {
    pimi_plan *P = W->plan;
    dim3 g((unsigned)P->nxtiles), b(256);  // one 16x16 tile per workgroup (split K), XCD-grouped order
    const cx<R> *nul = nullptr;
    const cx<R> *U, *Inv;
    const R *Ur;
    if constexpr (sizeof(R) == 8) { U = P->pU64.p; Inv = P->pInv64.p; Ur = P->pUr64.p; }
    else { U = P->pU32.p; Inv = P->pInv32.p; Ur = P->pUr32.p; }
    const int64_t nt = P->nxtiles;
    const MTile *xt = P->xtiles.p;
    if (P->u_real) {
        typedef cx<R> C;
        hipLaunchKernelGGL((k_sector_gemm_mfma<R, true, false, R, C>), g, b, 0, W->stream, Ur, r, t1, xt, nt, nul, (R)1);
        hipLaunchKernelGGL((k_sector_gemm_mfma<R, false, false, C, R>), g, b, 0, W->stream, t1, Ur, t2, xt, nt, Inv,
                           (R)scale);
        hipLaunchKernelGGL((k_sector_gemm_mfma<R, false, false, R, C>), g, b, 0, W->stream, Ur, t2, t1, xt, nt, nul, (R)1);
        hipLaunchKernelGGL((k_sector_gemm_mfma<R, false, true, C, R>), g, b, 0, W->stream, t1, Ur, out, xt, nt, nul, (R)1);
    } else {
        hipLaunchKernelGGL((k_sector_gemm_mfma<R, true, false>), g, b, 0, W->stream, U, r, t1, xt, nt, nul, (R)1);
        hipLaunchKernelGGL((k_sector_gemm_mfma<R, false, false>), g, b, 0, W->stream, t1, U, t2, xt, nt, Inv,
                           (R)scale);
        hipLaunchKernelGGL((k_sector_gemm_mfma<R, false, false>), g, b, 0, W->stream, U, t2, t1, xt, nt, nul, (R)1);
        hipLaunchKernelGGL((k_sector_gemm_mfma<R, false, true>), g, b, 0, W->stream, t1, U, out, xt, nt, nul, (R)1);
    }
    HIPCHK(hipGetLastError());
}

static void apply_precond(pimi_workspace *W, const cx<double> *r, cx<double> *out, double scale) {
    apply_precond_t<double>(W, r, out, scale, W->pt1.p, W->pt2.p);
}

struct GmresResult {
    bool converged;
    int64_t iterations, restarts;
    double residual;
};

// Restarted GMRES(m) on the bordered operator (src/krylov.jl:732-786); x, b on device.
// Restarted GMRES(m) on the bordered operator (src/krylov.jl:732-786); x, b on
// device.  With `precond` the sector preconditioner is applied on the right
// (x = x0 + M^{-1} V y), so the GMRES residual is the true residual.  With
// `backward` the stopping threshold is (atol + rtol |x_k|) / 2, where |x_k| of
// the current iterate is formed on the device every kXnormEvery steps: since
// t^H L = 0 and t.v = 1, the bordered residual r bounds |L x| / s <= 2 |r|,
// so the threshold guarantees |L x| / s <= atol + rtol |x|.
constexpr int kXnormEvery = 25;

static double iterate_norm(pimi_workspace *W, const std::vector<cd> &Hm, const std::vector<cd> &g, int m, int k,
                           bool precond, double invscale) {
    const int64_t n = W->plan->n;
    std::vector<cd> yv(k);
    for (int i = k - 1; i >= 0; --i) {
        cd acc = g[i];
        for (int jj = i + 1; jj < k; ++jj) acc -= Hm[(size_t)i + (size_t)jj * (m + 1)] * yv[jj];
        cd piv = Hm[(size_t)i + (size_t)i * (m + 1)];
        if (piv == cd(0, 0)) throw Error(PIMI_EINVAL, "singular GMRES Hessenberg");
        yv[i] = acc / piv;
    }
    std::vector<cx<double>> yc(k);
    for (int i = 0; i < k; ++i) yc[i] = cx<double>{yv[i].real(), yv[i].imag()};
    h2d(W, W->coef.p, yc.data(), sizeof(cx<double>) * k);
    HIPCHK(hipMemcpyAsync(W->trial.p, W->x.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, W->stream));
    if (precond) {
        hipLaunchKernelGGL(k_gemv<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->tmp.p, W->V.p, n, k,
                           W->coef.p, n, 1.0, 1);
        apply_precond(W, W->tmp.p, W->zbuf.p, 1.0 / invscale);
        hipLaunchKernelGGL(k_axpy<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->trial.p, W->zbuf.p,
                           n, 1.0, 0.0);
    } else {
        hipLaunchKernelGGL(k_gemv<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->trial.p, W->V.p, n,
                           k, W->coef.p, n, 1.0, 0);
    }
    dev_norm<double>(W, W->trial.p, n, W->nscal.p + 5);
    return read_scalar(W, W->nscal.p + 5);
}

static GmresResult gmres_device(pimi_workspace *W, double invscale, double atol, double rtol, int64_t maxiter,
                                int64_t &applications, bool precond, bool backward) {
    const int64_t n = W->plan->n;
    const int m = W->m;
    cx<double> *V = W->V.p;
    std::vector<cd> H((size_t)(m + 1) * m), g(m + 1), sn(m), yv(m);
    std::vector<double> cs(m);
    auto Hij = [&](int i, int j) -> cd & { return H[(size_t)i + (size_t)j * (m + 1)]; };
    dev_norm<double>(W, W->b.p, n, W->nscal.p);
    dev_norm<double>(W, W->x.p, n, W->nscal.p + 1);
    d2h(W, W->hpin, W->nscal.p, 2 * sizeof(double));
    const double bnorm = W->hpin[0];
    double xnorm = W->hpin[1];
    auto threshold = [&]() { return backward ? 0.5 * (atol + rtol * xnorm) : atol + rtol * bnorm; };
    int64_t iterations = 0, restarts = 0;
    steady_apply(W, W->x.p, W->w.p, invscale);
    applications++;
    hipLaunchKernelGGL(k_sub<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->r.p, W->b.p, W->w.p, n);
    dev_norm<double>(W, W->r.p, n, W->nscal.p);
    double beta = read_scalar(W, W->nscal.p);
    if (beta <= threshold()) return {true, 0, 0, beta};
    while (iterations < maxiter) {
        std::fill(H.begin(), H.end(), cd(0, 0));
        std::fill(g.begin(), g.end(), cd(0, 0));
        // V0 = r / beta
        W->hpin[0] = beta;
        h2d(W, W->nscal.p + 1, W->hpin, sizeof(double));
        hipLaunchKernelGGL(k_scale_div<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, V, W->r.p, n,
                           W->nscal.p + 1);
        g[0] = beta;
        int kdone = 0;
        int jmax = (int)std::min<int64_t>(m, maxiter - iterations);
        auto step = [&](int j) {
            const cx<double> *vin = V + (int64_t)j * n;
            if (precond) {
                apply_precond(W, vin, W->zbuf.p, 1.0 / invscale);
                vin = W->zbuf.p;
            }
            steady_apply(W, vin, W->w.p, invscale);
            HIPCHK(hipMemsetAsync(W->hacc.p, 0, sizeof(cx<double>) * (j + 1), W->stream));
            dev_cgs2<double>(W, V, n, j + 1, W->w.p, n, W->hacc.p);
            dev_scale_normalize<double>(W, V + (int64_t)(j + 1) * n, W->w.p, n, W->nscal.p);
            look_post(W, j);
        };
        if (jmax > 0) step(0);
        for (int j = 0; j < jmax; ++j) {
            if (j + 1 < jmax) step(j + 1);  // lookahead (wasted only when column j converges)
            applications++;
            double hnext = 0.0;
            const cx<double> *hc = look_wait(W, j, hnext);
            for (int i = 0; i <= j; ++i) Hij(i, j) = cd(hc[i].x, hc[i].y);
            Hij(j + 1, j) = hnext;
            for (int i = 0; i < j; ++i) {
                cd h1 = Hij(i, j), h2 = Hij(i + 1, j);
                Hij(i, j) = cs[i] * h1 + sn[i] * h2;
                Hij(i + 1, j) = -std::conj(sn[i]) * h1 + cs[i] * h2;
            }
            double c;
            cd s;
            complex_givens(Hij(j, j), Hij(j + 1, j), c, s);
            cs[j] = c;
            sn[j] = s;
            cd h1 = Hij(j, j), h2 = Hij(j + 1, j);
            Hij(j, j) = c * h1 + s * h2;
            Hij(j + 1, j) = 0;
            cd g1 = g[j], g2 = g[j + 1];
            g[j] = c * g1 + s * g2;
            g[j + 1] = -std::conj(s) * g1 + c * g2;
            iterations++;
            kdone = j + 1;
            const double est = std::abs(g[j + 1]);
            if (backward && (kdone % kXnormEvery == 0 || est <= threshold()) && Hij(j, j) != cd(0, 0))
                xnorm = iterate_norm(W, H, g, m, kdone, precond, invscale);
            if (W->trace_progress && (iterations % 50 == 0))
                std::fprintf(stderr, "[pimi] gmres iteration %lld  residual %.3e  |b| %.3e  |x| %.3e\n",
                             (long long)iterations, est, bnorm, xnorm);
            if (est <= threshold()) break;
        }
        for (int i = kdone - 1; i >= 0; --i) {
            cd acc = g[i];
            for (int jj = i + 1; jj < kdone; ++jj) acc -= Hij(i, jj) * yv[jj];
            if (Hij(i, i) == cd(0, 0)) throw Error(PIMI_EINVAL, "singular GMRES Hessenberg");
            yv[i] = acc / Hij(i, i);
        }
        std::vector<cx<double>> yc(kdone);
        for (int i = 0; i < kdone; ++i) yc[i] = cx<double>{yv[i].real(), yv[i].imag()};
        h2d(W, W->coef.p, yc.data(), sizeof(cx<double>) * kdone);
        if (precond) {
            // x += M^{-1} (V y): the preconditioner is fixed, so Z = M^{-1} V is not stored
            hipLaunchKernelGGL(k_gemv<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->tmp.p, V, n,
                               kdone, W->coef.p, n, 1.0, 1);
            apply_precond(W, W->tmp.p, W->zbuf.p, 1.0 / invscale);
            hipLaunchKernelGGL(k_axpy<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->x.p,
                               W->zbuf.p, n, 1.0, 0.0);
        } else {
            hipLaunchKernelGGL(k_gemv<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->x.p, V, n,
                               kdone, W->coef.p, n, 1.0, 0);
        }
        restarts++;
        steady_apply(W, W->x.p, W->w.p, invscale);
        applications++;
        hipLaunchKernelGGL(k_sub<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->r.p, W->b.p, W->w.p,
                           n);
        dev_norm<double>(W, W->r.p, n, W->nscal.p);
        dev_norm<double>(W, W->x.p, n, W->nscal.p + 1);
        d2h(W, W->hpin, W->nscal.p, 2 * sizeof(double));
        beta = W->hpin[0];
        xnorm = W->hpin[1];
        if (beta <= threshold()) return {true, iterations, restarts, beta};
    }
    return {false, iterations, restarts, beta};
}

// ---------------------------------------------------------------------------
// Inner-outer stationary solve.  The outer iteration is flexible GMRES on the
// bordered operator; each outer direction z_j approximately solves D z = v_j
// with `kin` steps of right-preconditioned GMRES (sector preconditioner) on
// D/s, D the sector-block-diagonal part of L.  The inner solve runs in the
// workspace precision (complex FP32 for PIMI_C32: the mixed-precision form),
// the outer iteration, its basis and all residuals in FP64.
// ---------------------------------------------------------------------------
static void ensure_inner(pimi_workspace *W, int kin) {
    const int64_t n = W->plan->n;
    const size_t es = W->prec == PIMI_C32 ? sizeof(cx<float>) : sizeof(cx<double>);
    if (W->inner_k < kin) {
        if (W->inner_graph) { HIPCHK(hipGraphExecDestroy(W->inner_graph)); W->inner_graph = nullptr; }
        // V (kin + 1 vectors), Z (kin), then four work vectors
        W->inner_mem.alloc(es * (size_t)n * (size_t)(2 * kin + 1 + 4));
        HIPCHK(hipMemsetAsync(W->inner_mem.p, 0, es * (size_t)n * (size_t)(2 * kin + 1 + 4), W->stream));
        W->Hin.alloc((size_t)(kin + 1) * kin);
        W->inner_norms.alloc((size_t)kin + 2);
        W->inner_io.alloc((size_t)2 * n);
        W->inner_y.alloc((size_t)kin);
        W->inner_k = kin;
    }
    if (!W->Zo.p) W->Zo.alloc((size_t)n * W->m);
}

// The inner solve with no host interaction: norms, Hessenberg, least squares
// and the combination all stay on the device.  Right-preconditioned GMRES on
// D/s: the directions z_i = M^-1 v_i are kept (Z), so the result M^-1 V y is
// the combination Z y and needs no preconditioner application of its own.
// Reads rin, writes zout (n entries each, FP64).
template <typename R>
static void inner_solve_dev(pimi_workspace *W, int kin, double scale, const cx<double> *rin, cx<double> *zout) {
    const int64_t n = W->plan->n;
    cx<R> *base = reinterpret_cast<cx<R> *>(W->inner_mem.p);
    cx<R> *Vi = base, *Zi = base + (size_t)n * (kin + 1), *zt = Zi + (size_t)n * kin, *wt = zt + n, *t1 = wt + n,
          *t2 = t1 + n;
    double *nrm = W->inner_norms.p;
    const dim3 gv(grid_for(n)), bv(kRedBlock);
    // v_0 = r / ||r||: norm partials, then the scaling kernel reduces them itself
    if constexpr (sizeof(R) == 8) {
        hipLaunchKernelGGL(k_norm_partial<double>, dim3(kNormBlocks), bv, 0, W->stream, rin, n, W->npartial.p);
        dev_scale_normalize<double>(W, Vi, rin, n, nrm + kin);
    } else {
        hipLaunchKernelGGL((k_convert<R, double>), dim3(1024), dim3(256), 0, W->stream, Vi, rin, n);
        hipLaunchKernelGGL(k_norm_partial<R>, dim3(kNormBlocks), bv, 0, W->stream, Vi, n, W->npartial.p);
        dev_scale_normalize<R>(W, Vi, Vi, n, nrm + kin);
    }
    // the fused steps (i < kFusedDotCols) write their Hessenberg column's rows
    // 0..i (all that the least squares reads); wider steps accumulate
    if (kin > kFusedDotCols)
        hipLaunchKernelGGL(k_zero_c64, dim3(1), dim3(256), 0, W->stream, W->Hin.p, (int64_t)(kin + 1) * kin);
    for (int i = 0; i < kin; ++i) {
        cx<R> *zi = Zi + (size_t)i * n, *wi = Vi + (size_t)(i + 1) * n;
        cx<double> *hcol = W->Hin.p + (size_t)i * (kin + 1);
        apply_precond_t<R>(W, Vi + (size_t)i * n, zi, scale, t1, t2);
        launch_apply<R>(W, zi, wi, 1.0 / scale, 1, true);
        // one Gram-Schmidt pass: the inner solve only preconditions the
        // flexible outer iteration, and at <= 32 vectors one classical pass
        // keeps the inner basis orthogonal to far below its tolerance.  The
        // candidate w sits in the basis slot it will occupy, so one reduction
        // pass gives V^H w and ||w||^2 and one kernel updates and normalizes.
        const bool last = i + 1 == kin;
        if (i + 1 <= kFusedDotCols) {
            hipLaunchKernelGGL(k_dots_partial<R>, dim3(kDotBlocks, i + 2), bv, 0, W->stream, Vi, n, wi, n,
                               W->partial.p);
            if (last && kin <= kLsqRegs && W->inner_finish) {
                // the last normalization, the least squares and z = Z y in one launch
                hipLaunchKernelGGL(k_inner_finish<R>, gv, bv, 0, W->stream, W->partial.p, kDotBlocks, W->Hin.p, nrm,
                                   kin, Zi, n, zout);
                HIPCHK(hipGetLastError());
                return;
            }
            hipLaunchKernelGGL(k_gs_update_scale_r<R>, dim3(last ? 1 : kNormBlocks), bv, 0, W->stream, wi, Vi, n,
                               i + 1, W->partial.p, kDotBlocks, hcol, nrm + i, n, last ? 0 : 1);
        } else {
            dev_cgs2<R>(W, Vi, n, i + 1, wi, n, hcol, 1);
            dev_scale_normalize<R>(W, wi, wi, n, nrm + i);
        }
    }
    hipLaunchKernelGGL(k_inner_lsq, dim3(1), dim3(64), 0, W->stream, W->Hin.p, nrm, kin, W->inner_y.p);
    if constexpr (sizeof(R) == 8) {
        hipLaunchKernelGGL(k_gemv<R>, gv, bv, 0, W->stream, zout, Zi, n, kin, W->inner_y.p, n, 1.0, 1);
    } else {
        hipLaunchKernelGGL(k_gemv<R>, gv, bv, 0, W->stream, zt, Zi, n, kin, W->inner_y.p, n, 1.0, 1);
        hipLaunchKernelGGL((k_convert<double, R>), dim3(1024), dim3(256), 0, W->stream, zout, zt, n);
    }
    HIPCHK(hipGetLastError());
}

// z = inner(r).  Eager launches (the default) read r and write z in place; the
// captured-graph form (PIMI_INNER_GRAPH=1) runs on the fixed pair inner_io and
// copies r in and z out around the replay.
static void run_inner(pimi_workspace *W, int kin, double scale, const cx<double> *r, cx<double> *z) {
    auto body = [&](const cx<double> *rin, cx<double> *zout) {
        if (W->prec == PIMI_C32) inner_solve_dev<float>(W, kin, scale, rin, zout);
        else inner_solve_dev<double>(W, kin, scale, rin, zout);
    };
    if (!W->use_graphs) { body(r, z); return; }
    const int64_t n = W->plan->n;
    // the key covers everything the captured launches bake in: the inner
    // Krylov dimension, the precision, the scale and the banded apply's tile
    // width (which the autotuner may change after a capture)
    const int width = W->plan->dia_w_single[W->prec == PIMI_C32 ? 1 : 0];
    if (!W->inner_graph || W->inner_graph_k != kin || W->inner_graph_prec != W->prec ||
        W->inner_graph_scale != scale || W->inner_graph_width != width) {
        if (W->inner_graph) { HIPCHK(hipGraphExecDestroy(W->inner_graph)); W->inner_graph = nullptr; }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(W->stream, hipStreamCaptureModeThreadLocal));
        try {
            body(W->inner_io.p, W->inner_io.p + n);
        } catch (...) {
            (void)hipStreamEndCapture(W->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHK(hipStreamEndCapture(W->stream, &g));
        hipError_t e = hipGraphInstantiate(&W->inner_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
        W->inner_graph_k = kin;
        W->inner_graph_prec = W->prec;
        W->inner_graph_scale = scale;
        W->inner_graph_width = width;
    }
    HIPCHK(hipMemcpyAsync(W->inner_io.p, r, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, W->stream));
    HIPCHK(hipGraphLaunch(W->inner_graph, W->stream));
    HIPCHK(hipMemcpyAsync(z, W->inner_io.p + n, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, W->stream));
}

static GmresResult fgmres_device_cgs2(pimi_workspace *W, double invscale, double atol, double rtol,
                                      int64_t maxiter, int64_t &applications, bool backward, int kin) {
    const int64_t n = W->plan->n;
    const int m = W->m;
    const double scale = 1.0 / invscale;
    ensure_inner(W, kin);
    cx<double> *V = W->V.p, *Z = W->Zo.p;
    auto inner = [&](const cx<double> *r, cx<double> *z) {
        run_inner(W, kin, scale, r, z);
    };
    std::vector<cd> H((size_t)(m + 1) * m), g(m + 1), sn(m), yv(m);
    std::vector<double> cs(m);
    auto Hij = [&](int i, int j) -> cd & { return H[(size_t)i + (size_t)j * (m + 1)]; };
    dev_norm<double>(W, W->b.p, n, W->nscal.p);
    dev_norm<double>(W, W->x.p, n, W->nscal.p + 1);
    d2h(W, W->hpin, W->nscal.p, 2 * sizeof(double));
    const double bnorm = W->hpin[0];
    double xnorm = W->hpin[1];
    auto threshold = [&]() { return backward ? 0.5 * (atol + rtol * xnorm) : atol + rtol * bnorm; };
    auto form_x = [&](int k, const std::vector<cd> &Hr, const std::vector<cd> &gr, cx<double> *dst, bool add_to_x) {
        std::vector<cd> yy(k);
        for (int i = k - 1; i >= 0; --i) {
            cd acc = gr[i];
            for (int jj = i + 1; jj < k; ++jj) acc -= Hr[(size_t)i + (size_t)jj * (m + 1)] * yy[jj];
            yy[i] = acc / Hr[(size_t)i + (size_t)i * (m + 1)];
        }
        std::vector<cx<double>> yc(k);
        for (int i = 0; i < k; ++i) yc[i] = cx<double>{yy[i].real(), yy[i].imag()};
        h2d(W, W->coef.p, yc.data(), sizeof(cx<double>) * k);
        if (!add_to_x) HIPCHK(hipMemcpyAsync(dst, W->x.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, W->stream));
        hipLaunchKernelGGL(k_gemv<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, dst, Z, n, k, W->coef.p,
                           n, 1.0, 0);
    };
    int64_t iterations = 0, restarts = 0;
    steady_apply(W, W->x.p, W->w.p, invscale);
    applications++;
    hipLaunchKernelGGL(k_sub<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->r.p, W->b.p, W->w.p, n);
    dev_norm<double>(W, W->r.p, n, W->nscal.p);
    double beta = read_scalar(W, W->nscal.p);
    if (beta <= threshold()) return {true, 0, 0, beta};
    while (iterations < maxiter) {
        std::fill(H.begin(), H.end(), cd(0, 0));
        std::fill(g.begin(), g.end(), cd(0, 0));
        W->hpin[0] = beta;
        h2d(W, W->nscal.p + 1, W->hpin, sizeof(double));
        hipLaunchKernelGGL(k_scale_div<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, V, W->r.p, n,
                           W->nscal.p + 1);
        g[0] = beta;
        int kdone = 0;
        int jmax = (int)std::min<int64_t>(m, maxiter - iterations);
        auto step = [&](int j) {
            inner(V + (int64_t)j * n, Z + (int64_t)j * n);
            steady_apply(W, Z + (int64_t)j * n, W->w.p, invscale);
            HIPCHK(hipMemsetAsync(W->hacc.p, 0, sizeof(cx<double>) * (j + 1), W->stream));
            dev_cgs2<double>(W, V, n, j + 1, W->w.p, n, W->hacc.p);
            dev_scale_normalize<double>(W, V + (int64_t)(j + 1) * n, W->w.p, n, W->nscal.p);
            look_post(W, j);
        };
        if (jmax > 0) step(0);
        for (int j = 0; j < jmax; ++j) {
            if (j + 1 < jmax) step(j + 1);  // lookahead (wasted only when column j converges)
            applications += 1 + kin;
            double hnext = 0.0;
            const cx<double> *hc = look_wait(W, j, hnext);
            for (int i = 0; i <= j; ++i) Hij(i, j) = cd(hc[i].x, hc[i].y);
            Hij(j + 1, j) = hnext;
            for (int i = 0; i < j; ++i) {
                cd h1 = Hij(i, j), h2 = Hij(i + 1, j);
                Hij(i, j) = cs[i] * h1 + sn[i] * h2;
                Hij(i + 1, j) = -std::conj(sn[i]) * h1 + cs[i] * h2;
            }
            double c;
            cd s;
            complex_givens(Hij(j, j), Hij(j + 1, j), c, s);
            cs[j] = c;
            sn[j] = s;
            cd h1 = Hij(j, j), h2 = Hij(j + 1, j);
            Hij(j, j) = c * h1 + s * h2;
            Hij(j + 1, j) = 0;
            cd g1 = g[j], g2 = g[j + 1];
            g[j] = c * g1 + s * g2;
            g[j + 1] = -std::conj(s) * g1 + c * g2;
            iterations++;
            kdone = j + 1;
            const double est = std::abs(g[j + 1]);
            if (backward && (kdone % 5 == 0 || est <= threshold()) && Hij(j, j) != cd(0, 0)) {
                form_x(kdone, H, g, W->trial.p, false);
                dev_norm<double>(W, W->trial.p, n, W->nscal.p + 5);
                xnorm = read_scalar(W, W->nscal.p + 5);
            }
            if (W->trace_progress && (iterations % 10 == 0))
                std::fprintf(stderr, "[pimi] fgmres iteration %lld  residual %.3e  |b| %.3e  |x| %.3e\n",
                             (long long)iterations, est, bnorm, xnorm);
            if (est <= threshold()) break;
        }
        form_x(kdone, H, g, W->x.p, true);
        restarts++;
        steady_apply(W, W->x.p, W->w.p, invscale);
        applications++;
        hipLaunchKernelGGL(k_sub<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->r.p, W->b.p, W->w.p,
                           n);
        dev_norm<double>(W, W->r.p, n, W->nscal.p);
        dev_norm<double>(W, W->x.p, n, W->nscal.p + 1);
        d2h(W, W->hpin, W->nscal.p, 2 * sizeof(double));
        beta = W->hpin[0];
        xnorm = W->hpin[1];
        if (beta <= threshold()) return {true, iterations, restarts, beta};
    }
    return {false, iterations, restarts, beta};
}

// Pinned read-back slot of the delayed outer loop: the step's mailbox
// (2 (m + 1) entries) followed by its 16 scalars.
static size_t dc_slot_entries(int m) { return 2 * ((size_t)m + 1) + 8; }

static void ensure_dcgs(pimi_workspace *W) {
    if (W->dc_ready) return;
    const int m = W->m;
    const int64_t n = W->plan->n;
    const int nblk = dcgs_single_nblk(n);
    W->dc_partial.alloc((size_t)2 * (m + 1) * nblk);
    W->dc_mail.alloc((size_t)2 * (m + 1));
    W->dc_h.alloc((size_t)2 * m + 2);
    W->dc_slot.alloc(16);
    // fine-grained pinned memory: the coefficient kernel writes each step's
    // mail and scalars here, readable once the step's event has completed
    HIPCHK(hipHostMalloc((void **)&W->dc_look, 2 * dc_slot_entries(m) * sizeof(cx<double>), hipHostMallocCoherent));
    std::vector<char> host(dcgs_single_desc_bytes(m));
    dcgs_single_descs(host.data(), m, W->V.p, W->w.p, W->dc_partial.p, W->dc_mail.p, W->dc_h.p, W->dc_slot.p, n,
                      W->dc_look, (int64_t)dc_slot_entries(m));
    W->dc_desc.alloc(host.size());
    h2d(W, W->dc_desc.p, host.data(), host.size());
    W->dc_ready = true;
}

// Flexible GMRES, outer loop with delayed reorthogonalization (after
// Swirydowicz, Langou, Ananthan, Yang and Thomas, "Low synchronization
// Gram-Schmidt and GMRES", 2020; own derivation).  Step j takes the
// once-orthogonalized candidate p_j = V[:, j], forms z'_j = inner(p_j) and
// w = A z'_j, and one pass over [Q, p_j] (Q = V[:, 0:j]) gives s = Q^H p_j,
// t = Q^H w, rho_j and e_j; a second pass writes q_j = (p_j - Q s)/rho_j and
// the next candidate p_{j+1} = (w - e p_j - Q(t - e s))/rho_j.  With
// z_j = z'_j / rho_j the flexible Arnoldi relation holds with Hessenberg
// column j = (t/rho_j + s_{j+1}[0:j]; e_j + s_{j+1}[j]; rho_{j+1}), final
// once step j+1 has run: two passes over the basis per step instead of
// classical Gram-Schmidt's four.  Column j is finalized one step late, so a
// converged cycle computes one (with the lookahead two) extra steps.
static GmresResult fgmres_device(pimi_workspace *W, double invscale, double atol, double rtol, int64_t maxiter,
                                 int64_t &applications, bool backward, int kin) {
    {
        const char *e = std::getenv("PIMI_OUTER_GS");
        if (e && std::strcmp(e, "cgs2") == 0)
            return fgmres_device_cgs2(W, invscale, atol, rtol, maxiter, applications, backward, kin);
    }
    const int64_t n = W->plan->n;
    const int m = W->m;
    const double scale = 1.0 / invscale;
    ensure_inner(W, kin);
    ensure_dcgs(W);
    cx<double> *V = W->V.p, *Z = W->Zo.p;
    const size_t se = dc_slot_entries(m);
    auto inner = [&](const cx<double> *r, cx<double> *z) {
        run_inner(W, kin, scale, r, z);
    };
    std::vector<cd> H((size_t)(m + 1) * m), g(m + 1), sn(m);
    std::vector<double> cs(m), rho_col(m + 1);
    auto Hij = [&](int i, int j) -> cd & { return H[(size_t)i + (size_t)j * (m + 1)]; };
    dev_norm<double>(W, W->b.p, n, W->nscal.p);
    dev_norm<double>(W, W->x.p, n, W->nscal.p + 1);
    d2h(W, W->hpin, W->nscal.p, 2 * sizeof(double));
    const double bnorm = W->hpin[0];
    double xnorm = W->hpin[1];
    auto threshold = [&]() { return backward ? 0.5 * (atol + rtol * xnorm) : atol + rtol * bnorm; };
    // dst = (add_to_x ? dst : x) + Z' (y ./ rho), y the least-squares solution
    auto form_x = [&](int k, cx<double> *dst, bool add_to_x) {
        std::vector<cd> yy(k);
        for (int i = k - 1; i >= 0; --i) {
            cd acc = g[i];
            for (int jj = i + 1; jj < k; ++jj) acc -= Hij(i, jj) * yy[jj];
            yy[i] = acc / Hij(i, i);
        }
        std::vector<cx<double>> yc(k);
        for (int i = 0; i < k; ++i) {
            const cd v = yy[i] / rho_col[i];
            yc[i] = cx<double>{v.real(), v.imag()};
        }
        h2d(W, W->coef.p, yc.data(), sizeof(cx<double>) * k);
        if (!add_to_x) HIPCHK(hipMemcpyAsync(dst, W->x.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, W->stream));
        hipLaunchKernelGGL(k_gemv<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, dst, Z, n, k, W->coef.p,
                           n, 1.0, 0);
    };
    // column c of H from its rows 0..c+1, rotated into the triangular factor
    auto finish_column = [&](int c, const std::vector<cd> &col) {
        for (int i = 0; i <= c + 1; ++i) Hij(i, c) = col[i];
        for (int i = 0; i < c; ++i) {
            cd h1 = Hij(i, c), h2 = Hij(i + 1, c);
            Hij(i, c) = cs[i] * h1 + sn[i] * h2;
            Hij(i + 1, c) = -std::conj(sn[i]) * h1 + cs[i] * h2;
        }
        double cc;
        cd ss;
        complex_givens(Hij(c, c), Hij(c + 1, c), cc, ss);
        cs[c] = cc;
        sn[c] = ss;
        cd h1 = Hij(c, c), h2 = Hij(c + 1, c);
        Hij(c, c) = cc * h1 + ss * h2;
        Hij(c + 1, c) = 0;
        cd g1 = g[c], g2 = g[c + 1];
        g[c] = cc * g1 + ss * g2;
        g[c + 1] = -std::conj(ss) * g1 + cc * g2;
    };
    int64_t iterations = 0, restarts = 0;
    steady_apply(W, W->x.p, W->w.p, invscale);
    applications++;
    hipLaunchKernelGGL(k_sub<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->r.p, W->b.p, W->w.p, n);
    dev_norm<double>(W, W->r.p, n, W->nscal.p);
    double beta = read_scalar(W, W->nscal.p);
    if (beta <= threshold()) return {true, 0, 0, beta};
    while (iterations < maxiter) {
        std::fill(H.begin(), H.end(), cd(0, 0));
        std::fill(g.begin(), g.end(), cd(0, 0));
        W->hpin[0] = beta;
        h2d(W, W->nscal.p + 1, W->hpin, sizeof(double));
        hipLaunchKernelGGL(k_scale_div<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, V, W->r.p, n,
                           W->nscal.p + 1);
        g[0] = beta;
        const int jmax = (int)std::min<int64_t>(m, maxiter - iterations);
        auto step = [&](int j) {
            inner(V + (int64_t)j * n, Z + (int64_t)j * n);
            steady_apply(W, Z + (int64_t)j * n, W->w.p, invscale);
            // the step's coefficient kernel writes its mail and scalars into
            // dc_look region j mod 2 (pinned host memory)
            dcgs_single_step(W->stream, W->dc_desc.p, j, n);
            HIPCHK(hipEventRecord(W->look_ev[j & 1], W->stream));
        };
        std::vector<cd> tprev, col;
        cd eprev(0, 0);
        int kdone = 0;
        bool stop = false;
        // column c = j - 1 from the previous step's t, e, rho and this step's s, rho
        auto close_column = [&](int j, const std::vector<cd> &sj, double rho_j) {
            col.assign((size_t)j + 1, cd(0, 0));
            for (int i = 0; i < j - 1; ++i) col[i] = tprev[i] / rho_col[j - 1] + sj[i];
            col[j - 1] = eprev + sj[j - 1];
            col[j] = rho_j;
            finish_column(j - 1, col);
            iterations++;
            applications += 1 + kin;
            kdone = j;
            const double est = std::abs(g[j]);
            if (backward && (kdone % 5 == 0 || est <= threshold()) && Hij(j - 1, j - 1) != cd(0, 0)) {
                form_x(kdone, W->trial.p, false);
                dev_norm<double>(W, W->trial.p, n, W->nscal.p + 5);
                xnorm = read_scalar(W, W->nscal.p + 5);
            }
            if (W->trace_progress && (iterations % 10 == 0))
                std::fprintf(stderr, "[pimi] fgmres iteration %lld  residual %.3e  |b| %.3e  |x| %.3e\n",
                             (long long)iterations, est, bnorm, xnorm);
            return est <= threshold();
        };
        if (jmax > 0) step(0);
        for (int j = 0; j < jmax && !stop; ++j) {
            if (j + 1 < jmax) step(j + 1);  // lookahead
            HIPCHK(hipEventSynchronize(W->look_ev[j & 1]));
            const cx<double> *sl = W->dc_look + (size_t)(j & 1) * se;
            const double *sc = reinterpret_cast<const double *>(sl + 2 * ((size_t)m + 1));
            std::vector<cd> sj(j), tj(j);
            for (int i = 0; i < j; ++i) {
                sj[i] = cd(sl[i].x, sl[i].y);
                tj[i] = cd(sl[j + 1 + i].x, sl[j + 1 + i].y);
            }
            const double rho_j = sc[10];
            const cd e_j(sc[12], sc[13]);
            if (j == 0) g[0] = beta * rho_j;  // r = beta p_0 = beta rho_0 q_0
            if (j >= 1) stop = close_column(j, sj, rho_j);
            tprev = tj;
            eprev = e_j;
            rho_col[j] = rho_j;
        }
        if (!stop && jmax > 0) {
            // the cycle ran its jmax steps: column jmax-1 needs s and rho of
            // the last candidate p = V[:, jmax] against V[:, 0:jmax]
            dev_dots<double>(W, V, n, jmax + 1, V + (int64_t)jmax * n, n, W->hvec.p, nullptr);
            std::vector<cx<double>> hv((size_t)jmax + 1);
            d2h(W, hv.data(), W->hvec.p, sizeof(cx<double>) * (jmax + 1));
            std::vector<cd> sj(jmax);
            double ss = 0;
            for (int i = 0; i < jmax; ++i) {
                sj[i] = cd(hv[i].x, hv[i].y);
                ss += std::norm(sj[i]);
            }
            const double r2 = hv[jmax].x - ss;
            close_column(jmax, sj, r2 > 0 ? std::sqrt(r2) : 0.0);
        }
        form_x(kdone, W->x.p, true);
        restarts++;
        steady_apply(W, W->x.p, W->w.p, invscale);
        applications++;
        hipLaunchKernelGGL(k_sub<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->r.p, W->b.p, W->w.p,
                           n);
        dev_norm<double>(W, W->r.p, n, W->nscal.p);
        dev_norm<double>(W, W->x.p, n, W->nscal.p + 1);
        d2h(W, W->hpin, W->nscal.p, 2 * sizeof(double));
        beta = W->hpin[0];
        xnorm = W->hpin[1];
        if (beta <= threshold()) return {true, iterations, restarts, beta};
    }
    return {false, iterations, restarts, beta};
}

static double operator_scale(pimi_workspace *W) {
    const int64_t n = W->plan->n;
    uint64_t st = 0x51ca1eULL;
    std::vector<cx<double>> h((size_t)n);
    double scale = 0;
    for (int k = 0; k < 3; ++k) {
        probe_vector(st, h);
        h2d(W, W->tmp.p, h.data(), sizeof(cx<double>) * n);
        dev_norm<double>(W, W->tmp.p, n, W->nscal.p);
        launch_apply<double>(W, W->tmp.p, W->w.p, 1.0);
        dev_norm<double>(W, W->w.p, n, W->nscal.p + 1);
        d2h(W, W->hpin, W->nscal.p, 2 * sizeof(double));
        scale = std::max(scale, W->hpin[1] / W->hpin[0]);
    }
    return scale;
}

static void steady_solve(pimi_workspace *W, const pimi_gmres_opts *o, const double *x0, double *rho_out,
                         pimi_solve_info *info) {
    auto t0 = std::chrono::steady_clock::now();
    pimi_plan *P = W->plan;
    const int64_t n = P->n;
    // the workspace holds min(krylovdim, n) columns, as KrylovWorkspace (src/krylov.jl:74-82)
    if (o->krylovdim < 1 || std::min<int64_t>(o->krylovdim, n) != W->m)
        throw Error(PIMI_EINVAL, "opts.krylovdim must match the workspace Krylov dimension");
    if (!(o->atol >= 0) || !(o->rtol >= 0)) throw Error(PIMI_EINVAL, "atol and rtol must be nonnegative");
    // v = t / (t . t) (src/source_protocol.jl:107-113)
    double tt = 0;
    for (double t : P->tval) tt += t * t;
    std::vector<double> vv(P->tval.size());
    for (size_t k = 0; k < vv.size(); ++k) vv[k] = P->tval[k] / tt;
    if (o->weighting != PIMI_WEIGHT_NONE && o->weighting != PIMI_WEIGHT_SECTOR)
        throw Error(PIMI_EINVAL, "unknown weighting");
    const bool weighted = o->weighting == PIMI_WEIGHT_SECTOR;
    if (weighted && o->criterion == PIMI_CRITERION_REFERENCE)
        throw Error(PIMI_EINVAL, "sector weighting needs the backward or the weighted criterion");
    if (!weighted && o->criterion == PIMI_CRITERION_WEIGHTED)
        throw Error(PIMI_EINVAL, "the weighted criterion needs sector weighting");
    const bool weighted_accept = o->criterion == PIMI_CRITERION_WEIGHTED;
    W->weighted = false;
    std::vector<double> vw(vv);
    if (weighted) {  // v -> S v on the trace support (ordered by sector, m entries each)
        ensure_plan_weighting(P);
        size_t k = 0;
        for (int q = 0; q < P->ns; ++q)
            for (int64_t a = 0; a < P->host.dims[(size_t)q]; ++a, ++k) vw[k] = vv[k] * P->sigma[(size_t)q];
    }
    W->d_vval.upload(vw.data(), vw.size(), W->stream);
    double scale = o->operator_scale > 0 ? o->operator_scale : operator_scale(W);
    if (!(scale > 0) || !std::isfinite(scale))
        throw Error(PIMI_EINVAL, "operator_scale must be finite and positive");
    HIPCHK(hipMemsetAsync(W->b.p, 0, sizeof(cx<double>) * n, W->stream));
    int64_t nt = (int64_t)vv.size();
    hipLaunchKernelGGL(k_scatter<double>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, W->stream, W->b.p,
                       P->d_tidx.p, W->d_vval.p, nt);
    if (x0) {  // initial_state (src/krylov.jl:818-821); in weighted coordinates u0 = S x0
        std::vector<cx<double>> xs((size_t)n);
        std::memcpy(xs.data(), x0, sizeof(cx<double>) * (size_t)n);
        for (const cx<double> &z : xs)
            if (!std::isfinite(z.x) || !std::isfinite(z.y)) throw Error(PIMI_EINVAL, "initial_state must be finite");
        if (weighted)
            for (int q = 0; q < P->ns; ++q) {
                const double sg = P->sigma[(size_t)q];
                const int64_t m = P->host.dims[(size_t)q];
                for (int64_t i = P->host.offs[(size_t)q]; i < P->host.offs[(size_t)q] + m * m; ++i)
                    xs[(size_t)i] = cx<double>{xs[(size_t)i].x * sg, xs[(size_t)i].y * sg};
            }
        h2d(W, W->x.p, xs.data(), sizeof(cx<double>) * (size_t)n);
    } else {
        HIPCHK(hipMemcpyAsync(W->x.p, W->b.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, W->stream));
    }
    int64_t apps = 3;
    const bool precond = o->preconditioner == PIMI_PRECOND_SECTOR;
    const bool inner_outer = o->preconditioner == PIMI_PRECOND_SECTOR_INNER;
    if (o->preconditioner != PIMI_PRECOND_NONE && !precond && !inner_outer)
        throw Error(PIMI_EINVAL, "unknown preconditioner");
    if (inner_outer && (o->inner_steps < 1 || o->inner_steps > 32))
        throw Error(PIMI_EINVAL, "inner_steps must lie in 1..32");
    if (W->prec == PIMI_C32 && !inner_outer)
        throw Error(PIMI_EINVAL, "complex FP32 workspaces solve with the inner-outer preconditioner");
    if (precond || inner_outer) ensure_preconditioner(W);
    if (inner_outer) ensure_diag_stencil(P);
    const bool backward = o->criterion == PIMI_CRITERION_BACKWARD || weighted_accept;
    if (o->criterion != PIMI_CRITERION_REFERENCE && !backward) throw Error(PIMI_EINVAL, "unknown criterion");
    W->weighted = weighted;
    // With sector weighting the iteration controls |S A S^{-1} u - S v| against
    // |u|.  Under the backward criterion the acceptance below measures
    // |L x| / (s |x|) on x = S^{-1} u; if that has not followed, the iteration
    // continues with a tighter threshold.  Under the weighted criterion the
    // acceptance is the weighted relative residual |S L x| / (s |S x|), the
    // quantity the iteration controls.
    GmresResult gr{false, 0, 0, 0.0};
    double tight = 1.0;
    for (int pass = 0;; ++pass) {
        const int64_t left = o->maxiter - gr.iterations;
        GmresResult g = inner_outer ? fgmres_device(W, 1.0 / scale, o->atol * tight, o->rtol * tight, left, apps,
                                                    backward, o->inner_steps)
                                    : gmres_device(W, 1.0 / scale, o->atol * tight, o->rtol * tight, left, apps,
                                                   precond, backward);
        gr.converged = g.converged;
        gr.iterations += g.iterations;
        gr.restarts += g.restarts;
        gr.residual = g.residual;
        if (!weighted || weighted_accept || !g.converged || pass >= 4 || gr.iterations >= o->maxiter) break;
        hipLaunchKernelGGL(k_sector_scale<double>, dim3((unsigned)P->ntiles), dim3(kApplyBlock), 0, W->stream,
                           W->x.p, W->tmp.p, n, P->offs.p, P->tile_sector.p, P->d_inv_sigma.p);
        launch_apply<double>(W, W->tmp.p, W->w.p, 1.0);
        apps++;
        dev_norm<double>(W, W->w.p, n, W->nscal.p);
        dev_norm<double>(W, W->tmp.p, n, W->nscal.p + 1);
        d2h(W, W->hpin, W->nscal.p, 2 * sizeof(double));
        if (W->hpin[0] / scale <= o->atol + o->rtol * W->hpin[1]) break;
        tight *= 0.1;
    }
    W->weighted = false;
    if (backward) {
        // The trace row of the bordered system carries weight |v| = |t|^{-1}; under the
        // scale-invariant rule it is restored exactly by x <- x / (t . x), which leaves
        // |L x| / |x| unchanged.  Sector weighting is undone first (x = S^{-1} u).
        std::vector<cx<double>> xs((size_t)n);
        d2h(W, xs.data(), W->x.p, sizeof(cx<double>) * n);
        if (weighted)
            for (int q = 0; q < P->ns; ++q) {
                const double inv = 1.0 / P->sigma[(size_t)q];
                const int64_t m = P->host.dims[(size_t)q];
                for (int64_t i = P->host.offs[(size_t)q]; i < P->host.offs[(size_t)q] + m * m; ++i)
                    xs[(size_t)i] = cx<double>{xs[(size_t)i].x * inv, xs[(size_t)i].y * inv};
            }
        cd tr0(0, 0);
        for (size_t k = 0; k < P->tidx.size(); ++k) tr0 += P->tval[k] * cd(xs[P->tidx[k]].x, xs[P->tidx[k]].y);
        if (tr0 == cd(0, 0)) throw Error(PIMI_ENOTCONVERGED, "stationary iterate has zero trace");
        for (auto &z : xs) {
            cd v = cd(z.x, z.y) / tr0;
            z = cx<double>{v.real(), v.imag()};
        }
        h2d(W, W->x.p, xs.data(), sizeof(cx<double>) * n);
    }
    // acceptance (src/krylov.jl:868-873)
    launch_apply<double>(W, W->x.p, W->w.p, 1.0);
    apps++;
    dev_norm<double>(W, W->w.p, n, W->nscal.p);
    dev_norm<double>(W, W->x.p, n, W->nscal.p + 1);
    std::vector<cx<double>> xh((size_t)n);
    HIPCHK(hipMemcpyAsync(xh.data(), W->x.p, sizeof(cx<double>) * n, hipMemcpyDeviceToHost, W->stream));
    HIPCHK(hipMemcpyAsync(W->hpin, W->nscal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, W->stream));
    sync(W);
    double resid = W->hpin[0], xn = W->hpin[1];
    // the weighted relative residual |S L x| / (s |S x|), S = sqrt(f^nu) per sector
    double wres = NAN;
    if (P->compiled && !P->host.geom.sqrt_f.empty()) {
        ensure_plan_weighting(P);
        hipLaunchKernelGGL(k_sector_scale<double>, dim3((unsigned)P->ntiles), dim3(kApplyBlock), 0, W->stream,
                           W->w.p, W->tmp.p, n, P->offs.p, P->tile_sector.p, P->d_sigma.p);
        dev_norm<double>(W, W->tmp.p, n, W->nscal.p + 2);
        hipLaunchKernelGGL(k_sector_scale<double>, dim3((unsigned)P->ntiles), dim3(kApplyBlock), 0, W->stream,
                           W->x.p, W->tmp.p, n, P->offs.p, P->tile_sector.p, P->d_sigma.p);
        dev_norm<double>(W, W->tmp.p, n, W->nscal.p + 3);
        d2h(W, W->hpin + 2, W->nscal.p + 2, 2 * sizeof(double));
        wres = W->hpin[3] > 0 ? W->hpin[2] / (scale * W->hpin[3]) : INFINITY;
    }
    cd tr(0, 0);
    for (size_t k = 0; k < P->tidx.size(); ++k) tr += P->tval[k] * cd(xh[P->tidx[k]].x, xh[P->tidx[k]].y);
    double terr = std::abs(tr - 1.0);
    double nres = resid / scale;
    double tol = backward ? o->atol + o->rtol * xn : o->atol + o->rtol * std::max(xn, 1.0);
    std::memcpy(rho_out, xh.data(), sizeof(cx<double>) * n);
    std::memset(info, 0, sizeof(*info));
    info->converged = gr.converged && terr <= o->atol + o->rtol &&
                      (weighted_accept ? wres <= o->atol + o->rtol : nres <= tol);
    info->iterations = gr.iterations;
    info->restarts = gr.restarts;
    info->applications = apps;
    info->residual = resid;
    info->normalized_residual = nres;
    info->trace_error = terr;
    info->operator_scale = scale;
    info->linear_residual = gr.residual;
    info->relative_residual = xn > 0 ? nres / xn : INFINITY;
    info->weighted_relative_residual = wres;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!info->converged) {
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "matrix-free Krylov steady-state solve did not converge in %lld iterations; "
                      "linear residual %.3e, normalized_residual=%.3e, trace_error=%.3e",
                      (long long)gr.iterations, gr.residual, nres, terr);
        throw Error(PIMI_ENOTCONVERGED, buf);
    }
}

// ---------------------------------------------------------------------------
// adaptive Krylov exponential (src/krylov_extensions.jl:1346-1568)
// ---------------------------------------------------------------------------
static void ensure_expv_arnoldi(pimi_workspace *W) {
    if (W->ex_ready) return;
    const int m = W->m, steps = m + 2;
    const int64_t n = W->plan->n;
    W->ex_stride = kArnMailHead + 2 * ((int64_t)steps + 1);
    W->ex_partial.alloc((size_t)2 * steps * dcgs_single_nblk(n));
    W->ex_mail.alloc((size_t)(W->ex_stride * steps));
    W->ex_h.alloc((size_t)2 * steps);
    std::vector<char> host(dcgs_single_desc_bytes(steps));
    dcgs_arnoldi_descs(host.data(), steps, W->V.p, W->w.p, W->ex_partial.p, W->ex_mail.p, W->ex_stride, W->ex_h.p,
                       n);
    W->ex_desc.alloc(host.size());
    h2d(W, W->ex_desc.p, host.data(), host.size());
    HIPCHK(hipHostMalloc((void **)&W->ex_host, sizeof(cx<double>) * (size_t)(W->ex_stride * steps),
                         hipHostMallocDefault));
    W->ex_ready = true;
}

// The Krylov factorization of krylov_expv!'s Arnoldi (src/krylov_extensions.jl:
// 1346-1401) from W->cur: beta = |cur|, V[:, 0] = cur / beta, then steps
// 0..m of the delayed-reorthogonalization Arnoldi (two passes over the basis
// per step; V[:, j] holds q_j once step j has run) on the device with no
// host round trip, each step's reductions in a mail region of its own; the
// host reads them once and finds the breakdowns (DelayedArnoldi).  Steps
// after an exact breakdown divide by rho = 0 as multiplication by zero and
// write columns that are never read.  After a near breakdown one more step
// (the enlargement) runs.
static void arnoldi(pimi_workspace *W, Factorization &F) {
    const int64_t n = W->plan->n;
    const int m = W->m;
    cx<double> *V = W->V.p;
    F = Factorization();
    dev_norm<double>(W, W->cur.p, n, W->nscal.p);
    const double beta = read_scalar(W, W->nscal.p);
    if (beta == 0) {
        F.S.assign((size_t)(m + 1) * (m + 1), cd(0, 0));
        F.zero_state = true;
        F.completed = true;
        return;
    }
    ensure_expv_arnoldi(W);
    hipLaunchKernelGGL(k_scale_div<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, V, W->cur.p, n,
                       W->nscal.p);
    const int64_t stride = W->ex_stride;
    auto run = [&](int j0, int j1) {  // steps j0 .. j1 - 1, then their regions to the host
        for (int j = j0; j < j1; ++j) {
            launch_apply<double>(W, V + (int64_t)j * n, W->w.p, 1.0);
            dcgs_single_step(W->stream, W->ex_desc.p, j, n);
        }
        HIPCHK(hipMemcpyAsync(W->ex_host + (size_t)(j0 * stride), W->ex_mail.p + (size_t)(j0 * stride),
                              sizeof(cx<double>) * (size_t)((j1 - j0) * stride), hipMemcpyDeviceToHost,
                              W->stream));
        sync(W);
    };
    DelayedArnoldi A;
    A.reset(m, n);
    auto feed = [&](int j) {
        const cx<double> *mb = W->ex_host + (size_t)(j * stride);
        const double *sc = reinterpret_cast<const double *>(mb);
        const cx<double> *sv = mb + kArnMailHead;
        return A.step(j, sv, sv + (j + 1), sc[kArnSlotRho], cd(sc[kArnSlotE], sc[kArnSlotE + 1]));
    };
    run(0, m + 1);
    int j = 0, r = 0;
    for (; j <= m; ++j)
        if ((r = feed(j)) != 0) break;
    if (r == 0) {  // a near breakdown at step m: the enlargement
        j = m + 1;
        run(j, j + 1);
        r = feed(j);
    }
    if (r < 0) throw Error(PIMI_EINVAL, "the Krylov exponential's state is not finite");
    A.finish(j, beta, F);
}

// trial = beta V exp(h S) e1 ; returns error estimate
static double project(pimi_workspace *W, const Factorization &F, double h) {
    const int64_t n = W->plan->n;
    if (F.zero_state) {
        HIPCHK(hipMemsetAsync(W->trial.p, 0, sizeof(cx<double>) * n, W->stream));
        return 0.0;
    }
    std::vector<cx<double>> c;
    const double err = expv_coefficients(F, W->m + 1, h, c);
    h2d(W, W->coef.p, c.data(), sizeof(cx<double>) * F.coeff_count);
    hipLaunchKernelGGL(k_gemv<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->trial.p, W->V.p, n,
                       F.coeff_count, W->coef.p, n, F.beta, 1);
    return err;
}

struct ExpvStats {
    int64_t accepted = 0, rejected = 0, applications = 0, factorizations = 0, attempts = 0;
    double error_sum = 0;
};

// W->cur holds b on entry and exp(t A) b on exit.
static bool expv_device(pimi_workspace *W, double t, const pimi_expv_opts *o, ExpvStats &S) {
    const int64_t n = W->plan->n;
    ExpvControl X;
    if (!X.begin(t)) return true;
    Factorization F;
    bool have = false;
    while (!X.done() && S.attempts < o->max_steps) {
        S.attempts++;
        const double hs = X.step();
        if (!have) {
            arnoldi(W, F);
            S.applications += F.used;
            S.factorizations++;
            have = true;
        }
        const double err = project(W, F, hs);
        dev_norm<double>(W, W->trial.p, n, W->nscal.p + 4);
        const double trial_norm = read_scalar(W, W->nscal.p + 4);
        auto cur_norm = [&]() {
            dev_norm<double>(W, W->cur.p, n, W->nscal.p + 4);
            return read_scalar(W, W->nscal.p + 4);
        };
        const int d = X.decide(err, trial_norm, cur_norm, F, *o);
        if (d > 0) {
            std::swap(W->cur.p, W->trial.p);
            S.accepted++;
            S.error_sum += err;
            have = false;
        } else {
            S.rejected++;
            if (d < 0) break;
        }
    }
    return X.done();
}

static void expv_series(pimi_workspace *W, const double *rho0, const double *ts, int64_t nts,
                        const pimi_expv_opts *o, int32_t nobs, const double *obs, double *obs_out,
                        double *rho_final, pimi_solve_info *info) {
    auto t0 = std::chrono::steady_clock::now();
    const int64_t n = W->plan->n;
    // the workspace holds min(krylovdim, n) columns, as KrylovWorkspace (src/krylov.jl:74-82)
    if (o->krylovdim < 1 || std::min<int64_t>(o->krylovdim, n) != W->m)
        throw Error(PIMI_EINVAL, "opts.krylovdim must match the workspace Krylov dimension");
    if (!(o->safety > 0 && o->safety < 1)) throw Error(PIMI_EINVAL, "safety must lie strictly between zero and one");
    if (o->max_steps <= 0) throw Error(PIMI_EINVAL, "max_steps must be positive");
    for (int64_t k = 1; k < nts; ++k)
        if (!(ts[k] >= ts[k - 1])) throw Error(PIMI_EINVAL, "Krylov exponential output times must be nondecreasing");
    h2d(W, W->cur.p, rho0, sizeof(cx<double>) * n);
    if (nobs > 0) {
        W->obs.alloc((size_t)nobs * n);
        h2d(W, W->obs.p, obs, sizeof(cx<double>) * n * nobs);
    }
    std::vector<cx<double>> hv((size_t)std::max(nobs, 1));
    ExpvStats S;
    bool ok = true;
    for (int64_t k = 0; k < nts; ++k) {
        if (k > 0) {
            double dt = ts[k] - ts[k - 1];
            if (dt != 0 && !expv_device(W, dt, o, S)) { ok = false; break; }
        }
        if (nobs > 0) {
            dev_dots<double>(W, W->obs.p, n, nobs, W->cur.p, n, W->hvec.p, nullptr);
            d2h(W, hv.data(), W->hvec.p, sizeof(cx<double>) * nobs);
            for (int32_t q = 0; q < nobs; ++q) {
                obs_out[2 * (k * nobs + q)] = hv[q].x;
                obs_out[2 * (k * nobs + q) + 1] = hv[q].y;
            }
        }
    }
    if (rho_final) d2h(W, rho_final, W->cur.p, sizeof(cx<double>) * n);
    else sync(W);
    std::memset(info, 0, sizeof(*info));
    info->converged = ok;
    info->accepted = S.accepted;
    info->rejected = S.rejected;
    info->applications = S.applications;
    info->factorizations = S.factorizations;
    info->iterations = S.attempts;
    info->estimated_error = S.error_sum;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!ok) throw Error(PIMI_ENOTCONVERGED, "adaptive Krylov exponential action did not reach the requested time");
}

}  // namespace pimi

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace pimi;

// ---------------------------------------------------------------------------
// conditional (no-jump) trajectories: batched_conditional_action!,
// batched_conditional_rk4!, batched_channel_intensities!, batched_apply_jumps!
// (src/batched_trajectories.jl:266-598).  The conditional drift of a state X is
//   -i[H, X] - {E, X}/2 + lambda X,   E = sum_k rate_k L_k^dagger L_k,
//   lambda = sum_nu sqrt(f^nu) Re tr(E_nu X_nu)  (total jump intensity),
// i.e. the left/right (H, Q) part of the stencil plus lambda X; a jump in
// channel k maps X to gain_k(X) / (t . gain_k(X)).
// ---------------------------------------------------------------------------
__global__ void k_cols_axpy_lam(cx<double> *__restrict__ Y, const cx<double> *__restrict__ X,
                                const cx<double> *__restrict__ h, cx<double> *__restrict__ lam_out, int64_t n,
                                int64_t ncols) {
    const int64_t total = n * ncols;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = k / n;
        const double lam = fmax(h[c].x, 0.0);  // intensities are clamped at zero after the sign check
        Y[k] = cx<double>{Y[k].x + lam * X[k].x, Y[k].y + lam * X[k].y};
        if (k % n == 0 && lam_out) lam_out[c] = h[c];
    }
}

// out = a + s b over len entries (RK4 stage combinations)
__global__ void k_lincomb(cx<double> *__restrict__ out, const cx<double> *a, const cx<double> *b, double s,
                          int64_t len) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x) {
        const cx<double> av = a[k], bv = b[k];
        out[k] = cx<double>{av.x + s * bv.x, av.y + s * bv.y};
    }
}

// z[c] = t . X[:, c] on the trace support, one block per column
__global__ __launch_bounds__(256) void k_cols_trace(const cx<double> *__restrict__ X, int64_t n,
                                                    const int64_t *__restrict__ idx, const double *__restrict__ t,
                                                    int64_t nt, cx<double> *__restrict__ z) {
    __shared__ double sh[256 / 64];
    const cx<double> *x = X + (int64_t)blockIdx.x * n;
    double re = 0, im = 0;
    for (int64_t k = threadIdx.x; k < nt; k += 256) {
        re += t[k] * x[idx[k]].x;
        im += t[k] * x[idx[k]].y;
    }
    re = block_sum(re, sh);
    __syncthreads();
    im = block_sum(im, sh);
    if (threadIdx.x == 0) z[blockIdx.x] = cx<double>{re, im};
}

// X[:, c] /= z[c]
__global__ void k_cols_div(cx<double> *__restrict__ X, const cx<double> *__restrict__ z, int64_t n, int64_t ncols) {
    const int64_t total = n * ncols;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
        const cx<double> d = z[k / n], v = X[k];
        const double den = d.x * d.x + d.y * d.y;
        X[k] = cx<double>{(v.x * d.x + v.y * d.y) / den, (v.y * d.x - v.x * d.y) / den};
    }
}

static void ensure_trajectory_tables(pimi_plan *P) {
    std::lock_guard<std::recursive_mutex> lk(P->lazy_mu);
    if (P->traj_ready) return;
    const FusedPlan &F = P->host;
    if (!P->compiled || F.geom.sqrt_f.empty())
        throw Error(PIMI_EINVAL, "trajectory operations need a plan compiled from a model");
    auto skeleton = [&](FusedPlan &G) {
        G.ns = F.ns;
        G.n = F.n;
        G.dims = F.dims;
        G.offs = F.offs;
    };
    {   // the H/Q-only (conditional) stencil
        FusedPlan C;
        skeleton(C);
        C.has_h = F.has_h;
        C.has_q = F.has_q;
        C.H = F.H;
        C.Q = F.Q;
        Stencil st;
        build_stencil(C, st);
        P->cond.upload(st, F.dims, F.offs);
    }
    P->gain.clear();
    for (const JumpChannel &ch : F.channels) {  // one gain stencil per channel, scaled by its rate
        FusedPlan G;
        skeleton(G);
        if (ch.kind == 0) {
            G.cg_scale = {F.cg_scale[(size_t)ch.index]};
            G.K = {F.K[(size_t)ch.index]};
        } else {
            G.ob_scale = {F.ob_scale[(size_t)ch.index]};
            G.ob = {F.ob[(size_t)ch.index]};
        }
        Stencil st;
        build_stencil(G, st);
        std::unique_ptr<DevStencil> ds(new DevStencil());
        ds->upload(st, F.dims, F.offs);
        P->gain.push_back(std::move(ds));
    }
    // intensity weights: tr(E_nu X_nu) = sum_{r,c} E[r,c] X[c,r], X[c,r] at offs + c + r m
    const size_t nch = F.channels.size(), n = (size_t)F.n;
    std::vector<cx<double>> eb((nch + 1) * n, cx<double>{0, 0});
    auto fill = [&](size_t slot, const std::vector<CSC> &Qs, double scale) {
        for (int s = 0; s < F.ns; ++s) {
            const double w = F.geom.sqrt_f[(size_t)s] * scale;
            const int64_t m = F.dims[(size_t)s];
            const CSC &q = Qs[(size_t)s];
            for (int64_t c = 0; c < q.n; ++c)
                for (int64_t p = q.colptr[(size_t)c]; p < q.colptr[(size_t)c + 1]; ++p) {
                    const int64_t rr = q.rows[(size_t)p];
                    const cd e = w * q.vals[(size_t)p];
                    eb[slot * n + (size_t)(F.offs[(size_t)s] + c + rr * m)] = cx<double>{e.real(), -e.imag()};
                }
        }
    };
    if (F.has_q) fill(0, F.Q, 1.0);
    for (size_t k = 0; k < nch; ++k) fill(k + 1, F.channels[k].q, F.channels[k].rate);
    P->ebar.upload(eb.data(), eb.size());
    P->traj_ready = true;
}

static void ensure_traj_ws(pimi_workspace *W, int64_t ncols) {
    if (W->traj_cap >= ncols) return;
    const size_t len = (size_t)W->plan->n * (size_t)ncols;
    W->tX.alloc(len);
    W->tY.alloc(len);
    W->tTmp.alloc(len);
    W->tK1.alloc(len);
    W->tK2.alloc(len);
    W->tPartial.alloc((size_t)kDotBlocks * (size_t)ncols);
    W->tH.alloc((size_t)ncols);
    W->tZ.alloc((size_t)ncols);
    W->tLam.alloc((size_t)4 * ncols);
    W->traj_cap = ncols;
}

static void launch_stencil(pimi_workspace *W, const DevStencil &S, const cx<double> *x, cx<double> *y,
                           int64_t ncols) {
    pimi_plan *P = W->plan;
    if (S.dia_ok) {
        launch_dia<double>(W->stream, P, S, false, x, y, 1.0, ncols,
                           ncols > 1 ? P->dia_w_multi[0] : P->dia_w_single[0]);
        return;
    }
    dim3 grid((unsigned)(P->atiles_per_xcd * kNumXcd), (unsigned)std::min<int64_t>(ncols, 65535));
    hipLaunchKernelGGL(k_apply<double>, grid, dim3(kApplyWave), 0, W->stream, x, y, P->n, P->atiles.p, P->natiles,
                       P->atiles_per_xcd, P->offs.p, P->dims.p, S.entry_base.p, S.entries.p, S.col.p, S.val64.p, 1.0,
                       ncols);
    HIPCHK(hipGetLastError());
}

// h[c] = sum_i conj(X[i, c]) ebar[i] = conj(sum_i X e), whose real part is the intensity
// (partial: kDotBlocks x ncols reduction partials)
static void cols_intensity(pimi_workspace *W, const cx<double> *X, int64_t ncols, const cx<double> *ebar,
                           cx<double> *h, cx<double> *partial) {
    const int64_t n = W->plan->n;
    hipLaunchKernelGGL(k_dots_partial<double>, dim3(kDotBlocks, (unsigned)ncols), dim3(kRedBlock), 0, W->stream, X,
                       n, ebar, n, partial);
    hipLaunchKernelGGL(k_reduce_cols, dim3((unsigned)ncols), dim3(kRedBlock), 0, W->stream, partial, kDotBlocks, h,
                       (cx<double> *)nullptr);
}
static void cols_intensity(pimi_workspace *W, const cx<double> *X, int64_t ncols, const cx<double> *ebar,
                           cx<double> *h) {
    cols_intensity(W, X, ncols, ebar, h, W->tPartial.p);
}

// Y = conditional drift of X; lam (ncols) receives the unclamped total
// intensities; h (ncols) and partial (kDotBlocks x ncols) are scratch
static void conditional_action(pimi_workspace *W, const cx<double> *X, cx<double> *Y, int64_t ncols,
                               cx<double> *lam, cx<double> *h, cx<double> *partial) {
    pimi_plan *P = W->plan;
    const int64_t n = P->n;
    launch_stencil(W, P->cond, X, Y, ncols);
    cols_intensity(W, X, ncols, P->ebar.p, h, partial);
    hipLaunchKernelGGL(k_cols_axpy_lam, dim3(grid_for(n * ncols)), dim3(kRedBlock), 0, W->stream, Y, X, h, lam, n,
                       ncols);
}
static void conditional_action(pimi_workspace *W, const cx<double> *X, cx<double> *Y, int64_t ncols,
                               cx<double> *lam) {
    conditional_action(W, X, Y, ncols, lam, W->tH.p, W->tPartial.p);
}

// ---------------------------------------------------------------------------
// The allocation-free entry point's scratch: one fixed array per device
// (each GPU the library runs on holds its own copy until the process ends).
// Host code never passes the array's name to a kernel: pimi_context_create
// resolves its device address with hipGetSymbolAddress, with the context's
// device current, and the kernels take that address.  Layout: the intensity
// reduction's partials (kDotBlocks x ncols), then h (ncols), lam (ncols) and
// one status slot.  Calls that overlap on one GPU share it and race.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMaxCols = 256;
constexpr int64_t kNoAllocScratch = kDotBlocks * kNoAllocMaxCols + 2 * kNoAllocMaxCols + 1;  // complex doubles
static __device__ __align__(256) cx<double> scratch[kNoAllocScratch];

// intensities[c] = max(0, Re lam[c]); *status = 1 when any Re lam[c] is
// nonfinite or below -1e-11 (the reference's sign check), else 0.  One block;
// the status is written unconditionally, so it needs no clearing.
__global__ __launch_bounds__(256) void k_lam_clamp_check(const cx<double> *__restrict__ lam, int64_t ncols,
                                                         double *__restrict__ intensities, cx<double> *status) {
    int bad = 0;
    for (int64_t c = threadIdx.x; c < ncols; c += 256) {
        const double v = lam[c].x;
        bad |= !(v == v) || v - v != 0.0 || v < -1e-11;  // NaN, +-inf, or a negative trace
        intensities[c] = v > 0.0 ? v : 0.0;
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) *status = cx<double>{bad ? 1.0 : 0.0, 0.0};
}

// Intensities are finite and not below -max(1e-11, 100 eps) (src/trajectories.jl:755)
static void check_intensities(const std::vector<cx<double>> &h, const char *what) {
    for (size_t c = 0; c < h.size(); ++c) {
        if (!std::isfinite(h[c].x)) throw Error(PIMI_EINVAL, std::string(what) + " is nonfinite");
        if (h[c].x < -1e-11) throw Error(PIMI_EINVAL, std::string(what) + " has a negative trace");
    }
}

// host (on_device = 0) or device input/output of n x ncols complex doubles
struct TrajIO {
    pimi_workspace *W;
    int64_t count;
    bool on_device;
    cx<double> *dev(const void *p, DevBuf<cx<double>> &buf) {
        if (on_device) return (cx<double> *)p;
        HIPCHK(hipMemcpyAsync(buf.p, p, sizeof(cx<double>) * count, hipMemcpyHostToDevice, W->stream));
        return buf.p;
    }
    void back(void *p, const cx<double> *d) {
        if (on_device) return;
        HIPCHK(hipMemcpyAsync(p, d, sizeof(cx<double>) * count, hipMemcpyDeviceToHost, W->stream));
    }
};

static void traj_prologue(pimi_workspace *W, int64_t ncols) {
    if (W->prec != PIMI_C64) throw Error(PIMI_EINVAL, "trajectory operations need a complex FP64 workspace");
    if (ncols < 1 || ncols > 65535) throw Error(PIMI_EINVAL, "ncols must lie in 1..65535");
    HIPCHK(hipSetDevice(W->plan->ctx->device));
    ensure_trajectory_tables(W->plan);
    ensure_traj_ws(W, ncols);
}

extern "C" {

const char *pimi_last_error(void) { return g_last_error.c_str(); }
const char *pimi_version(void) { return "pimi 0.1 gfx950"; }

pimi_status pimi_context_create(int32_t device_ordinal, pimi_context **out) {
    return guarded([&] {
        if (!out) throw Error(PIMI_EINVAL, "null output pointer");
        int count = 0;
        HIPCHK(hipGetDeviceCount(&count));
        if (device_ordinal < 0 || device_ordinal >= count) throw Error(PIMI_EINVAL, "device ordinal out of range");
        HIPCHK(hipSetDevice(device_ordinal));
        // the allocation-free entry points' scratch array: its address on
        // this device (each device has its own copy)
        void *sp = nullptr;
        const hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !sp)
            throw Error(PIMI_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        auto *c = new pimi_context();
        c->device = device_ordinal;
        c->noalloc_scratch = sp;
        *out = c;
    });
}

void pimi_context_destroy(pimi_context *ctx) { delete ctx; }

pimi_status pimi_compile(pimi_context *ctx, const pimi_model_desc *model, pimi_plan **out) {
    return guarded([&] {
        if (!model || !out) throw Error(PIMI_EINVAL, "null argument");
        if (model->nterms < 0 || (model->nterms > 0 && !model->terms)) throw Error(PIMI_EINVAL, "bad term list");
        const int d = model->d;
        std::vector<TermDesc> terms;
        for (int k = 0; k < model->nterms; ++k) {
            const pimi_term &t = model->terms[k];
            TermDesc td;
            td.kind = t.kind;
            td.rate = t.rate;
            td.hbar = t.hbar;
            if (!t.op) throw Error(PIMI_EINVAL, "term operator is null");
            td.op.resize((size_t)d * d);
            for (int q = 0; q < d * d; ++q) td.op[q] = cd(t.op[2 * q], t.op[2 * q + 1]);
            if (t.kind == PIMI_TERM_DIRECT_PRODUCT_H) {
                if (!t.op2) throw Error(PIMI_EINVAL, "direct product term needs op2");
                td.op2.resize((size_t)d * d);
                for (int q = 0; q < d * d; ++q) td.op2[q] = cd(t.op2[2 * q], t.op2[2 * q + 1]);
            }
            terms.push_back(std::move(td));
        }
        std::unique_ptr<pimi_plan> P(new pimi_plan());
        P->ctx = ctx;
        lower_model(model->N, model->d, terms, P->host);
        P->compiled = true;
        finalize_plan(P.get());
        *out = P.release();
    });
}

pimi_status pimi_plan_create_fused(pimi_context *ctx, const pimi_fused_desc *D, pimi_plan **out) {
    return guarded([&] {
        if (!ctx || !D || !out) throw Error(PIMI_EINVAL, "null argument");
        if (D->ns <= 0 || !D->dims) throw Error(PIMI_EINVAL, "empty fused plan");
        std::unique_ptr<pimi_plan> P(new pimi_plan());
        P->ctx = ctx;
        FusedPlan &F = P->host;
        F.ns = D->ns;
        int64_t off = 0;
        for (int s = 0; s < F.ns; ++s) {
            if (D->dims[s] <= 0) throw Error(PIMI_EDIM, "sector dimensions must be positive");
            F.dims.push_back(D->dims[s]);
            F.offs.push_back(off);
            off += D->dims[s] * D->dims[s];
        }
        F.n = off;
        auto read_set = [&](const pimi_csc_set &S, int64_t blk, int64_t m) {
            CSC C;
            C.zero(m, m);
            const int64_t *cp = S.colptr + S.cp_base[blk];
            for (int64_t c = 0; c < m; ++c) {
                for (int64_t p = cp[c]; p < cp[c + 1]; ++p) {
                    if (S.rows[p] < 0 || S.rows[p] >= m) throw Error(PIMI_EDIM, "CSC row index out of range");
                    C.rows.push_back(S.rows[p]);
                    C.vals.push_back(cd(S.vals[2 * p], S.vals[2 * p + 1]));
                }
                C.colptr[c + 1] = (int64_t)C.rows.size();
            }
            return C;
        };
        F.has_h = D->has_h;
        F.has_q = D->has_q;
        for (int s = 0; s < F.ns; ++s) {
            if (F.has_h) F.H.push_back(read_set(D->H, s, F.dims[s]));
            if (F.has_q) F.Q.push_back(read_set(D->Q, s, F.dims[s]));
        }
        for (int g = 0; g < D->ncg; ++g) {
            F.cg_scale.push_back(cd(D->cg_scale[2 * g], D->cg_scale[2 * g + 1]));
            std::vector<CSC> ks;
            for (int s = 0; s < F.ns; ++s) ks.push_back(read_set(D->K, (int64_t)g * F.ns + s, F.dims[s]));
            F.K.push_back(std::move(ks));
        }
        for (int g = 0; g < D->nob; ++g) {
            F.ob_scale.push_back(cd(D->ob_scale[2 * g], D->ob_scale[2 * g + 1]));
            std::vector<Branch> brs;
            for (int64_t b = D->ob_br_base[g]; b < D->ob_br_base[g + 1]; ++b) {
                Branch br;
                br.out = D->br_out[b];
                br.in = D->br_in[b];
                if (br.out < 0 || br.out >= F.ns || br.in < 0 || br.in >= F.ns)
                    throw Error(PIMI_EDIM, "branch sector out of range");
                br.r = D->br_r[b];
                br.use_support = D->br_use_support[b];
                int64_t ml = F.dims[br.out], mn = F.dims[br.in];
                br.dense.assign((size_t)(ml * mn), cd(0, 0));
                if (br.use_support) {
                    for (int64_t e = D->br_sup_base[b]; e < D->br_sup_base[b + 1]; ++e) {
                        int32_t r = D->sup_rows[e], c = D->sup_cols[e];
                        if (r < 0 || r >= ml || c < 0 || c >= mn) throw Error(PIMI_EDIM, "support index out of range");
                        cd v(D->sup_vals[2 * e], D->sup_vals[2 * e + 1]);
                        br.srows.push_back(r);
                        br.scols.push_back(c);
                        br.svals.push_back(v);
                        br.dense[(size_t)(r + c * ml)] += v;
                    }
                } else {
                    const double *src = D->dense + 2 * D->br_dense_base[b];
                    for (int64_t e = 0; e < ml * mn; ++e) br.dense[(size_t)e] = cd(src[2 * e], src[2 * e + 1]);
                }
                brs.push_back(std::move(br));
            }
            F.ob.push_back(std::move(brs));
        }
        // trace functional needs sqrt(f); without a partition list use 1 (callers that
        // upload a fused plan pass their own trace functional to the solvers)
        P->compiled = false;
        finalize_plan(P.get());
        *out = P.release();
    });
}

pimi_status pimi_plan_prepare(pimi_plan *P) {
    return guarded([&] {
        if (!P) throw Error(PIMI_EINVAL, "null argument");
        if (!P->ctx) throw Error(PIMI_EINVAL, "host-only plan");
        HIPCHK(hipSetDevice(P->ctx->device));
        ensure_plan_preconditioner(P);
        ensure_plan_weighting(P);
    });
}

pimi_status pimi_plan_info_get(const pimi_plan *P, pimi_plan_info *info) {
    return guarded([&] {
        if (!P || !info) throw Error(PIMI_EINVAL, "null argument");
        info->n = P->n;
        info->nsectors = P->ns;
        info->max_block = 0;
        for (auto m : P->host.dims) info->max_block = std::max<int32_t>(info->max_block, (int32_t)m);
        info->nterms = (int64_t)P->st.terms.size();
        info->stencil_entries = P->st.entries_per_apply;
        info->table_bytes_c64 = P->main.table_bytes(16);
        info->table_bytes_c32 = P->main.table_bytes(8);
        info->apply_form = P->main.dia_ok ? 1 : 0;
        info->banded_groups = P->main.dia_ok ? (int32_t)P->main.h_chains.size() : 0;
    });
}

pimi_status pimi_plan_dims(const pimi_plan *P, int64_t *dims) {
    return guarded([&] {
        if (!P || !dims) throw Error(PIMI_EINVAL, "null argument");
        for (int s = 0; s < P->ns; ++s) dims[s] = P->host.dims[s];
    });
}

pimi_status pimi_trace_functional(const pimi_plan *P, int64_t *idx, double *vals) {
    return guarded([&] {
        if (!P || !idx || !vals) throw Error(PIMI_EINVAL, "null argument");
        if (!P->compiled) throw Error(PIMI_EINVAL, "trace functional needs a compiled model");
        std::copy(P->tidx.begin(), P->tidx.end(), idx);
        std::copy(P->tval.begin(), P->tval.end(), vals);
    });
}

pimi_status pimi_collective_operator(const pimi_plan *P, const double *X, double *out) {
    return guarded([&] {
        if (!P || !X || !out) throw Error(PIMI_EINVAL, "null argument");
        if (!P->compiled) throw Error(PIMI_EINVAL, "collective operators need a compiled model");
        int d = P->host.geom.d;
        std::vector<cd> Xc((size_t)d * d);
        for (int q = 0; q < d * d; ++q) Xc[q] = cd(X[2 * q], X[2 * q + 1]);
        collective_operator(P->host, Xc.data(), reinterpret_cast<cd *>(out));
    });
}

pimi_status pimi_iid_state(const pimi_plan *P, const double *rho, double *out) {
    return guarded([&] {
        if (!P || !rho || !out) throw Error(PIMI_EINVAL, "null argument");
        if (!P->compiled) throw Error(PIMI_EINVAL, "iid_state needs a plan compiled from a model");
        const int d = P->host.geom.d;
        std::vector<cd> A((size_t)d * d);
        for (int q = 0; q < d * d; ++q) {
            A[(size_t)q] = cd(rho[2 * q], rho[2 * q + 1]);
            if (!std::isfinite(rho[2 * q]) || !std::isfinite(rho[2 * q + 1]))
                throw Error(PIMI_EINVAL, "rho must be finite");
        }
        iid_state(P->host, A.data(), reinterpret_cast<cd *>(out));
    });
}

pimi_status pimi_plan_banded_export(const pimi_plan *P, int64_t *counts, int32_t *cbase, int64_t *chains,
                                    int32_t *ents, double *coef) {
    return guarded([&] {
        if (!P || !counts) throw Error(PIMI_EINVAL, "null argument");
        const DevStencil &S = P->main;
        if (!S.dia_ok) {
            counts[0] = counts[1] = counts[2] = counts[3] = 0;
            return;
        }
        counts[0] = (int64_t)S.h_cbase.size();
        counts[1] = (int64_t)S.h_chains.size();
        counts[2] = (int64_t)S.h_dents.size();
        counts[3] = (int64_t)S.h_coef.size();
        if (cbase) std::copy(S.h_cbase.begin(), S.h_cbase.end(), cbase);
        if (chains)
            for (size_t g = 0; g < S.h_chains.size(); ++g) {
                const DiaChain &G = S.h_chains[g];
                const int64_t v[7] = {G.in_off, G.mn, G.dA, G.c0, G.span, G.e0, G.e1};
                std::copy(v, v + 7, chains + 7 * g);
            }
        if (ents)
            for (size_t e = 0; e < S.h_dents.size(); ++e) {
                const DiaEntry &D = S.h_dents[e];
                const int32_t v[7] = {D.kind, D.dB, D.qa, D.qb, D.qa2, D.qb2, D.sh};
                std::copy(v, v + 7, ents + 7 * e);
            }
        if (coef)
            for (size_t k = 0; k < S.h_coef.size(); ++k) {
                coef[2 * k] = S.h_coef[k].real();
                coef[2 * k + 1] = S.h_coef[k].imag();
            }
    });
}

pimi_status pimi_plan_export_sizes(const pimi_plan *P, int64_t *sz) {
    return guarded([&] {
        if (!P || !sz) throw Error(PIMI_EINVAL, "null argument");
        const FusedPlan &F = P->host;
        int64_t nblk = 0, ncp = 0, nnz = 0;
        auto add = [&](const CSC &C) { nblk++; ncp += C.m + 1; nnz += (int64_t)C.rows.size(); };
        for (auto &C : F.H) add(C);
        for (auto &C : F.Q) add(C);
        for (auto &g : F.K)
            for (auto &C : g) add(C);
        int64_t nbr = 0, nsup = 0, ndense = 0;
        for (auto &g : F.ob)
            for (auto &b : g) {
                nbr++;
                if (b.use_support) nsup += (int64_t)b.svals.size();
                else ndense += (int64_t)b.dense.size();
            }
        sz[0] = nblk; sz[1] = ncp; sz[2] = nnz; sz[3] = (int64_t)F.cg_scale.size(); sz[4] = (int64_t)F.ob_scale.size();
        sz[5] = nbr; sz[6] = nsup; sz[7] = ndense; sz[8] = F.has_h; sz[9] = F.has_q; sz[10] = F.ns; sz[11] = F.n;
    });
}

pimi_status pimi_plan_export(const pimi_plan *P, int64_t *cp_base, int64_t *colptr, int32_t *rows, double *vals,
                             double *cg_scale, double *ob_scale, int64_t *ob_br_base, int32_t *br_out, int32_t *br_in,
                             double *br_r, int32_t *br_use_support, int64_t *br_sup_base, int32_t *sup_rows,
                             int32_t *sup_cols, double *sup_vals, int64_t *br_dense_base, double *dense) {
    return guarded([&] {
        if (!P) throw Error(PIMI_EINVAL, "null argument");
        const FusedPlan &F = P->host;
        int64_t blk = 0, cp = 0, nz = 0;
        auto put = [&](const CSC &C) {
            cp_base[blk++] = cp;
            for (int64_t c = 0; c <= C.m; ++c) colptr[cp++] = nz + C.colptr[c];
            for (size_t k = 0; k < C.rows.size(); ++k) {
                rows[nz] = C.rows[k];
                vals[2 * nz] = C.vals[k].real();
                vals[2 * nz + 1] = C.vals[k].imag();
                nz++;
            }
        };
        for (auto &C : F.H) put(C);
        for (auto &C : F.Q) put(C);
        for (auto &g : F.K)
            for (auto &C : g) put(C);
        for (size_t g = 0; g < F.cg_scale.size(); ++g) {
            cg_scale[2 * g] = F.cg_scale[g].real();
            cg_scale[2 * g + 1] = F.cg_scale[g].imag();
        }
        int64_t br = 0, sp = 0, dn = 0;
        for (size_t g = 0; g < F.ob.size(); ++g) {
            ob_scale[2 * g] = F.ob_scale[g].real();
            ob_scale[2 * g + 1] = F.ob_scale[g].imag();
            ob_br_base[g] = br;
            for (auto &b : F.ob[g]) {
                br_out[br] = b.out;
                br_in[br] = b.in;
                br_r[br] = b.r;
                br_use_support[br] = b.use_support;
                br_sup_base[br] = sp;
                br_dense_base[br] = dn;
                if (b.use_support) {
                    for (size_t k = 0; k < b.svals.size(); ++k) {
                        sup_rows[sp] = b.srows[k];
                        sup_cols[sp] = b.scols[k];
                        sup_vals[2 * sp] = b.svals[k].real();
                        sup_vals[2 * sp + 1] = b.svals[k].imag();
                        sp++;
                    }
                } else {
                    for (size_t k = 0; k < b.dense.size(); ++k) {
                        dense[2 * dn] = b.dense[k].real();
                        dense[2 * dn + 1] = b.dense[k].imag();
                        dn++;
                    }
                }
                br++;
            }
        }
        ob_br_base[F.ob.size()] = br;
        br_sup_base[br] = sp;
    });
}

pimi_status pimi_sector_preconditioner_factors(const pimi_plan *P, double *U, double *inv_den) {
    return guarded([&] {
        if (!P || !U || !inv_den) throw Error(PIMI_EINVAL, "null argument");
        std::vector<cd> u, inv;
        build_sector_preconditioner(P->host, u, inv);
        std::memcpy(U, u.data(), sizeof(cd) * u.size());
        std::memcpy(inv_den, inv.data(), sizeof(cd) * inv.size());
    });
}

void pimi_plan_destroy(pimi_plan *P) { delete P; }

pimi_status pimi_workspace_create(pimi_plan *P, int32_t precision, int32_t krylovdim, pimi_workspace **out) {
    return guarded([&] {
        if (!P || !out) throw Error(PIMI_EINVAL, "null argument");
        if (!P->ctx) throw Error(PIMI_EINVAL, "plan was compiled without a device context");
        if (precision != PIMI_C64 && precision != PIMI_C32) throw Error(PIMI_EINVAL, "unknown precision");
        if (krylovdim < 1 || krylovdim > kMaxKrylov) throw Error(PIMI_EINVAL, "krylovdim must lie in 1..4000");
        // budget check before allocating: basis + work vectors must fit in device memory
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipSetDevice(P->ctx->device));
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        double need = (double)P->n * 16.0 * (std::min<int64_t>(krylovdim, P->n) + 14) *
                      (precision == PIMI_C32 ? 1.5 : 1.0);
        if (need > 0.9 * (double)free_b) throw Error(PIMI_ENOMEM, "Krylov workspace exceeds free device memory");
        HIPCHK(hipSetDevice(P->ctx->device));
        std::unique_ptr<pimi_workspace> W(new pimi_workspace());
        W->plan = P;
        W->prec = precision;
        W->m = (int)std::min<int64_t>(krylovdim, P->n);
        const char *tr = std::getenv("PIMI_TRACE");
        W->trace_progress = tr && tr[0] == '1';
        // The inner solve launches eagerly by default: measured at N = 10..120,
        // the captured graph was 1-4 % slower per solve than the same kernels
        // launched one by one behind the outer loop's lookahead, which already
        // keeps the queue full.  PIMI_INNER_GRAPH=1 replays a captured graph.
        const char *ig = std::getenv("PIMI_INNER_GRAPH");
        const char *ng = std::getenv("PIMI_NO_GRAPH");
        W->use_graphs = ig && ig[0] == '1' && !(ng && ng[0] == '1');
        const char *fi = std::getenv("PIMI_INNER_FINISH");
        W->inner_finish = !(fi && fi[0] == '0');
        HIPCHK(hipStreamCreateWithFlags(&W->stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&W->ev0));
        HIPCHK(hipEventCreate(&W->ev1));
        HIPCHK(hipHostMalloc((void **)&W->hpin, 64 * sizeof(double), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void **)&W->hlook, 2 * ((size_t)W->m + 2) * sizeof(cx<double>), hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&W->look_ev[0], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&W->look_ev[1], hipEventDisableTiming));
        ensure_workspace(W.get());
        *out = W.release();
    });
}

void pimi_workspace_destroy(pimi_workspace *W) {
    if (!W) return;
    (void)hipSetDevice(W->plan->ctx->device);
    if (W->stream) (void)hipStreamSynchronize(W->stream);
    if (W->inner_graph) (void)hipGraphExecDestroy(W->inner_graph);
    if (W->hpin) (void)hipHostFree(W->hpin);
    if (W->hlook) (void)hipHostFree(W->hlook);
    if (W->dc_look) (void)hipHostFree(W->dc_look);
    if (W->ex_host) (void)hipHostFree(W->ex_host);
    for (hipEvent_t e : W->look_ev)
        if (e) (void)hipEventDestroy(e);
    if (W->ev0) (void)hipEventDestroy(W->ev0);
    if (W->ev1) (void)hipEventDestroy(W->ev1);
    if (W->stream) (void)hipStreamDestroy(W->stream);
    delete W;
}

static pimi_status apply_entry(pimi_workspace *W, const void *x, void *y, int64_t ncols, int32_t on_device,
                               bool adjoint) {
    return guarded([&] {
        if (!W || !x || !y) throw Error(PIMI_EINVAL, "null argument");
        if (ncols < 1) throw Error(PIMI_EINVAL, "ncols must be positive");
        if (x == y) throw Error(PIMI_EINVAL, "Liouvillian source and destination must not alias");
        BusyGuard guard(W->busy);
        HIPCHK(hipSetDevice(W->plan->ctx->device));
        if (adjoint) ensure_adjoint_stencil(W->plan);
        const int64_t n = W->plan->n;
        if (on_device) {
            HIPCHK(hipEventRecord(W->ev0, W->stream));
            if (W->prec == PIMI_C64)
                launch_apply<double>(W, (const cx<double> *)x, (cx<double> *)y, 1.0, ncols, false, false, adjoint);
            else
                launch_apply<float>(W, (const cx<float> *)x, (cx<float> *)y, 1.0, ncols, false, false, adjoint);
            HIPCHK(hipEventRecord(W->ev1, W->stream));
            sync(W);
        } else {
            DevBuf<cx<double>> dx, dy;
            dx.alloc((size_t)n * ncols);
            dy.alloc((size_t)n * ncols);
            HIPCHK(hipMemcpyAsync(dx.p, x, sizeof(cx<double>) * n * ncols, hipMemcpyHostToDevice, W->stream));
            if (W->prec == PIMI_C64) {
                HIPCHK(hipEventRecord(W->ev0, W->stream));
                launch_apply<double>(W, dx.p, dy.p, 1.0, ncols, false, false, adjoint);
                HIPCHK(hipEventRecord(W->ev1, W->stream));
            } else {
                DevBuf<cx<float>> fx, fy;
                fx.alloc((size_t)n * ncols);
                fy.alloc((size_t)n * ncols);
                hipLaunchKernelGGL((k_convert<float, double>), dim3(1024), dim3(256), 0, W->stream, fx.p, dx.p,
                                   n * ncols);
                HIPCHK(hipEventRecord(W->ev0, W->stream));
                launch_apply<float>(W, fx.p, fy.p, 1.0, ncols, false, false, adjoint);
                HIPCHK(hipEventRecord(W->ev1, W->stream));
                hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(256), 0, W->stream, dy.p, fy.p,
                                   n * ncols);
                sync(W);
            }
            HIPCHK(hipMemcpyAsync(y, dy.p, sizeof(cx<double>) * n * ncols, hipMemcpyDeviceToHost, W->stream));
            sync(W);
        }
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, W->ev0, W->ev1));
        W->last_apply_ms = ms;
    });
}

pimi_status pimi_apply(pimi_workspace *W, const void *x, void *y, int64_t ncols, int32_t on_device) {
    return apply_entry(W, x, y, ncols, on_device, false);
}

pimi_status pimi_apply_adjoint(pimi_workspace *W, const void *x, void *y, int64_t ncols, int32_t on_device) {
    return apply_entry(W, x, y, ncols, on_device, true);
}

pimi_status pimi_workspace_autotune_apply(pimi_workspace *W, int64_t ncols, int32_t reps, int32_t *widths_out) {
    return guarded([&] {
        if (!W || !widths_out || ncols < 1 || reps < 1) throw Error(PIMI_EINVAL, "bad argument");
        BusyGuard guard(W->busy);
        pimi_plan *P = W->plan;
        HIPCHK(hipSetDevice(P->ctx->device));
        widths_out[0] = widths_out[1] = 0;
        if (!P->main.dia_ok) return;
        const int pi = W->prec == PIMI_C64 ? 0 : 1;
        const size_t es = W->prec == PIMI_C64 ? sizeof(cx<double>) : sizeof(cx<float>);
        const int64_t n = P->n;
        DevBuf<char> dx, dy;
        dx.alloc(es * (size_t)(n * ncols));
        dy.alloc(es * (size_t)(n * ncols));
        // finite synthetic input: every byte 0x3f (FP64 ~0.49, FP32 ~0.75)
        HIPCHK(hipMemsetAsync(dx.p, 0x3f, es * (size_t)(n * ncols), W->stream));
        auto time_width = [&](int w, int64_t nc) {
            auto run = [&]() {
                if (W->prec == PIMI_C64)
                    launch_dia<double>(W->stream, P, P->main, false, (const cx<double> *)dx.p, (cx<double> *)dy.p,
                                       1.0, nc, w);
                else
                    launch_dia<float>(W->stream, P, P->main, false, (const cx<float> *)dx.p, (cx<float> *)dy.p, 1.0,
                                      nc, w);
            };
            run();  // warm
            HIPCHK(hipEventRecord(W->ev0, W->stream));
            for (int r = 0; r < reps; ++r) run();
            HIPCHK(hipEventRecord(W->ev1, W->stream));
            sync(W);
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, W->ev0, W->ev1));
            return ms;
        };
        for (int pass = 0; pass < 2; ++pass) {
            const int64_t nc = pass == 0 ? 1 : ncols;
            int best = 0;
            float best_ms = INFINITY;
            for (int w = 0; w < kNumDiaWidths; ++w) {
                const float ms = time_width(w, nc);
                if (ms < best_ms) {
                    best_ms = ms;
                    best = w;
                }
            }
            {
                std::lock_guard<std::recursive_mutex> lk(P->lazy_mu);
                if (pass == 0) P->dia_w_single[pi] = best;
                else P->dia_w_multi[pi] = best;
            }
            widths_out[pass] = kDiaWidths[best];
        }
    });
}

pimi_status pimi_apply_bench(pimi_workspace *W, int64_t ncols, int32_t reps, double *ms_avg) {
    return guarded([&] {
        if (!W || !ms_avg || ncols < 1 || reps < 1) throw Error(PIMI_EINVAL, "bad argument");
        BusyGuard guard(W->busy);
        HIPCHK(hipSetDevice(W->plan->ctx->device));
        const int64_t n = W->plan->n;
        const size_t elems = (size_t)(n * ncols);
        uint64_t st = 0x51ca1eULL;
        std::vector<cx<double>> h(elems);
        probe_vector(st, h);
        DevBuf<cx<double>> dx, dy;
        dx.alloc(elems);
        dy.alloc(elems);
        h2d(W, dx.p, h.data(), sizeof(cx<double>) * elems);
        DevBuf<cx<float>> fx, fy;
        if (W->prec == PIMI_C32) {
            fx.alloc(elems);
            fy.alloc(elems);
            hipLaunchKernelGGL((k_convert<float, double>), dim3(1024), dim3(256), 0, W->stream, fx.p, dx.p,
                               (int64_t)elems);
        }
        auto run = [&]() {
            if (W->prec == PIMI_C64) launch_apply<double>(W, dx.p, dy.p, 1.0, ncols);
            else launch_apply<float>(W, fx.p, fy.p, 1.0, ncols);
        };
        run();  // warm
        sync(W);
        HIPCHK(hipEventRecord(W->ev0, W->stream));
        for (int r = 0; r < reps; ++r) run();
        HIPCHK(hipEventRecord(W->ev1, W->stream));
        sync(W);
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, W->ev0, W->ev1));
        *ms_avg = (double)ms / reps;
    });
}

pimi_status pimi_copy_bench(pimi_workspace *W, int64_t bytes, int32_t reps, double *ms_avg) {
    return guarded([&] {
        if (!W || !ms_avg || bytes < 1 || reps < 1) throw Error(PIMI_EINVAL, "bad argument");
        BusyGuard guard(W->busy);
        HIPCHK(hipSetDevice(W->plan->ctx->device));
        DevBuf<char> a, b;
        a.alloc((size_t)bytes);
        b.alloc((size_t)bytes);
        HIPCHK(hipMemsetAsync(a.p, 1, (size_t)bytes, W->stream));
        HIPCHK(hipMemcpyAsync(b.p, a.p, (size_t)bytes, hipMemcpyDeviceToDevice, W->stream));
        sync(W);
        HIPCHK(hipEventRecord(W->ev0, W->stream));
        for (int r = 0; r < reps; ++r)
            HIPCHK(hipMemcpyAsync(b.p, a.p, (size_t)bytes, hipMemcpyDeviceToDevice, W->stream));
        HIPCHK(hipEventRecord(W->ev1, W->stream));
        sync(W);
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, W->ev0, W->ev1));
        *ms_avg = (double)ms / reps;
    });
}

pimi_status pimi_precond_apply(pimi_workspace *W, const double *x, double *y, double scale) {
    return guarded([&] {
        if (!W || !x || !y) throw Error(PIMI_EINVAL, "null argument");
        BusyGuard guard(W->busy);
        HIPCHK(hipSetDevice(W->plan->ctx->device));
        ensure_preconditioner(W);
        const int64_t n = W->plan->n;
        h2d(W, W->tmp.p, x, sizeof(cx<double>) * n);
        if (W->prec == PIMI_C32) {
            DevBuf<cx<float>> a, b, t1, t2;
            a.alloc(n); b.alloc(n); t1.alloc(n); t2.alloc(n);
            hipLaunchKernelGGL((k_convert<float, double>), dim3(1024), dim3(256), 0, W->stream, a.p, W->tmp.p, n);
            apply_precond_t<float>(W, a.p, b.p, scale, t1.p, t2.p);
            hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(256), 0, W->stream, W->w.p, b.p, n);
        } else {
            apply_precond(W, W->tmp.p, W->w.p, scale);
        }
        d2h(W, y, W->w.p, sizeof(cx<double>) * n);
    });
}

pimi_status pimi_precond_bench(pimi_workspace *W, int32_t reps, double *ms_avg, double *flops) {
    return guarded([&] {
        if (!W || reps < 1 || !ms_avg || !flops) throw Error(PIMI_EINVAL, "invalid argument");
        BusyGuard guard(W->busy);
        HIPCHK(hipSetDevice(W->plan->ctx->device));
        ensure_preconditioner(W);
        const int64_t n = W->plan->n;
        HIPCHK(hipMemsetAsync(W->w.p, 0x3f, sizeof(cx<double>) * n, W->stream));
        auto once = [&] {
            if (W->prec == PIMI_C32)
                apply_precond_t<float>(W, (const cx<float> *)W->w.p, (cx<float> *)W->zbuf.p, 1.0,
                                       (cx<float> *)W->pt1.p, (cx<float> *)W->pt2.p);
            else
                apply_precond_t<double>(W, W->w.p, W->zbuf.p, 1.0, W->pt1.p, W->pt2.p);
        };
        once();  // warm-up
        HIPCHK(hipEventRecord(W->ev0, W->stream));
        for (int32_t r = 0; r < reps; ++r) once();
        HIPCHK(hipEventRecord(W->ev1, W->stream));
        HIPCHK(hipEventSynchronize(W->ev1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, W->ev0, W->ev1));
        *ms_avg = ms / reps;
        double m3 = 0;
        for (auto m : W->plan->host.dims) m3 += (double)m * m * m;
        *flops = 4 * (W->plan->u_real ? 4.0 : 8.0) * m3;  // complex-by-real products: 4 flops per complex entry fma
    });
}

pimi_status pimi_workspace_last_apply_ms(const pimi_workspace *W, double *ms) {
    return guarded([&] {
        if (!W || !ms) throw Error(PIMI_EINVAL, "null argument");
        *ms = W->last_apply_ms;
    });
}

pimi_status pimi_steady_gmres(pimi_workspace *W, const pimi_gmres_opts *opts, const double *x0, double *rho_out,
                              pimi_solve_info *info) {
    return guarded([&] {
        if (!W || !opts || !rho_out || !info) throw Error(PIMI_EINVAL, "null argument");
        if (!W->plan->compiled) throw Error(PIMI_EINVAL, "stationary solves need a compiled model");
        BusyGuard guard(W->busy);
        HIPCHK(hipSetDevice(W->plan->ctx->device));
        steady_solve(W, opts, x0, rho_out, info);
    });
}

pimi_status pimi_expv_series(pimi_workspace *W, const double *rho0, const double *ts, int64_t nts,
                             const pimi_expv_opts *opts, int32_t nobs, const double *obs, double *obs_out,
                             double *rho_final, pimi_solve_info *info) {
    return guarded([&] {
        if (!W || !rho0 || !ts || !opts || !info || nts < 1) throw Error(PIMI_EINVAL, "bad argument");
        if (nobs < 0 || (nobs > 0 && (!obs || !obs_out))) throw Error(PIMI_EINVAL, "bad observable arguments");
        if (nobs > kMaxObservables) throw Error(PIMI_EINVAL, "at most 100 observables per call");
        BusyGuard guard(W->busy);
        HIPCHK(hipSetDevice(W->plan->ctx->device));
        expv_series(W, rho0, ts, nts, opts, nobs, obs, obs_out, rho_final, info);
    });
}

pimi_status pimi_trajectory_channels(const pimi_plan *P, int32_t *count) {
    return guarded([&] {
        if (!P || !count) throw Error(PIMI_EINVAL, "null argument");
        *count = (int32_t)P->host.channels.size();
    });
}

pimi_status pimi_conditional_action(pimi_workspace *W, const void *X, void *Y, int64_t ncols, double *intensities,
                                    int32_t on_device) {
    return guarded([&] {
        if (!W || !X || !Y || !intensities) throw Error(PIMI_EINVAL, "null argument");
        if (X == Y) throw Error(PIMI_EINVAL, "batched trajectory source and destination must not alias");
        BusyGuard guard(W->busy);
        traj_prologue(W, ncols);
        TrajIO io{W, W->plan->n * ncols, on_device != 0};
        const cx<double> *dx = io.dev(X, W->tX);
        cx<double> *dy = on_device ? (cx<double> *)Y : W->tY.p;
        conditional_action(W, dx, dy, ncols, W->tLam.p);
        std::vector<cx<double>> h((size_t)ncols);
        HIPCHK(hipMemcpyAsync(h.data(), W->tLam.p, sizeof(cx<double>) * ncols, hipMemcpyDeviceToHost, W->stream));
        io.back(Y, dy);
        sync(W);
        check_intensities(h, "total jump intensity");
        for (int64_t c = 0; c < ncols; ++c) intensities[c] = std::max(0.0, h[(size_t)c].x);
    });
}

pimi_status pimi_conditional_action_noalloc(pimi_workspace *W, const void *X, void *Y, int64_t ncols,
                                            double *intensities) {
    return guarded([&] {
        if (!W || !X || !Y || !intensities) throw Error(PIMI_EINVAL, "null argument");
        if (X == Y) throw Error(PIMI_EINVAL, "batched trajectory source and destination must not alias");
        if (ncols < 1 || ncols > kNoAllocMaxCols)
            throw Error(PIMI_EINVAL, "pimi_conditional_action_noalloc takes 1..256 columns (its fixed scratch size)");
        if (W->prec != PIMI_C64) throw Error(PIMI_EINVAL, "trajectory operations need a complex FP64 workspace");
        BusyGuard guard(W->busy);
        pimi_plan *P = W->plan;
        HIPCHK(hipSetDevice(P->ctx->device));
        {
            std::lock_guard<std::recursive_mutex> lk(P->lazy_mu);
            if (!P->traj_ready)
                throw Error(PIMI_EINVAL, "the plan's trajectory tables are not built yet (building them allocates): "
                                         "call pimi_conditional_action or pimi_channel_intensities on it first");
        }
        cx<double> *base = (cx<double> *)P->ctx->noalloc_scratch;
        if (!base) throw Error(PIMI_EDEVICE, "no scratch address for this device");
        cx<double> *partial = base, *h = base + kDotBlocks * kNoAllocMaxCols, *lam = h + kNoAllocMaxCols;
        cx<double> *status = lam + kNoAllocMaxCols;
        conditional_action(W, (const cx<double> *)X, (cx<double> *)Y, ncols, lam, h, partial);
        hipLaunchKernelGGL(k_lam_clamp_check, dim3(1), dim3(256), 0, W->stream, lam, ncols, intensities, status);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(W->hpin, status, sizeof(double), hipMemcpyDeviceToHost, W->stream));  // pinned
        sync(W);
        if (W->hpin[0] != 0.0) throw Error(PIMI_EINVAL, "total jump intensity is nonfinite or has a negative trace");
    });
}

pimi_status pimi_conditional_rk4(pimi_workspace *W, void *X, int64_t ncols, double h, int32_t on_device) {
    return guarded([&] {
        if (!W || !X) throw Error(PIMI_EINVAL, "null argument");
        if (!(h > 0) || !std::isfinite(h)) throw Error(PIMI_EINVAL, "batched trajectory step must be positive");
        BusyGuard guard(W->busy);
        traj_prologue(W, ncols);
        pimi_plan *P = W->plan;
        const int64_t len = P->n * ncols;
        TrajIO io{W, len, on_device != 0};
        cx<double> *x = io.dev(X, W->tX);
        cx<double> *tmp = W->tTmp.p, *k1 = W->tK1.p, *k2 = W->tK2.p, *lam = W->tLam.p;
        const dim3 g(grid_for(len)), b(kRedBlock);
        // the stage sequence of batched_conditional_rk4! (src/batched_trajectories.jl:335-368)
        conditional_action(W, x, k1, ncols, lam);
        HIPCHK(hipMemcpyAsync(k2, k1, sizeof(cx<double>) * len, hipMemcpyDeviceToDevice, W->stream));
        hipLaunchKernelGGL(k_lincomb, g, b, 0, W->stream, tmp, x, k1, h / 2, len);
        conditional_action(W, tmp, k1, ncols, lam + ncols);
        hipLaunchKernelGGL(k_lincomb, g, b, 0, W->stream, k2, k2, k1, 2.0, len);
        hipLaunchKernelGGL(k_lincomb, g, b, 0, W->stream, tmp, x, k1, h / 2, len);
        conditional_action(W, tmp, k1, ncols, lam + 2 * ncols);
        hipLaunchKernelGGL(k_lincomb, g, b, 0, W->stream, k2, k2, k1, 2.0, len);
        hipLaunchKernelGGL(k_lincomb, g, b, 0, W->stream, tmp, x, k1, h, len);
        conditional_action(W, tmp, k1, ncols, lam + 3 * ncols);
        hipLaunchKernelGGL(k_lincomb, g, b, 0, W->stream, k2, k2, k1, 1.0, len);
        hipLaunchKernelGGL(k_lincomb, g, b, 0, W->stream, x, x, k2, h / 6, len);
        hipLaunchKernelGGL(k_cols_trace, dim3((unsigned)ncols), dim3(256), 0, W->stream, x, P->n, P->d_tidx.p,
                           P->d_tval.p, (int64_t)P->tidx.size(), W->tZ.p);
        std::vector<cx<double>> z((size_t)ncols), hl((size_t)(4 * ncols));
        HIPCHK(hipMemcpyAsync(z.data(), W->tZ.p, sizeof(cx<double>) * ncols, hipMemcpyDeviceToHost, W->stream));
        HIPCHK(hipMemcpyAsync(hl.data(), lam, sizeof(cx<double>) * 4 * ncols, hipMemcpyDeviceToHost, W->stream));
        sync(W);
        check_intensities(hl, "total jump intensity");
        for (auto &v : z)
            if (!(std::abs(cd(v.x, v.y)) > 2.220446049250313e-16))
                throw Error(PIMI_EINVAL, "conditional state acquired zero trace");
        hipLaunchKernelGGL(k_cols_div, g, b, 0, W->stream, x, W->tZ.p, P->n, ncols);
        io.back(X, x);
        sync(W);
    });
}

pimi_status pimi_channel_intensities(pimi_workspace *W, const void *X, int64_t ncols, double *rates,
                                     int32_t on_device) {
    return guarded([&] {
        if (!W || !X || !rates) throw Error(PIMI_EINVAL, "null argument");
        BusyGuard guard(W->busy);
        traj_prologue(W, ncols);
        pimi_plan *P = W->plan;
        const int64_t n = P->n, nch = (int64_t)P->host.channels.size();
        TrajIO io{W, n * ncols, on_device != 0};
        const cx<double> *x = io.dev(X, W->tX);
        std::vector<cx<double>> h((size_t)ncols);
        for (int64_t k = 0; k < nch; ++k) {
            cols_intensity(W, x, ncols, P->ebar.p + (size_t)(k + 1) * (size_t)n, W->tH.p);
            HIPCHK(hipMemcpyAsync(h.data(), W->tH.p, sizeof(cx<double>) * ncols, hipMemcpyDeviceToHost, W->stream));
            sync(W);
            check_intensities(h, "jump intensity");
            for (int64_t c = 0; c < ncols; ++c) rates[k + c * nch] = std::max(0.0, h[(size_t)c].x);
        }
    });
}

pimi_status pimi_apply_jumps(pimi_workspace *W, void *X, int64_t ncols, const int32_t *channels, int32_t on_device) {
    return guarded([&] {
        if (!W || !X || !channels) throw Error(PIMI_EINVAL, "null argument");
        BusyGuard guard(W->busy);
        traj_prologue(W, ncols);
        pimi_plan *P = W->plan;
        const int64_t n = P->n, nch = (int64_t)P->host.channels.size();
        for (int64_t c = 0; c < ncols; ++c)
            if (channels[c] < 0 || channels[c] > nch) throw Error(PIMI_EINVAL, "selected jump channel out of range");
        TrajIO io{W, n * ncols, on_device != 0};
        cx<double> *x = io.dev(X, W->tX);
        const size_t col = sizeof(cx<double>) * (size_t)n;
        for (int64_t k = 1; k <= nch; ++k) {
            std::vector<int64_t> sel;
            for (int64_t c = 0; c < ncols; ++c)
                if (channels[c] == k) sel.push_back(c);
            if (sel.empty()) continue;
            const int64_t cnt = (int64_t)sel.size();
            for (int64_t q = 0; q < cnt; ++q)  // gather the cohort
                HIPCHK(hipMemcpyAsync(W->tTmp.p + q * n, x + sel[(size_t)q] * n, col, hipMemcpyDeviceToDevice,
                                      W->stream));
            launch_stencil(W, *P->gain[(size_t)(k - 1)], W->tTmp.p, W->tK1.p, cnt);
            hipLaunchKernelGGL(k_cols_trace, dim3((unsigned)cnt), dim3(256), 0, W->stream, W->tK1.p, n,
                               P->d_tidx.p, P->d_tval.p, (int64_t)P->tidx.size(), W->tZ.p);
            std::vector<cx<double>> z((size_t)cnt);
            HIPCHK(hipMemcpyAsync(z.data(), W->tZ.p, sizeof(cx<double>) * cnt, hipMemcpyDeviceToHost, W->stream));
            sync(W);
            for (auto &v : z)
                if (!(std::abs(cd(v.x, v.y)) > 2.220446049250313e-16))
                    throw Error(PIMI_EINVAL, "selected jump has zero probability");
            hipLaunchKernelGGL(k_cols_div, dim3(grid_for(n * cnt)), dim3(kRedBlock), 0, W->stream, W->tK1.p, W->tZ.p,
                               n, cnt);
            for (int64_t q = 0; q < cnt; ++q)  // scatter back
                HIPCHK(hipMemcpyAsync(x + sel[(size_t)q] * n, W->tK1.p + q * n, col, hipMemcpyDeviceToDevice,
                                      W->stream));
        }
        io.back(X, x);
        sync(W);
    });
}

// ---------------------------------------------------------------------------
// Arnoldi spectrum (src/krylov.jl:914-961) and fixed-step RK4 evolution
// (src/evolution.jl:75-118)
// ---------------------------------------------------------------------------
pimi_status pimi_arnoldi(pimi_workspace *W, const double *q0, int32_t m, double breakfactor, double *H,
                         int32_t *k_out) {
    return guarded([&] {
        if (!W || !q0 || !H || !k_out) throw Error(PIMI_EINVAL, "null argument");
        if (W->prec != PIMI_C64) throw Error(PIMI_EINVAL, "the Arnoldi spectrum needs an FP64 workspace");
        if (m < 1 || m > W->m) throw Error(PIMI_EINVAL, "Arnoldi dimension must lie in 1..workspace krylovdim");
        if (!(breakfactor >= 0)) throw Error(PIMI_EINVAL, "breakfactor must be nonnegative");
        BusyGuard guard(W->busy);
        HIPCHK(hipSetDevice(W->plan->ctx->device));
        const int64_t n = W->plan->n;
        cx<double> *V = W->V.p;
        h2d(W, W->tmp.p, q0, sizeof(cx<double>) * n);
        dev_norm<double>(W, W->tmp.p, n, W->nscal.p);
        const double nq = read_scalar(W, W->nscal.p);
        if (!(nq > 0) || !std::isfinite(nq)) throw Error(PIMI_EINVAL, "initial_vector must be nonzero and finite");
        hipLaunchKernelGGL(k_scale_div<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, V, W->tmp.p, n,
                           W->nscal.p);
        std::memset(H, 0, sizeof(cx<double>) * (size_t)(m + 1) * m);
        std::vector<cx<double>> hcol((size_t)m + 1);
        int k = 0;
        for (int j = 0; j < m; ++j) {
            launch_apply<double>(W, V + (int64_t)j * n, W->w.p, 1.0);
            dev_norm<double>(W, W->w.p, n, W->nscal.p + 1);  // image norm |L v_j|
            HIPCHK(hipMemsetAsync(W->hacc.p, 0, sizeof(cx<double>) * (j + 1), W->stream));
            dev_cgs2<double>(W, V, n, j + 1, W->w.p, n, W->hacc.p);
            hipLaunchKernelGGL(k_reduce_norm, dim3(1), dim3(kRedBlock), 0, W->stream, W->npartial.p, kNormBlocks,
                               W->nscal.p + 2);
            HIPCHK(hipMemcpyAsync(hcol.data(), W->hacc.p, sizeof(cx<double>) * (j + 1), hipMemcpyDeviceToHost,
                                  W->stream));
            HIPCHK(hipMemcpyAsync(W->hpin, W->nscal.p + 1, 2 * sizeof(double), hipMemcpyDeviceToHost, W->stream));
            sync(W);
            const double image = W->hpin[0], hnext = W->hpin[1];
            for (int i = 0; i <= j; ++i) {
                H[2 * ((size_t)i + (size_t)j * (m + 1))] = hcol[(size_t)i].x;
                H[2 * ((size_t)i + (size_t)j * (m + 1)) + 1] = hcol[(size_t)i].y;
            }
            H[2 * ((size_t)(j + 1) + (size_t)j * (m + 1))] = hnext;
            k = j + 1;
            if (j + 1 == m || std::abs(hnext) <= breakfactor * std::max(image, DBL_MIN)) break;
            hipLaunchKernelGGL(k_scale_div<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream,
                               V + (int64_t)(j + 1) * n, W->w.p, n, W->nscal.p + 2);
        }
        sync(W);
        *k_out = k;
    });
}

pimi_status pimi_arnoldi_vectors(pimi_workspace *W, int32_t k, const double *Y, int32_t ncols, double *out) {
    return guarded([&] {
        if (!W || !Y || !out) throw Error(PIMI_EINVAL, "null argument");
        if (W->prec != PIMI_C64) throw Error(PIMI_EINVAL, "the Arnoldi spectrum needs an FP64 workspace");
        if (k < 1 || k > W->m) throw Error(PIMI_EINVAL, "k must lie in 1..workspace krylovdim");
        if (ncols < 0) throw Error(PIMI_EINVAL, "ncols must be nonnegative");
        BusyGuard guard(W->busy);
        HIPCHK(hipSetDevice(W->plan->ctx->device));
        const int64_t n = W->plan->n;
        for (int32_t c = 0; c < ncols; ++c) {
            h2d(W, W->coef.p, Y + 2 * (size_t)c * k, sizeof(cx<double>) * k);
            hipLaunchKernelGGL(k_gemv<double>, dim3(grid_for(n)), dim3(kRedBlock), 0, W->stream, W->tmp.p, W->V.p, n,
                               k, W->coef.p, n, 1.0, 1);
            d2h(W, out + 2 * (size_t)c * n, W->tmp.p, sizeof(cx<double>) * n);
        }
    });
}

pimi_status pimi_evolve_rk4(pimi_workspace *W, const double *x0, double t0, double t1, int64_t steps, double *x_out) {
    return guarded([&] {
        if (!W || !x0 || !x_out) throw Error(PIMI_EINVAL, "null argument");
        if (W->prec != PIMI_C64) throw Error(PIMI_EINVAL, "RK4 evolution needs an FP64 workspace");
        if (steps <= 0) throw Error(PIMI_EINVAL, "steps must be positive");
        if (!std::isfinite(t0) || !std::isfinite(t1)) throw Error(PIMI_EINVAL, "tspan must be finite");
        BusyGuard guard(W->busy);
        HIPCHK(hipSetDevice(W->plan->ctx->device));
        const int64_t n = W->plan->n;
        const double h = (t1 - t0) / (double)steps;
        cx<double> *x = W->cur.p, *tmp = W->tmp.p, *k1 = W->w.p, *k2 = W->trial.p;
        h2d(W, x, x0, sizeof(cx<double>) * n);
        const dim3 g(grid_for(n)), b(kRedBlock);
        for (int64_t s = 0; s < steps; ++s) {
            launch_apply<double>(W, x, k1, 1.0);
            hipLaunchKernelGGL(k_rk4_stage<double>, g, b, 0, W->stream, 0, x, tmp, k2, k1, n, h / 2);
            launch_apply<double>(W, tmp, k1, 1.0);
            hipLaunchKernelGGL(k_rk4_stage<double>, g, b, 0, W->stream, 1, x, tmp, k2, k1, n, h / 2);
            launch_apply<double>(W, tmp, k1, 1.0);
            hipLaunchKernelGGL(k_rk4_stage<double>, g, b, 0, W->stream, 1, x, tmp, k2, k1, n, h);
            launch_apply<double>(W, tmp, k1, 1.0);
            hipLaunchKernelGGL(k_rk4_stage<double>, g, b, 0, W->stream, 2, x, tmp, k2, k1, n, h / 6);
        }
        HIPCHK(hipGetLastError());
        d2h(W, x_out, x, sizeof(cx<double>) * n);
    });
}

}  // extern "C"
