// Internal declarations shared by the libpimi translation units: error
// plumbing, device buffers, the context and the device-resident plan.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <complex>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pimi.h"
#include "pimi_kernels.hpp"
#include "pimi_lower.hpp"

namespace pimi {

constexpr int kMaxKrylov = 4000;
constexpr int kColChunk = 256;  // basis columns staged through LDS per pass

// thread-local message behind pimi_last_error()
inline thread_local std::string g_last_error;


#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw Error(PIMI_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t count = 0;
    void alloc(size_t n) {
        release();
        if (n == 0) n = 1;
        HIPCHK(hipMalloc((void **)&p, n * sizeof(T)));
        count = n;
    }
    void upload(const T *h, size_t n, hipStream_t s = 0) {
        alloc(n);
        HIPCHK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        count = 0;
    }
    ~DevBuf() { release(); }
};

// A stencil's device tables (FP64 and FP32 values).
struct DevStencil {
    DevBuf<int64_t> term_base;
    DevBuf<StencilTermDev> terms;
    DevBuf<int32_t> col;
    DevBuf<cx<double>> val64;
    DevBuf<cx<float>> val32;
    void upload(const Stencil &st) {
        std::vector<StencilTermDev> t(st.terms.size());
        for (size_t k = 0; k < t.size(); ++k) {
            const StencilTerm &a = st.terms[k];
            t[k] = StencilTermDev{a.in_sector, a.flags, a.wa, a.wb, a.a_base, a.b_base, a.s_re, a.s_im};
        }
        std::vector<cx<double>> v64(st.val.size());
        std::vector<cx<float>> v32(st.val.size());
        for (size_t k = 0; k < v64.size(); ++k) {
            v64[k] = cx<double>{st.val[k].real(), st.val[k].imag()};
            v32[k] = cx<float>{(float)st.val[k].real(), (float)st.val[k].imag()};
        }
        term_base.upload(st.term_base.data(), st.term_base.size());
        terms.upload(t.data(), t.size());
        col.upload(st.col.data(), st.col.size());
        val64.upload(v64.data(), v64.size());
        val32.upload(v32.data(), v32.size());
    }
};

}  // namespace pimi

struct pimi_context {
    int device = 0;
};

struct pimi_plan {
    pimi_context *ctx = nullptr;
    pimi::FusedPlan host;
    pimi::Stencil st;
    bool compiled = false;  // true: geometry available (collective operators)
    int64_t n = 0;
    int ns = 0;
    int64_t ntiles = 0, tiles_per_xcd = 0;        // 256-coordinate tiles (sector scaling)
    int64_t natiles = 0, atiles_per_xcd = 0;      // single-column apply tiles (one wave each)
    int64_t natiles4 = 0;                         // kApplyCols-column apply tiles (batched launches)
    std::vector<int64_t> tidx;
    std::vector<double> tval;
    pimi::DevBuf<int64_t> offs, term_base, d_tidx;
    pimi::DevBuf<int32_t> dims, tile_sector, col;
    pimi::DevBuf<pimi::ApplyTile> atiles, atiles4;
    pimi::DevBuf<pimi::StencilTermDev> terms;
    pimi::DevBuf<pimi::cx<double>> val64;
    pimi::DevBuf<pimi::cx<float>> val32;
    pimi::DevBuf<double> d_tval;
    // stationary-solve preconditioner factors (built on first use)
    bool prec_ready = false;
    pimi::DevBuf<pimi::cx<double>> pU64, pInv64;
    pimi::DevBuf<pimi::cx<float>> pU32, pInv32;
    pimi::DevBuf<pimi::GemmTile> gtiles;
    int64_t ngtiles = 0;
    pimi::DevBuf<pimi::GemmTile> mtiles;  // 16x16 tiles for the matrix-core GEMM
    int64_t nmtiles = 0;
    pimi::DevBuf<pimi::MTile> xtiles;  // the same tiles grouped by XCD, padded with m = 0
    int64_t nxtiles = 0;
    // sector-block-diagonal part of L (inner solver), built on first use
    bool diag_ready = false;
    pimi::DevBuf<int64_t> dg_term_base;
    pimi::DevBuf<pimi::StencilTermDev> dg_terms;
    pimi::DevBuf<int32_t> dg_col;
    pimi::DevBuf<pimi::cx<double>> dg_val64;
    pimi::DevBuf<pimi::cx<float>> dg_val32;
    // conditional (no-jump) trajectory tables (built on first use): the
    // H/Q-only stencil, one gain stencil per jump channel, and conj of the
    // intensity weights (n complex each; the total first, then per channel)
    bool traj_ready = false;
    pimi::DevStencil cond;
    std::vector<std::unique_ptr<pimi::DevStencil>> gain;
    pimi::DevBuf<pimi::cx<double>> ebar;
    // stencil of the adjoint Liouvillian (built on first use)
    bool adj_ready = false;
    pimi::DevBuf<int64_t> adj_term_base;
    pimi::DevBuf<pimi::StencilTermDev> adj_terms;
    pimi::DevBuf<int32_t> adj_col;
    pimi::DevBuf<pimi::cx<double>> adj_val64;
    pimi::DevBuf<pimi::cx<float>> adj_val32;
    // sector weighting u = S x, S = sqrt(f^nu) on sector nu (built on first use):
    // stencil terms of S L S^{-1} (scale times sigma_out / sigma_in), the
    // transformed trace functional (ones on the block diagonals) and 1/sigma
    bool weight_ready = false;
    std::vector<double> sigma;  // per sector
    pimi::DevBuf<pimi::StencilTermDev> terms_w;
    pimi::DevBuf<double> d_ones, d_inv_sigma;
};

namespace pimi {

inline void set_error(const std::string &m) { g_last_error = m; }

template <typename F>
inline pimi_status guarded(F &&f) {
    try {
        f();
        return PIMI_OK;
    } catch (const Error &e) {
        set_error(e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return PIMI_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return PIMI_EINVAL;
    }
}

struct BusyGuard {
    std::atomic<int> &f;
    explicit BusyGuard(std::atomic<int> &flag) : f(flag) {
        int expected = 0;
        if (!f.compare_exchange_strong(expected, 1))
            throw Error(PIMI_EBUSY, "a pimi workspace cannot be used concurrently");
    }
    ~BusyGuard() { f.store(0); }
};

inline int grid_for(int64_t n) {
    int64_t g = (n + kRedBlock - 1) / kRedBlock;
    return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 2048);
}

// Complex Givens rotation [c s; -conj(s) c] with c real, mapping (a, b) to
// (r, 0) (the convention of src/krylov.jl:710-716).
inline void complex_givens(cd a, cd b, double &c, cd &s) {
    if (b == cd(0, 0)) { c = 1.0; s = 0; return; }
    if (a == cd(0, 0)) { c = 0.0; s = 1.0; return; }
    double scale = std::abs(a) + std::abs(b);
    cd as = a / scale, bs = b / scale;
    double normab = scale * std::sqrt(std::norm(as) + std::norm(bs));
    cd alpha = a / std::abs(a);
    c = std::abs(a) / normab;
    s = alpha * std::conj(b) / normab;
}

// Plan-level preconditioner factors and GEMM tile lists (built once, on
// first use by any solver).
void ensure_plan_preconditioner(pimi_plan *P);
// Plan-level sector-weighting tables (built once, on first use).
void ensure_plan_weighting(pimi_plan *P);

}  // namespace pimi
