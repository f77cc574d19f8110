// Internal declarations shared by the libpimi translation units: error
// plumbing, device buffers, the context and the device-resident plan.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
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

// A stencil's apply tables: per output sector a list of ApplyEntry (padded to
// whole chunks), and the ELL slot tables with the identity slot appended
// (FP64 and FP32 values).  build() lowers a host Stencil; push() uploads.
// entries_w holds the same entries with the sector-weighted scales
// s * sigma_out / sigma_in (built by push_weighted when first needed).
struct DevStencil {
    std::vector<int32_t> h_entry_base;  // ns + 1
    std::vector<ApplyEntry> h_entries;
    std::vector<int32_t> h_col;
    std::vector<cd> h_val;
    DevBuf<int32_t> entry_base;
    DevBuf<ApplyEntry> entries, entries_w;
    DevBuf<int32_t> col;
    DevBuf<cx<double>> val64;
    DevBuf<cx<float>> val32;

    void build(const Stencil &st, const std::vector<int64_t> &dims, const std::vector<int64_t> &offs) {
        int64_t maxm = 1;
        for (int64_t m : dims) maxm = std::max(maxm, m);
        const int64_t id = (int64_t)st.col.size();
        if (id + maxm >= (int64_t)INT32_MAX) throw Error(PIMI_EINVAL, "stencil tables exceed 32-bit indexing");
        h_col = st.col;
        h_val = st.val;
        for (int64_t r = 0; r < maxm; ++r) {
            h_col.push_back((int32_t)r);
            h_val.push_back(cd(1.0, 0.0));
        }
        h_entries.clear();
        h_entry_base.assign(1, 0);
        const int ns = (int)dims.size();
        for (int q = 0; q < ns; ++q) {
            const int64_t m = dims[(size_t)q];
            for (int64_t t = st.term_base[(size_t)q]; t < st.term_base[(size_t)q + 1]; ++t) {
                const StencilTerm &T = st.terms[(size_t)t];
                std::vector<int32_t> qa, qb;
                if (T.flags & 1) qa.push_back((int32_t)id);
                else
                    for (int k = 0; k < T.wa; ++k) qa.push_back((int32_t)(T.a_base + k * m));
                if (T.flags & 2) qb.push_back((int32_t)id);
                else
                    for (int k = 0; k < T.wb; ++k) qb.push_back((int32_t)(T.b_base + k * m));
                for (int32_t b : qb)
                    for (int32_t a : qa)
                        h_entries.push_back(ApplyEntry{offs[(size_t)T.in_sector], a, b,
                                                       (int32_t)dims[(size_t)T.in_sector], T.in_sector, T.s_re,
                                                       T.s_im});
            }
            while ((h_entries.size() - (size_t)h_entry_base.back()) % kEntryChunk)
                h_entries.push_back(ApplyEntry{offs[(size_t)q], (int32_t)id, (int32_t)id, (int32_t)m, q, 0.0, 0.0});
            if (h_entries.size() >= (size_t)INT32_MAX) throw Error(PIMI_EINVAL, "stencil entry list too long");
            h_entry_base.push_back((int32_t)h_entries.size());
        }
    }
    void push() {
        std::vector<cx<double>> v64(h_val.size());
        std::vector<cx<float>> v32(h_val.size());
        for (size_t k = 0; k < h_val.size(); ++k) {
            v64[k] = cx<double>{h_val[k].real(), h_val[k].imag()};
            v32[k] = cx<float>{(float)h_val[k].real(), (float)h_val[k].imag()};
        }
        entry_base.upload(h_entry_base.data(), h_entry_base.size());
        entries.upload(h_entries.data(), h_entries.size());
        col.upload(h_col.data(), h_col.size());
        val64.upload(v64.data(), v64.size());
        val32.upload(v32.data(), v32.size());
    }
    void upload(const Stencil &st, const std::vector<int64_t> &dims, const std::vector<int64_t> &offs) {
        build(st, dims, offs);
        push();
    }
    void push_weighted(const std::vector<double> &sigma) {
        std::vector<ApplyEntry> w(h_entries);
        for (size_t q = 0; q + 1 < h_entry_base.size(); ++q)
            for (int32_t k = h_entry_base[q]; k < h_entry_base[q + 1]; ++k) {
                const double r = sigma[q] / sigma[(size_t)w[(size_t)k].in_sector];
                w[(size_t)k].s_re *= r;
                w[(size_t)k].s_im *= r;
            }
        entries_w.upload(w.data(), w.size());
    }
    // bytes of the tables one apply reads (entries, slot tables, sector sizes)
    int64_t table_bytes(int elem) const {
        return (int64_t)(h_entries.size() * sizeof(ApplyEntry) + h_entry_base.size() * 4 + h_col.size() * 4 +
                         (h_entry_base.size() - 1) * 4) +
               (int64_t)h_val.size() * elem;
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
    pimi::DevBuf<int64_t> offs, d_tidx;
    pimi::DevBuf<int32_t> dims, tile_sector;
    pimi::DevBuf<pimi::ApplyTile> atiles, atiles4;
    pimi::DevStencil main;  // the Liouvillian's apply tables
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
    pimi::DevStencil diag;
    // conditional (no-jump) trajectory tables (built on first use): the
    // H/Q-only stencil, one gain stencil per jump channel, and conj of the
    // intensity weights (n complex each; the total first, then per channel)
    bool traj_ready = false;
    pimi::DevStencil cond;
    std::vector<std::unique_ptr<pimi::DevStencil>> gain;
    pimi::DevBuf<pimi::cx<double>> ebar;
    // stencil of the adjoint Liouvillian (built on first use)
    bool adj_ready = false;
    pimi::DevStencil adj;
    // sector weighting u = S x, S = sqrt(f^nu) on sector nu (built on first use):
    // entries of S L S^{-1} (main.entries_w: scale times sigma_out / sigma_in),
    // the transformed trace functional (ones on the block diagonals) and 1/sigma
    bool weight_ready = false;
    std::vector<double> sigma;  // per sector
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
