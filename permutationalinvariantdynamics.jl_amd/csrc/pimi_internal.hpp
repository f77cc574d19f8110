// Internal declarations shared by the libpimi translation units: error
// plumbing, device buffers, the context and the device-resident plan.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <complex>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
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
        build_dia(st, dims, offs);
    }
    // Banded form (k_apply_dia): built when every factor is a sum of
    // diagonals and each (input sector, row shift) group's column shifts
    // span at most three diagonals; dia_ok false otherwise (the ELL entries
    // above are then the only form).  Groups of one input sector whose row
    // shifts lie within kDiaShift and whose columns fit one strip of TB + 2
    // share a chain (one strip load per tile).
    bool dia_ok = false;
    std::vector<int32_t> h_cbase;  // ns + 1: chains of each output sector
    std::vector<int32_t> h_pid;    // ns: compiled program of each output sector (-1: interpreted)
    std::vector<DiaChain> h_chains;
    std::vector<DiaEntry> h_dents;
    std::vector<cd> h_coef;
    std::vector<int32_t> h_dent_sector, h_dent_in;  // per entry: output and input sector (weighting)
    DevBuf<int32_t> cbase;
    DevBuf<DiaChain> chains;
    DevBuf<int32_t> pid;  // per output sector: compiled program (-1: interpreted)
    // compiled banded programs: the structure signature (empty: none, the
    // tables are interpreted), the programs' source, and the kernels
    // compiled so far, [0: single-vector, 1: batched][0: FP64, 1: FP32][tile width]
    std::string prog_sig, prog_src;
    int prog_device = -1;
    mutable std::atomic<hipFunction_t> prog_fn[2][2][kNumDiaWidths] = {};
    DiaTables tables() const { return DiaTables{cbase.p, chains.p, dents.p, pid.p}; }
    DevBuf<DiaEntry> dents;
    DevBuf<cx<double>> coef64, coefw64;
    DevBuf<cx<float>> coef32, coefw32;
    bool dia_w_ready = false;

    void build_dia(const Stencil &st, const std::vector<int64_t> &dims, const std::vector<int64_t> &offs) {
        dia_ok = false;
        // PIMI_APPLY_FORM=ell keeps every plan on the ELL tables (a second,
        // independent kernel for the same operator)
        if (const char *f = std::getenv("PIMI_APPLY_FORM"))
            if (std::string(f) == "ell") return;
        h_cbase.assign(1, 0);
        h_chains.clear();
        h_dents.clear();
        h_coef.clear();
        h_dent_sector.clear();
        h_dent_in.clear();
        h_dims = dims;
        const int ns = (int)dims.size();
        // diagonals of one factor: d -> coefficient per output row (m rows)
        auto diagonals = [&](bool identity, int64_t base, int32_t w, int64_t m) {
            std::map<int, std::vector<cd>> D;
            if (identity) {
                D[0].assign((size_t)m, cd(1.0, 0.0));
                return D;
            }
            for (int64_t a = 0; a < m; ++a)
                for (int k = 0; k < w; ++k) {
                    const size_t q = (size_t)(base + (int64_t)k * m + a);
                    if (st.val[q] == cd(0.0, 0.0)) continue;
                    auto &v = D[st.col[q] - (int)a];
                    if (v.empty()) v.assign((size_t)m, cd(0.0, 0.0));
                    v[(size_t)a] += st.val[q];
                }
            return D;
        };
        struct PendingEntry {
            int kind, dB;
            std::vector<cd> ca, wb, ca2, wb2;
        };
        for (int q = 0; q < ns; ++q) {
            const int64_t m = dims[(size_t)q];
            // groups of this sector in first-appearance order
            std::vector<std::pair<int, int>> keys;
            std::vector<std::vector<PendingEntry>> pend;
            for (int64_t t = st.term_base[(size_t)q]; t < st.term_base[(size_t)q + 1]; ++t) {
                const StencilTerm &T = st.terms[(size_t)t];
                const cd s(T.s_re, T.s_im);
                const bool ia = T.flags & 1, ib = T.flags & 2;
                auto DA = diagonals(ia, T.a_base, T.wa, m);
                auto DB = diagonals(ib, T.b_base, T.wb, m);
                for (auto &pa : DA)
                    for (auto &pb : DB) {
                        PendingEntry E;
                        E.dB = pb.first;
                        E.ca.assign((size_t)m, cd(0.0, 0.0));
                        E.wb.assign((size_t)m, cd(0.0, 0.0));
                        if (ib) {  // left (both identity included): scale on the row side
                            E.kind = 0;
                            for (int64_t a = 0; a < m; ++a) E.ca[(size_t)a] = s * pa.second[(size_t)a];
                        } else if (ia) {
                            E.kind = 1;
                            for (int64_t b = 0; b < m; ++b) E.wb[(size_t)b] = s * std::conj(pb.second[(size_t)b]);
                        } else {
                            E.kind = 2;
                            bool real = true;
                            for (int64_t a = 0; a < m; ++a) {
                                E.ca[(size_t)a] = s * pa.second[(size_t)a];
                                E.wb[(size_t)a] = std::conj(pb.second[(size_t)a]);
                                real = real && E.ca[(size_t)a].imag() == 0.0 && E.wb[(size_t)a].imag() == 0.0;
                            }
                            if (real) E.kind = 3;
                        }
                        const std::pair<int, int> key(T.in_sector, pa.first);
                        size_t g = 0;
                        while (g < keys.size() && keys[g] != key) ++g;
                        if (g == keys.size()) {
                            keys.push_back(key);
                            pend.emplace_back();
                        }
                        pend[g].push_back(std::move(E));
                    }
            }
            struct GroupOut {  // one (input sector, row shift) group, entries ready
                int64_t in_off;
                int32_t mn, dA, c0, span, nu;
                std::vector<DiaEntry> ents;
            };
            std::vector<GroupOut> outs;
            for (size_t g = 0; g < keys.size(); ++g) {
                int lo = INT32_MAX, hi = INT32_MIN;
                for (auto &E : pend[g]) {
                    lo = std::min(lo, E.dB);
                    hi = std::max(hi, E.dB);
                }
                if (hi - lo > 2) return;  // not banded enough: ELL form only
                const int nu = keys[g].first;
                GroupOut O{offs[(size_t)nu], (int32_t)dims[(size_t)nu], keys[g].second, lo, hi - lo, nu, {}};
                // every coefficient vector starts on a kDiaCoefPad boundary and
                // is zero-padded to a multiple of it, so a tile reads its TB
                // column coefficients as one unclamped contiguous (scalar) load
                auto put = [&](const std::vector<cd> &v) {
                    const int32_t q = (int32_t)h_coef.size();
                    h_coef.insert(h_coef.end(), v.begin(), v.end());
                    h_coef.resize((size_t)q + (v.size() + kDiaCoefPad - 1) / kDiaCoefPad * kDiaCoefPad,
                                  cd(0.0, 0.0));
                    return q;
                };
                // a real sandwich keeps real parts only, two per complex slot
                // (element i of the real vector at real index i of the slot array)
                auto packed = [](const std::vector<cd> &v) {
                    std::vector<cd> p((v.size() + 1) / 2, cd(0.0, 0.0));
                    for (size_t i = 0; i < v.size(); ++i)
                        p[i / 2] = i % 2 ? cd(p[i / 2].real(), v[i].real()) : cd(v[i].real(), p[i / 2].imag());
                    return p;
                };
                // specialised kinds (arithmetic only; every output still sums
                // the group's entries in one fixed order for every tile width):
                // purely imaginary left/right coefficients, a complex left and
                // right entry on one strip column merged, and real sandwiches
                // on one strip column paired
                std::vector<PendingEntry> &P = pend[g];
                auto imag_only = [](const std::vector<cd> &v) {
                    for (const cd &z : v)
                        if (z.real() != 0.0) return false;
                    return true;
                };
                for (auto &E : P) {
                    if (E.kind == 0 && imag_only(E.ca)) {
                        E.kind = 4;
                        for (cd &z : E.ca) z = cd(z.imag(), 0.0);
                    } else if (E.kind == 1 && imag_only(E.wb)) {
                        E.kind = 5;
                        for (cd &z : E.wb) z = cd(z.imag(), 0.0);
                    }
                }
                std::vector<PendingEntry> merged;
                std::vector<bool> used(P.size(), false);
                for (size_t i = 0; i < P.size(); ++i) {
                    if (used[i]) continue;
                    PendingEntry E = P[i];
                    used[i] = true;
                    const int want = E.kind == 0 ? 1 : E.kind == 1 ? 0 : E.kind == 3 ? 3 : -1;
                    if (want >= 0)
                        for (size_t k = i + 1; k < P.size(); ++k) {
                            if (used[k] || P[k].kind != want || P[k].dB != E.dB) continue;
                            used[k] = true;
                            if (E.kind == 3) {  // pair: the second sandwich in ca2/wb2
                                E.kind = 7;
                                E.ca2 = P[k].ca;
                                E.wb2 = P[k].wb;
                            } else {            // left + right: the left's row vector, the right's column vector
                                if (E.kind == 1) E.ca = P[k].ca;
                                else E.wb = P[k].wb;
                                E.kind = 6;
                            }
                            break;
                        }
                    merged.push_back(std::move(E));
                }
                for (auto &E : merged) {
                    DiaEntry D{E.kind, E.dB - lo, 0, 0, 0, 0, 0, 0};
                    const bool real_rows = E.kind == 3 || E.kind == 4 || E.kind == 7;
                    const bool real_cols = E.kind == 3 || E.kind == 5 || E.kind == 7;
                    if (E.kind != 1 && E.kind != 5) D.qa = put(real_rows ? packed(E.ca) : E.ca);
                    if (E.kind != 0 && E.kind != 4) D.qb = put(real_cols ? packed(E.wb) : E.wb);
                    if (E.kind == 7) {
                        D.qa2 = put(packed(E.ca2));
                        D.qb2 = put(packed(E.wb2));
                    }
                    if (h_coef.size() >= (size_t)INT32_MAX) return;
                    O.ents.push_back(D);
                }
                outs.push_back(std::move(O));
            }
            // chains: the sector's groups ordered by (input block, row shift),
            // first-appearance order among equals; a chain takes the next
            // group while it reads the same block, its row shift stays within
            // kDiaShift of the chain's first and the column union stays
            // within TB + 2
            std::vector<size_t> order(outs.size());
            for (size_t k = 0; k < order.size(); ++k) order[k] = k;
            std::stable_sort(order.begin(), order.end(), [&](size_t u, size_t v) {
                return outs[u].in_off != outs[v].in_off ? outs[u].in_off < outs[v].in_off : outs[u].dA < outs[v].dA;
            });
            for (size_t i = 0; i < order.size();) {
                const GroupOut &F = outs[order[i]];
                int lo = F.c0, hi = F.c0 + F.span;
                size_t j = i + 1;
                for (; j < order.size(); ++j) {
                    const GroupOut &H = outs[order[j]];
                    const int nlo = std::min(lo, H.c0), nhi = std::max(hi, H.c0 + H.span);
                    if (H.in_off != F.in_off || H.dA - F.dA > kDiaShift || nhi - nlo > 2) break;
                    lo = nlo;
                    hi = nhi;
                }
                DiaChain C{F.in_off, F.mn, F.dA, lo, hi - lo, (int32_t)h_dents.size(), 0};
                for (size_t k = i; k < j; ++k) {
                    const GroupOut &H = outs[order[k]];
                    for (DiaEntry D : H.ents) {
                        D.dB += H.c0 - lo;  // strip column relative to the chain's first
                        D.sh = H.dA - F.dA;
                        h_dents.push_back(D);
                        h_dent_sector.push_back(q);
                        h_dent_in.push_back(H.nu);
                    }
                }
                C.e1 = (int32_t)h_dents.size();
                h_chains.push_back(C);
                i = j;
            }
            h_cbase.push_back((int32_t)h_chains.size());
        }
        if (h_coef.empty()) h_coef.push_back(cd(0.0, 0.0));
        dia_ok = true;
    }
    void push_dia() {
        if (!dia_ok) return;
        std::vector<cx<double>> c64(h_coef.size());
        std::vector<cx<float>> c32(h_coef.size());
        for (size_t k = 0; k < h_coef.size(); ++k) {
            c64[k] = cx<double>{h_coef[k].real(), h_coef[k].imag()};
            c32[k] = cx<float>{(float)h_coef[k].real(), (float)h_coef[k].imag()};
        }
        cbase.upload(h_cbase.data(), h_cbase.size());
        if (h_pid.size() + 1 != h_cbase.size()) h_pid.assign(h_cbase.size() - 1, -1);
        pid.upload(h_pid.data(), h_pid.size());
        if (h_chains.empty()) chains.alloc(1);  // no entries: the kernel reads no chain
        else chains.upload(h_chains.data(), h_chains.size());
        if (h_dents.empty()) dents.alloc(1);
        else dents.upload(h_dents.data(), h_dents.size());
        coef64.upload(c64.data(), c64.size());
        coef32.upload(c32.data(), c32.size());
    }
    // coefficients of S L S^{-1}: every entry's scaled side (rows for left
    // and sandwich, columns for right) times sigma_out / sigma_in
    void push_dia_weighted(const std::vector<double> &sigma) {
        if (!dia_ok || dia_w_ready) return;
        std::vector<cd> w(h_coef);
        for (size_t e = 0; e < h_dents.size(); ++e) {
            const double r = sigma[(size_t)h_dent_sector[e]] / sigma[(size_t)h_dent_in[e]];
            if (r == 1.0) continue;
            const DiaEntry &D = h_dents[e];
            const size_t m = (size_t)dims_of_entry(e);
            auto scale = [&](int32_t q, bool packed_reals) {  // packed reals: two per slot
                const size_t len = packed_reals ? (m + 1) / 2 : m;
                for (size_t k = 0; k < len; ++k) w[(size_t)q + k] *= r;
            };
            switch (D.kind) {  // the side that carries the scale
                case 1: scale(D.qb, false); break;
                case 5: scale(D.qb, true); break;
                case 3: case 4: scale(D.qa, true); break;
                case 6: scale(D.qa, false); scale(D.qb, false); break;  // (ca + w): both sides
                case 7: scale(D.qa, true); scale(D.qa2, true); break;
                default: scale(D.qa, false); break;
            }
        }
        std::vector<cx<double>> c64(w.size());
        std::vector<cx<float>> c32(w.size());
        for (size_t k = 0; k < w.size(); ++k) {
            c64[k] = cx<double>{w[k].real(), w[k].imag()};
            c32[k] = cx<float>{(float)w[k].real(), (float)w[k].imag()};
        }
        coefw64.upload(c64.data(), c64.size());
        coefw32.upload(c32.data(), c32.size());
        dia_w_ready = true;
    }
    std::vector<int64_t> h_dims;  // sector sizes (banded coefficient vector lengths)
    int64_t dims_of_entry(size_t e) const { return h_dims[(size_t)h_dent_sector[e]]; }
    int64_t dia_table_bytes(int elem) const {
        return (int64_t)(h_chains.size() * sizeof(DiaChain) + h_dents.size() * sizeof(DiaEntry) +
                         h_cbase.size() * 4) +
               (int64_t)h_coef.size() * elem;
    }

    void push() {
        std::vector<cx<double>> v64(h_val.size());
        std::vector<cx<float>> v32(h_val.size());
        for (size_t k = 0; k < h_val.size(); ++k) {
            v64[k] = cx<double>{h_val[k].real(), h_val[k].imag()};
            v32[k] = cx<float>{(float)h_val[k].real(), (float)h_val[k].imag()};
        }
        push_dia();
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
        push_dia_weighted(sigma);
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
        if (dia_ok) return dia_table_bytes(elem);
        return (int64_t)(h_entries.size() * sizeof(ApplyEntry) + h_entry_base.size() * 4 + h_col.size() * 4 +
                         (h_entry_base.size() - 1) * 4) +
               (int64_t)h_val.size() * elem;
    }
};

}  // namespace pimi

struct pimi_context {
    int device = 0;
    // device address of the library's fixed scratch array on this device
    // (hipGetSymbolAddress at creation); used by the allocation-free entry
    // points only
    void *noalloc_scratch = nullptr;
};

struct pimi_plan {
    pimi_context *ctx = nullptr;
    // serializes the tables built on first use (preconditioner, weighting,
    // sector-diagonal, adjoint and trajectory stencils) and the autotuner's
    // width updates, so workspaces of one plan on several threads build each
    // table once and never replace a buffer another stream reads
    std::recursive_mutex lazy_mu;
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
    // banded-apply tiles (k_apply_dia) of kDiaWidths[w] columns, and the
    // width index chosen for single-vector and for multi-column launches,
    // per precision (set by pimi_workspace_autotune_apply, or the defaults)
    pimi::DevBuf<pimi::ApplyTile> dtiles[pimi::kNumDiaWidths];
    int64_t ndtiles[pimi::kNumDiaWidths] = {};
    int dia_w_single[2] = {0, 0}, dia_w_multi[2] = {1, 1};  // [0]: FP64 launches, [1]: FP32
    pimi::DevStencil main;  // the Liouvillian's apply tables
    pimi::DevBuf<double> d_tval;
    // stationary-solve preconditioner factors (built on first use)
    bool prec_ready = false;
    pimi::DevBuf<pimi::cx<double>> pU64, pInv64;
    pimi::DevBuf<pimi::cx<float>> pU32, pInv32;
    // every U entry real (a real Hamiltonian block): the GEMMs then take U as a
    // real operand, two MFMAs per complex-by-real product instead of three
    bool u_real = false;
    pimi::DevBuf<double> pUr64;
    pimi::DevBuf<float> pUr32;
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
    pimi::DevBuf<double> d_ones, d_inv_sigma, d_sigma;
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

// Delayed-reorthogonalization Arnoldi step on one complex FP64 vector (the
// batch engine's kernels, one job).  With Q = V[:, 0:j] orthonormal, p =
// V[:, j] and w given, one pass over [Q, p] forms mail[0:j+1] = [Q p]^H p and
// mail[j+1:2j+2] = [Q p]^H w, then slot[10] = rho = sqrt(p^H p - |Q^H p|^2),
// slot[11] = 1/rho, slot[12:14] = e = (p^H w - s^H t)/rho^2, and a second
// pass writes V[:, j] = (p - Q s)/rho and V[:, j+1] = (w - e p - Q(t - e s))/rho.
// Descriptors for j = 0 .. m-1 are built once on the host (dcgs_single_descs,
// dcgs_single_desc_bytes) and uploaded; partial holds 2 (m + 1) nblk, mail
// 2 (m + 1), h 2 m complex doubles, slot 16 doubles.  look (pinned host
// memory, two regions of look_stride complex doubles) receives step j's mail
// and scalars in region j mod 2, written by the step's coefficient kernel:
// mail entries first, the scalars as doubles from entry 2 (m + 1).
int dcgs_single_nblk(int64_t n);
size_t dcgs_single_desc_bytes(int m);
void dcgs_single_descs(void *host, int m, void *V, const void *w, void *partial, void *mail, void *h, double *slot,
                       int64_t n, cx<double> *look, int64_t look_stride);
void dcgs_single_step(hipStream_t st, const void *desc_dev, int j, int64_t n);
// The same steps for a Krylov exponential's factorization, every step
// reporting in a region of its own so that all of them run before one read
// back: region j (stride complex doubles from mail) holds the step's scalars
// as doubles (rho at kArnSlotRho, e at kArnSlotE, kArnSlotE + 1) in its first
// kArnMailHead complex entries, then [Q p]^H p (j + 1 entries) and [Q p]^H w
// (j + 1 entries).  Descriptors (dcgs_single_desc_bytes(steps) bytes) are
// stepped with dcgs_single_step; partial holds 2 steps nblk, h 2 steps
// complex doubles.
constexpr int kArnMailHead = 9, kArnSlotRho = 10, kArnSlotE = 12;
void dcgs_arnoldi_descs(void *host, int steps, void *V, const void *w, void *partial, cx<double> *mail,
                        int64_t stride, void *h, int64_t n);

// Set up the banded programs of stencil S (of a plan on device `device`)
// when wanted: per-sector program ids and the programs' source; kernels are
// compiled on first use (dia_program).  No-op for stencils on the ELL form.
void dia_setup_programs(DevStencil &S, int device);
// The compiled kernel of S's programs for (form 0 single-vector / 1 batched,
// precision 0 FP64 / 1 FP32, tile width index w), compiled now if need be;
// null when S has no programs or `stream` is capturing a graph and the
// kernel is not compiled yet (the caller interprets the tables).
hipFunction_t dia_program(const DevStencil &S, int form, int prec, int w, hipStream_t stream);

// Launch the banded apply of y = out_scale * S x on ncols vectors through
// S's compiled programs (true) or report that it has none for this
// precision and width (false).
template <typename R>
bool launch_dia_program(hipStream_t stream, const DevStencil &S, int w, const cx<R> *x, cx<R> *y, int64_t n,
                        const ApplyTile *tiles, int64_t nt, int64_t per_xcd, const int64_t *offs, const int32_t *dims,
                        const cx<R> *coef, R out_scale, int64_t ncols, dim3 grid, dim3 block) {
    hipFunction_t f = dia_program(S, 0, sizeof(R) == 8 ? 0 : 1, w, stream);
    if (!f) return false;
    DiaTables T = S.tables();
    void *args[] = {(void *)&x,    (void *)&y,    (void *)&n, (void *)&tiles, (void *)&nt,        (void *)&per_xcd,
                    (void *)&offs, (void *)&dims, (void *)&T, (void *)&coef,  (void *)&out_scale, (void *)&ncols};
    HIPCHK(hipModuleLaunchKernel(f, grid.x, grid.y, grid.z, block.x, block.y, block.z, 0, stream, args, nullptr));
    return true;
}

}  // namespace pimi
