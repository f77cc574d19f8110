// Host lowering: exact combinatorics, one-box Clebsch--Gordan geometry,
// fused static plan and device stencil tables.
//
// Reference algorithms restated here (paths relative to the reference root):
//   partitions / GT patterns / content   src/partitions.jl:29-45, src/gtpatterns.jl:1-106
//   one-box CG coefficient               src/cgc.jl:6-39
//   sqrt of an exact ratio               src/partitions.jl:445-484
//   branching weight |l| f^mu / f^l      src/partitions.jl:519-541
//   one-body geometry                    src/geometry.jl:405-432,465-490,519-562
//   collective sparse blocks             src/geometry.jl:820-865
//   compile_term + static fusion         src/liouvillian.jl:270-326,1290-1371,1543-1604,1700-1738
//   trace functional                     src/source_protocol.jl:15-34
//
// Exact rationals use checked __int128 arithmetic; a value outside the
// exactly handled range raises PIMI_ERANGE instead of being rounded early.
// Summation orders of every contraction, collective block and fused block
// follow the reference so that the lowered tables are bitwise those of the
// reference's lowering (the tests compare them against the oracle's).
#include "pimi_lower.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>

#include "../../include/pimi.h"

namespace pimi {

// ---------------------------------------------------------------------------
// checked exact rationals
// ---------------------------------------------------------------------------
typedef __int128 i128;

static i128 iabs(i128 x) { return x < 0 ? -x : x; }
static i128 igcd(i128 a, i128 b) {
    a = iabs(a); b = iabs(b);
    while (b) { i128 t = a % b; a = b; b = t; }
    return a;
}
static i128 mulchk(i128 a, i128 b) {
    i128 r;
    if (__builtin_mul_overflow(a, b, &r)) throw Error(PIMI_ERANGE, "exact rational overflow in PI lowering");
    return r;
}

struct Rat {
    i128 p = 0, q = 1;
    Rat() {}
    Rat(i128 pp, i128 qq = 1) {
        if (qq == 0) throw Error(PIMI_EINVAL, "singular CG product");
        if (qq < 0) { pp = -pp; qq = -qq; }
        i128 g = igcd(pp, qq);
        if (g > 1) { pp /= g; qq /= g; }
        p = pp; q = qq;
    }
};
static Rat operator*(const Rat &a, const Rat &b) {
    i128 g1 = igcd(a.p, b.q), g2 = igcd(b.p, a.q);
    if (g1 == 0) g1 = 1;
    if (g2 == 0) g2 = 1;
    return Rat(mulchk(a.p / g1, b.p / g2), mulchk(a.q / g2, b.q / g1));
}

// Correctly rounded p/q: exact when both fit in 53 bits (IEEE division of
// exact operands).
static double rn(const Rat &r) {
    const i128 lim = (i128)1 << 53;
    if (r.p == 0) return 0.0;
    if (iabs(r.p) >= lim || r.q >= lim)
        throw Error(PIMI_ERANGE, "exact ratio beyond the exactly rounded range");
    return (double)r.p / (double)r.q;
}
static double sqrt_exact(const Rat &r) {
    if (r.p < 0) throw Error(PIMI_EINVAL, "negative exact ratio under square root");
    return std::sqrt(rn(r));
}

// ---------------------------------------------------------------------------
// big unsigned integers (hook-length dimensions exceed 128 bits at N >= 130)
// ---------------------------------------------------------------------------
struct BigU {
    std::vector<uint32_t> w{1};
    void mul(uint32_t f) {
        uint64_t c = 0;
        for (auto &x : w) { uint64_t t = (uint64_t)x * f + c; x = (uint32_t)t; c = t >> 32; }
        if (c) w.push_back((uint32_t)c);
    }
    int bits() const {
        int top = (int)w.size() - 1;
        while (top > 0 && w[top] == 0) --top;
        uint32_t x = w[top];
        int b = 0;
        while (x) { ++b; x >>= 1; }
        return top * 32 + b;
    }
    int bit(int k) const { return (k / 32 < (int)w.size()) ? (w[k / 32] >> (k % 32)) & 1 : 0; }
    double to_double_rn() const {  // round to nearest, ties to even
        int nb = bits();
        if (nb <= 53) {
            double v = 0;
            for (int k = nb - 1; k >= 0; --k) v = v * 2 + bit(k);
            return v;
        }
        uint64_t m = 0;
        for (int k = nb - 1; k >= nb - 53; --k) m = (m << 1) | bit(k);
        int guard = bit(nb - 54);
        bool sticky = false;
        for (int k = nb - 55; k >= 0 && !sticky; --k) sticky = bit(k);
        if (guard && (sticky || (m & 1))) ++m;
        return std::ldexp((double)m, nb - 53);
    }
};

// f^p = N! / prod(hooks), multiplied out from prime exponents.
static BigU symmetric_group_dimension(const Partition &p) {
    int n = 0;
    for (int x : p) n += x;
    std::vector<int> ex(n + 2, 0);  // exponent of each integer factor, then factor primes
    auto add = [&](int v, int s) { ex[v] += s; };
    for (int k = 2; k <= n; ++k) add(k, 1);
    int D = (int)p.size();
    for (int i = 0; i < D; ++i)
        for (int j = 1; j <= p[i]; ++j) {
            int below = 0;
            for (int k = i + 1; k < D; ++k) below += p[k] >= j;
            int h = p[i] - j + below + 1;
            if (h >= 2) add(h, -1);
        }
    // reduce composite factors to primes
    std::vector<int> pe(n + 2, 0);
    for (int v = 2; v <= n; ++v) {
        if (!ex[v]) continue;
        int x = v;
        for (int q = 2; q * q <= x; ++q)
            while (x % q == 0) { pe[q] += ex[v]; x /= q; }
        if (x > 1) pe[x] += ex[v];
    }
    BigU r;
    for (int q = 2; q <= n; ++q) {
        if (pe[q] < 0) throw Error(PIMI_EINVAL, "internal hook-length error");
        for (int k = 0; k < pe[q]; ++k) r.mul((uint32_t)q);
    }
    return r;
}

double sqrt_symmetric_group_dimension(const Partition &p) {
    return std::sqrt(symmetric_group_dimension(p).to_double_rn());
}

// ---------------------------------------------------------------------------
// partitions and GT patterns
// ---------------------------------------------------------------------------
static std::vector<Partition> partitions(int N, int d) {
    std::vector<Partition> out;
    Partition pre;
    std::function<void(int, int)> rec = [&](int left, int cap) {
        if ((int)pre.size() == d) { if (left == 0) out.push_back(pre); return; }
        for (int x = std::min(cap, left); x >= 0; --x) { pre.push_back(x); rec(left - x, x); pre.pop_back(); }
    };
    rec(N, N);
    return out;
}

static std::vector<int> removable_corners(const Partition &p) {
    std::vector<int> r;
    int D = (int)p.size();
    for (int i = 0; i < D; ++i)
        if (p[i] > (i == D - 1 ? 0 : p[i + 1])) r.push_back(i);
    return r;
}

static inline int gidx(int i, int k) { return k * (k - 1) / 2 + i - 1; }  // one-based (i,k)

static std::vector<Pattern> gt_patterns(const Partition &p) {
    int D = (int)p.size();
    std::vector<Pattern> out;
    std::vector<std::vector<int>> rows{p};
    std::function<void()> descend = [&]() {
        if ((int)rows.size() == D) {
            Pattern g;
            for (int r = D - 1; r >= 0; --r) g.insert(g.end(), rows[r].begin(), rows[r].end());
            out.push_back(g);
            return;
        }
        const std::vector<int> up = rows.back();
        int k = (int)up.size() - 1;
        std::vector<int> v;
        std::function<void(int)> rec = [&](int i) {
            if (i > k) { rows.push_back(v); descend(); rows.pop_back(); return; }
            for (int x = up[i - 1]; x >= up[i]; --x) { v.push_back(x); rec(i + 1); v.pop_back(); }
        };
        rec(1);
    };
    descend();
    std::sort(out.begin(), out.end());
    return out;
}

static std::vector<int> content(const Pattern &g, int D) {
    std::vector<int> c(D);
    for (int K = 1; K <= D; ++K) {
        int v = 0;
        for (int i = 1; i <= K; ++i) v += g[gidx(i, K)];
        for (int i = 1; i < K; ++i) v -= g[gidx(i, K - 1)];
        c[K - 1] = v;
    }
    return c;
}

static int top_weight(const Pattern &g, int D) {
    int v = 0;
    for (int i = 1; i <= D; ++i) v += g[gidx(i, D)];
    return v;
}

// one-box CG coefficient <mu; j | lam> (src/cgc.jl:20-39)
static double cgc(const Pattern &mu, int j, const Pattern &lam, int D) {
    if (top_weight(lam, D) != top_weight(mu, D) + 1) return 0.0;
    std::vector<int> tau;
    for (int k = 1; k <= D; ++k) {
        int ones = 0, pos = -1;
        bool ok = true;
        for (int i = 1; i <= k; ++i) {
            int dif = lam[gidx(i, k)] - mu[gidx(i, k)];
            if (k <= j) { if (dif != 0) return 0.0; continue; }
            if (dif == 1) { ++ones; pos = i; }
            else if (dif != 0) ok = false;
        }
        if (k > j) {
            if (!(ok && ones == 1)) return 0.0;
            tau.push_back(pos);
        }
    }
    auto P = [&](int i, int k) { return (i128)(mu[gidx(i, k)] + k - i); };
    int tj = tau[0];
    Rat q(1);
    for (int k = 1; k <= j; ++k) q = q * Rat(P(tj, j + 1) - P(k, j));
    for (int k = 1; k <= j + 1; ++k) {
        if (k == tj) continue;
        q = q * Rat(1, P(tj, j + 1) - P(k, j + 1));
    }
    int sign = 1;
    for (int l = j + 2; l <= D; ++l) {
        int tprev = tau[l - j - 2], tl = tau[l - j - 1];
        sign *= (tprev < tl) ? -1 : 1;
        for (int k = 1; k <= l; ++k) {
            if (k == tl) continue;
            q = q * Rat(P(tprev, l - 1) - P(k, l) + 1, P(tl, l) - P(k, l));
        }
        for (int k = 1; k < l; ++k) {
            if (k == tprev) continue;
            q = q * Rat(P(tl, l) - P(k, l - 1), P(tprev, l - 1) - P(k, l - 1) + 1);
        }
    }
    if (q.p < 0 && std::fabs(rn(q)) < 1e-14) q.p = -q.p;
    if (q.p < 0) throw Error(PIMI_EINVAL, "negative squared CG coefficient");
    return (double)sign * sqrt_exact(q);
}

static Rat branch_weight(const Partition &lam, const Partition &mu) {
    int D = (int)lam.size(), row = -1;
    for (int i = 0; i < D; ++i)
        if (lam[i] == mu[i] + 1) { row = i; break; }
    if (row < 0) throw Error(PIMI_EINVAL, "partition is not a one-box child");
    std::vector<i128> sh(D);
    for (int i = 0; i < D; ++i) sh[i] = lam[i] + D - (i + 1);
    i128 pivot = sh[row];
    Rat r(pivot);
    for (int i = 0; i < row; ++i) { i128 df = sh[i] - pivot; r = r * Rat(df + 1, df); }
    for (int i = row + 1; i < D; ++i) { i128 df = pivot - sh[i]; r = r * Rat(df - 1, df); }
    return r;
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
static Transitions onebody_transitions(const std::vector<Pattern> &parents, const std::vector<Pattern> &children,
                                       int D) {
    std::map<std::vector<int>, std::vector<int32_t>> by_weight;
    for (size_t i = 0; i < parents.size(); ++i) by_weight[content(parents[i], D)].push_back((int32_t)i);
    Transitions tr(children.size());
    for (size_t c = 0; c < children.size(); ++c) {
        std::vector<int> cw = content(children[c], D);
        for (int label = 0; label < D; ++label) {
            std::vector<int> pw = cw;
            pw[label] += 1;
            auto it = by_weight.find(pw);
            if (it == by_weight.end()) continue;
            for (int32_t pi : it->second) {
                double v = cgc(children[c], label, parents[pi], D);
                if (v != 0.0) tr[c].emplace_back(pi, label, v);
            }
        }
    }
    return tr;
}

void Geometry::build(int N_, int d_) {
    N = N_; d = d_;
    sectors = partitions(N, d);
    patterns.clear(); removed.clear(); trans.clear(); sqrt_f.clear();
    std::map<Partition, std::vector<Pattern>> child_patterns;
    for (auto &p : sectors) {
        patterns.push_back(gt_patterns(p));
        std::vector<Partition> ch;
        for (int r : removable_corners(p)) { Partition q = p; q[r] -= 1; ch.push_back(q); }
        removed.push_back(ch);
        sqrt_f.push_back(sqrt_symmetric_group_dimension(p));
    }
    for (size_t s = 0; s < sectors.size(); ++s)
        for (auto &mu : removed[s]) {
            auto it = child_patterns.find(mu);
            if (it == child_patterns.end()) it = child_patterns.emplace(mu, gt_patterns(mu)).first;
            trans[{(int)s, mu}] = onebody_transitions(patterns[s], it->second, d);
        }
}

std::vector<Partition> Geometry::common(int li, int ni) const {
    std::vector<Partition> c;
    for (auto &mu : removed[li])
        if (std::find(removed[ni].begin(), removed[ni].end(), mu) != removed[ni].end()) c.push_back(mu);
    return c;
}

double Geometry::scale(int li, const Partition &mu, int ni) const {
    Rat w = branch_weight(sectors[li], mu) * branch_weight(sectors[ni], mu);
    return sqrt_exact(w);
}

void Geometry::contract(int li, const Partition &mu, int ni, const cd *X, std::vector<cd> &tmp) const {
    const Transitions &L = trans.at({li, mu});
    const Transitions &R = trans.at({ni, mu});
    int64_t ml = dim(li);
    tmp.assign((size_t)(ml * dim(ni)), cd(0.0, 0.0));
    for (size_t c = 0; c < L.size(); ++c)
        for (auto &lt : L[c])
            for (auto &rt : R[c]) {
                double z = std::get<2>(lt) * std::get<2>(rt);
                if (z == 0.0) continue;
                int i = std::get<1>(lt), j = std::get<1>(rt);
                cd &cell = tmp[(size_t)(std::get<0>(lt) + (int64_t)std::get<0>(rt) * ml)];
                cell = cell + z * X[i + j * d];
            }
}

static CSC dense_to_csc(const std::vector<cd> &A, int64_t m, int64_t n) {
    CSC C;
    C.zero(m, n);
    for (int64_t c = 0; c < n; ++c) {
        for (int64_t r = 0; r < m; ++r) {
            cd v = A[(size_t)(r + c * m)];
            if (v.real() == 0.0 && v.imag() == 0.0) continue;
            C.rows.push_back((int32_t)r);
            C.vals.push_back(v);
        }
        C.colptr[c + 1] = (int64_t)C.rows.size();
    }
    return C;
}

static std::vector<cd> csc_to_dense(const CSC &C) {
    std::vector<cd> A((size_t)(C.m * C.n), cd(0, 0));
    for (int64_t c = 0; c < C.n; ++c)
        for (int64_t p = C.colptr[c]; p < C.colptr[c + 1]; ++p) A[(size_t)(C.rows[p] + c * C.m)] += C.vals[p];
    return A;
}

CSC Geometry::collective_block(int s, const cd *X) const {
    int64_t m = dim(s);
    std::vector<cd> K((size_t)(m * m), cd(0, 0)), tmp;
    for (auto &mu : common(s, s)) {
        double sc = scale(s, mu, s);
        contract(s, mu, s, X, tmp);
        for (size_t k = 0; k < K.size(); ++k) K[k] = K[k] + sc * tmp[k];
    }
    return dense_to_csc(K, m, m);
}

// ---------------------------------------------------------------------------
// compile + fusion
// ---------------------------------------------------------------------------
static std::vector<CSC> fuse(const std::vector<std::pair<cd, std::vector<CSC>>> &parts, const FusedPlan &P) {
    std::vector<CSC> out(P.ns);
    for (int s = 0; s < P.ns; ++s) {
        int64_t m = P.dims[s];
        CSC &O = out[s];
        O.zero(m, m);
        std::vector<cd> acc((size_t)m, cd(0, 0));
        std::vector<char> mark((size_t)m, 0);
        std::vector<int32_t> touched;
        for (int64_t col = 0; col < m; ++col) {
            touched.clear();
            for (auto &pr : parts) {
                const CSC &B = pr.second[s];
                for (int64_t p = B.colptr[col]; p < B.colptr[col + 1]; ++p) {
                    int32_t r = B.rows[p];
                    if (!mark[r]) { mark[r] = 1; touched.push_back(r); }
                    acc[r] = acc[r] + pr.first * B.vals[p];
                }
            }
            std::sort(touched.begin(), touched.end());
            for (int32_t r : touched) {
                cd v = acc[r];
                if (!(v.real() == 0.0 && v.imag() == 0.0)) { O.rows.push_back(r); O.vals.push_back(v); }
                acc[r] = cd(0, 0);
                mark[r] = 0;
            }
            O.colptr[col + 1] = (int64_t)O.rows.size();
        }
    }
    return out;
}

// K^H K, exact zeros dropped.  Q[a,c] = sum_k conj(K[k,a]) K[k,c] accumulated
// row by row in increasing k over the stored entries only: every nonzero
// product is added in the same order as in the dense triple loop, and the
// skipped exact-zero products cannot change a nonzero sum, so the blocks are
// the dense loop's to the bit (an exactly zero sum is dropped either way).
static CSC gram(const CSC &K) {
    std::vector<std::vector<std::pair<int64_t, cd>>> rows((size_t)K.m);
    for (int64_t col = 0; col < K.n; ++col)
        for (int64_t p = K.colptr[(size_t)col]; p < K.colptr[(size_t)col + 1]; ++p)
            rows[(size_t)K.rows[(size_t)p]].emplace_back(col, K.vals[(size_t)p]);
    std::vector<cd> Q((size_t)(K.n * K.n), cd(0, 0));
    for (int64_t k = 0; k < K.m; ++k)
        for (const auto &ec : rows[(size_t)k])
            for (const auto &ea : rows[(size_t)k])
                Q[(size_t)(ea.first + ec.first * K.n)] += std::conj(ea.second) * ec.second;
    return dense_to_csc(Q, K.n, K.n);
}

// The geometry depends on (N, d) only: models of the same size (a parameter
// scan, a stream of requests) share one build, copied out of this cache.
static const Geometry &cached_geometry(int N, int d) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::unique_ptr<Geometry>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto &slot = cache[{N, d}];
    if (!slot) {
        std::unique_ptr<Geometry> g(new Geometry());
        g->build(N, d);
        slot = std::move(g);
    }
    return *slot;
}

void lower_model(int N, int d, const std::vector<TermDesc> &terms, FusedPlan &P) {
    if (N < 1 || d < 1) throw Error(PIMI_EINVAL, "N and d must be positive");
    if (d > 6) throw Error(PIMI_EINVAL, "local dimension d > 6 is not supported");
    P.geom = cached_geometry(N, d);
    const Geometry &G = P.geom;
    P.ns = (int)G.sectors.size();
    P.dims.clear(); P.offs.clear();
    int64_t off = 0;
    for (int s = 0; s < P.ns; ++s) {
        P.dims.push_back(G.dim(s));
        P.offs.push_back(off);
        off += G.dim(s) * G.dim(s);
    }
    P.n = off;
    std::vector<std::pair<cd, std::vector<CSC>>> hams, losses;
    for (const TermDesc &t : terms) {
        if ((int)t.op.size() != d * d) throw Error(PIMI_EDIM, "term operator must be d x d");
        for (auto &v : t.op)
            if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
                throw Error(PIMI_EINVAL, "fixed operator prototypes must contain only finite coefficients");
        if (!std::isfinite(t.rate)) throw Error(PIMI_EINVAL, "rate must be finite");
        std::vector<CSC> blocks(P.ns);
        switch (t.kind) {
        case PIMI_TERM_LOCAL_H:
        case PIMI_TERM_COLLECTIVE_H: {
            if (t.hbar == 0 || !std::isfinite(t.hbar)) throw Error(PIMI_EINVAL, "hbar must be finite and nonzero");
            for (int s = 0; s < P.ns; ++s) blocks[s] = G.collective_block(s, t.op.data());
            hams.emplace_back(cd(t.rate / t.hbar, 0.0), std::move(blocks));
            break;
        }
        case PIMI_TERM_DIRECT_PRODUCT_H: {
            if (t.hbar == 0 || !std::isfinite(t.hbar)) throw Error(PIMI_EINVAL, "hbar must be finite and nonzero");
            if ((int)t.op2.size() != d * d) throw Error(PIMI_EDIM, "second factor must be d x d");
            for (int s = 0; s < P.ns; ++s) {
                int64_t m = P.dims[s];
                std::vector<cd> A = csc_to_dense(G.collective_block(s, t.op.data()));
                std::vector<cd> B = csc_to_dense(G.collective_block(s, t.op2.data()));
                std::vector<cd> C((size_t)(m * m), cd(0, 0));
                for (int64_t c = 0; c < m; ++c)
                    for (int64_t k = 0; k < m; ++k) {
                        cd b = B[(size_t)(k + c * m)];
                        if (b == cd(0, 0)) continue;
                        for (int64_t r = 0; r < m; ++r) C[(size_t)(r + c * m)] += A[(size_t)(r + k * m)] * b;
                    }
                blocks[s] = dense_to_csc(C, m, m);
            }
            hams.emplace_back(cd(t.rate / t.hbar, 0.0), std::move(blocks));
            break;
        }
        case PIMI_TERM_COLLECTIVE_JUMP: {
            std::vector<CSC> Q(P.ns);
            for (int s = 0; s < P.ns; ++s) {
                blocks[s] = G.collective_block(s, t.op.data());
                Q[s] = gram(blocks[s]);
            }
            P.channels.push_back(JumpChannel{0, (int32_t)P.K.size(), t.rate, Q});
            losses.emplace_back(cd(t.rate, 0.0), std::move(Q));
            P.cg_scale.push_back(cd(t.rate, 0.0));
            P.K.push_back(std::move(blocks));
            break;
        }
        case PIMI_TERM_LOCAL_JUMP: {
            std::vector<cd> LdL((size_t)(d * d), cd(0, 0));
            for (int j = 0; j < d; ++j)
                for (int i = 0; i < d; ++i) {
                    cd s(0, 0);
                    for (int k = 0; k < d; ++k) s += std::conj(t.op[k + i * d]) * t.op[k + j * d];
                    LdL[i + j * d] = s;
                }
            std::vector<CSC> Q(P.ns);
            for (int s = 0; s < P.ns; ++s) Q[s] = G.collective_block(s, LdL.data());
            P.channels.push_back(JumpChannel{1, (int32_t)P.ob.size(), t.rate, Q});
            losses.emplace_back(cd(t.rate, 0.0), std::move(Q));
            std::vector<Branch> brs;
            std::vector<cd> tmp;
            for (int li = 0; li < P.ns; ++li)
                for (int ni = 0; ni < P.ns; ++ni)
                    for (auto &mu : G.common(li, ni)) {
                        Branch b;
                        b.out = li; b.in = ni; b.r = G.scale(li, mu, ni);
                        G.contract(li, mu, ni, t.op.data(), tmp);
                        int64_t ml = P.dims[li], mn = P.dims[ni];
                        b.dense = tmp;
                        for (int64_t c = 0; c < mn; ++c)
                            for (int64_t r = 0; r < ml; ++r) {
                                cd v = tmp[(size_t)(r + c * ml)];
                                if (v.real() == 0.0 && v.imag() == 0.0) continue;
                                b.srows.push_back((int32_t)r); b.scols.push_back((int32_t)c); b.svals.push_back(v);
                            }
                        unsigned __int128 nnz = b.svals.size();
                        unsigned __int128 sparse_cost = 2 * nnz * nnz;
                        unsigned __int128 dense_cost = (unsigned __int128)ml * mn * (ml + mn);
                        b.use_support = sparse_cost <= dense_cost;
                        brs.push_back(std::move(b));
                    }
            P.ob_scale.push_back(cd(t.rate, 0.0));
            P.ob.push_back(std::move(brs));
            break;
        }
        default:
            throw Error(PIMI_EINVAL, "unsupported term kind");
        }
    }
    P.has_h = !hams.empty();
    P.has_q = !losses.empty();
    if (P.has_h) P.H = fuse(hams, P);
    if (P.has_q) P.Q = fuse(losses, P);
}

void collective_operator(const FusedPlan &P, const cd *X, cd *out) {
    for (int s = 0; s < P.ns; ++s) {
        CSC K = P.geom.collective_block(s, X);
        int64_t m = P.dims[s];
        double sc = P.geom.sqrt_f[s];
        cd *o = out + P.offs[s];
        for (int64_t k = 0; k < m * m; ++k) o[k] = cd(0, 0);
        for (int64_t c = 0; c < m; ++c)
            for (int64_t p = K.colptr[c]; p < K.colptr[c + 1]; ++p) o[K.rows[p] + c * m] = sc * K.vals[p];
    }
}

// iid_state (src/operators.jl:146-250): the physical Schur blocks of
// rho^{otimes N} by the one-box Clebsch-Gordan recurrence
//   D_lambda = sum_{i,o} A[i,o] U_i D_mu U_o^T,   D_(0,...,0) = [1],
// mu = lambda without its first removable corner, U_l[a, c] the one-box
// isometry of (mu, lambda) at local label l (the transition tables of the
// geometry of lambda's particle number), one particle number at a time; the
// coefficient block of sector lambda is sqrt(f^lambda) D_lambda.  A is the
// (validated, Hermitian-averaged) d x d one-particle matrix, column-major.
void iid_state(const FusedPlan &P, const cd *A, cd *out) {
    const int N = P.geom.N, d = P.geom.d;
    std::map<Partition, std::vector<cd>> prev;
    prev[Partition((size_t)d, 0)] = {cd(1.0, 0.0)};
    for (int nn = 1; nn <= N; ++nn) {
        Geometry Gl;
        const Geometry *G = &P.geom;
        if (nn < N) {
            Gl.build(nn, d);
            G = &Gl;
        }
        std::map<Partition, std::vector<cd>> cur;
        for (size_t s = 0; s < G->sectors.size(); ++s) {
            const Partition &mu = G->removed[s].front();
            const std::vector<cd> &D = prev.at(mu);
            const Transitions &T = G->trans.at({(int)s, mu});
            const int64_t m = G->dim((int)s), mm = (int64_t)T.size();
            // W_o[p, b] = sum_c D[p, c] U_o[b, c]
            std::vector<cd> W((size_t)d * (size_t)(mm * m), cd(0, 0));
            for (int64_t c = 0; c < mm; ++c)
                for (const auto &t : T[(size_t)c]) {
                    const int64_t b = std::get<0>(t);
                    const int o = std::get<1>(t);
                    const double v = std::get<2>(t);
                    if (v == 0.0) continue;
                    cd *w = W.data() + (size_t)o * (size_t)(mm * m) + (size_t)(b * mm);
                    for (int64_t p = 0; p < mm; ++p) w[p] += D[(size_t)(p + c * mm)] * v;
                }
            // D_lambda[a, b] = sum_{i,o} A[i,o] sum_p U_i[a, p] W_o[p, b]
            std::vector<cd> Dl((size_t)(m * m), cd(0, 0));
            for (int64_t p = 0; p < mm; ++p)
                for (const auto &t : T[(size_t)p]) {
                    const int64_t a = std::get<0>(t);
                    const int i = std::get<1>(t);
                    const double v = std::get<2>(t);
                    if (v == 0.0) continue;
                    for (int o = 0; o < d; ++o) {
                        const cd coef = A[i + o * d] * v;
                        if (coef == cd(0, 0)) continue;
                        const cd *w = W.data() + (size_t)o * (size_t)(mm * m);
                        for (int64_t b = 0; b < m; ++b) Dl[(size_t)(a + b * m)] += coef * w[(size_t)(p + b * mm)];
                    }
                }
            cur[G->sectors[s]] = std::move(Dl);
        }
        prev = std::move(cur);
    }
    for (int s = 0; s < P.ns; ++s) {
        const std::vector<cd> &D = prev.at(P.geom.sectors[(size_t)s]);
        const double sc = P.geom.sqrt_f[(size_t)s];
        const int64_t m = P.dims[(size_t)s];
        for (int64_t k = 0; k < m * m; ++k) out[P.offs[(size_t)s] + k] = sc * D[(size_t)k];
    }
}

// ---------------------------------------------------------------------------
// device stencil tables
// ---------------------------------------------------------------------------
namespace {
struct RowLists {
    std::vector<std::vector<std::pair<int32_t, cd>>> rows;
};

void append_ell(Stencil &st, const RowLists &R, int64_t m_out, int32_t &width, int64_t &base) {
    width = 1;
    for (auto &r : R.rows) width = std::max<int32_t>(width, (int32_t)r.size());
    base = (int64_t)st.col.size();
    st.col.resize((size_t)(base + (int64_t)width * m_out), 0);
    st.val.resize((size_t)(base + (int64_t)width * m_out), cd(0, 0));
    for (int64_t a = 0; a < m_out; ++a)
        for (size_t k = 0; k < R.rows[a].size(); ++k) {
            st.col[(size_t)(base + (int64_t)k * m_out + a)] = R.rows[a][k].first;
            st.val[(size_t)(base + (int64_t)k * m_out + a)] = R.rows[a][k].second;
        }
}

RowLists rows_of_dense(const std::vector<cd> &A, int64_t m, int64_t n) {
    RowLists R;
    R.rows.resize((size_t)m);
    // column order (contiguous in memory); each row still receives its entries
    // in increasing column order
    for (int64_t c = 0; c < n; ++c)
        for (int64_t a = 0; a < m; ++a) {
            cd v = A[(size_t)(a + c * m)];
            if (v.real() != 0.0 || v.imag() != 0.0) R.rows[a].emplace_back((int32_t)c, v);
        }
    return R;
}
}  // namespace

void build_stencil(const FusedPlan &P, Stencil &st, bool sector_diagonal_only) {
    st = Stencil();
    st.term_base.push_back(0);
    for (int l = 0; l < P.ns; ++l) {
        int64_t m = P.dims[l];
        if (P.has_h || P.has_q) {
            std::vector<cd> Hd = P.has_h ? csc_to_dense(P.H[l]) : std::vector<cd>((size_t)(m * m), cd(0, 0));
            std::vector<cd> Qd = P.has_q ? csc_to_dense(P.Q[l]) : std::vector<cd>((size_t)(m * m), cd(0, 0));
            std::vector<cd> Gl((size_t)(m * m)), Br((size_t)(m * m));
            const cd mi(0.0, -1.0), pi_(0.0, 1.0);
            for (int64_t c = 0; c < m; ++c)
                for (int64_t a = 0; a < m; ++a) {
                    size_t k = (size_t)(a + c * m);
                    Gl[k] = mi * Hd[k] - 0.5 * Qd[k];       // left factor  -iH - Q/2
                    cd R = pi_ * Hd[k] - 0.5 * Qd[k];       // right factor  iH - Q/2, as B = R^H
                    Br[(size_t)(c + a * m)] = std::conj(R);
                }
            StencilTerm tl{l, 2, 0, 1, 0, 0, 1.0, 0.0};
            append_ell(st, rows_of_dense(Gl, m, m), m, tl.wa, tl.a_base);
            st.terms.push_back(tl);
            StencilTerm tr{l, 1, 1, 0, 0, 0, 1.0, 0.0};
            append_ell(st, rows_of_dense(Br, m, m), m, tr.wb, tr.b_base);
            st.terms.push_back(tr);
            st.entries_per_apply += m * m * (tl.wa + tr.wb);
        }
        for (size_t g = 0; g < P.K.size(); ++g) {
            std::vector<cd> Kd = csc_to_dense(P.K[g][l]);
            RowLists R = rows_of_dense(Kd, m, m);
            StencilTerm t{l, 0, 0, 0, 0, 0, P.cg_scale[g].real(), P.cg_scale[g].imag()};
            append_ell(st, R, m, t.wa, t.a_base);
            t.wb = t.wa; t.b_base = t.a_base;
            st.terms.push_back(t);
            st.entries_per_apply += m * m * (int64_t)t.wa * t.wb;
        }
        for (size_t g = 0; g < P.ob.size(); ++g)
            for (const Branch &b : P.ob[g]) {
                if (b.out != l) continue;
                if (sector_diagonal_only && b.in != l) continue;
                int64_t mn = P.dims[b.in];
                RowLists R = rows_of_dense(b.dense, m, mn);
                cd s = P.ob_scale[g] * b.r;
                StencilTerm t{b.in, 0, 0, 0, 0, 0, s.real(), s.imag()};
                append_ell(st, R, m, t.wa, t.a_base);
                t.wb = t.wa; t.b_base = t.a_base;
                st.terms.push_back(t);
                st.entries_per_apply += m * m * (int64_t)t.wa * t.wb;
            }
        st.term_base.push_back((int64_t)st.terms.size());
    }
    // every table index must address the input block it reads
    for (int l = 0; l < P.ns; ++l)
        for (int64_t k = st.term_base[l]; k < st.term_base[l + 1]; ++k) {
            const StencilTerm &t = st.terms[(size_t)k];
            if (t.in_sector < 0 || t.in_sector >= P.ns) throw Error(PIMI_EINVAL, "stencil sector out of range");
            int64_t m = P.dims[l], mn = P.dims[t.in_sector];
            if ((t.flags & 1) && mn != m) throw Error(PIMI_EINVAL, "identity factor needs a square block");
            auto chk = [&](int64_t base, int32_t w) {
                if (base < 0 || base + (int64_t)w * m > (int64_t)st.col.size()) throw Error(PIMI_EINVAL, "stencil base");
                for (int64_t q = base; q < base + (int64_t)w * m; ++q)
                    if (st.col[(size_t)q] < 0 || st.col[(size_t)q] >= mn) throw Error(PIMI_EINVAL, "stencil column");
            };
            if (!(t.flags & 1)) chk(t.a_base, t.wa);
            if (!(t.flags & 2)) chk(t.b_base, t.wb);
        }
}

// The adjoint Liouvillian under the PI inner product (src/liouvillian.jl:3275-3293):
// L X = -i[H, X] - {Q, X}/2 + sum_g c_g K X K^H + sum_branches s C X_nu C^H has
// L^dagger Y = i[H, Y] - {Q, Y}/2 + sum_g conj(c_g) K^H Y K
//              + sum_branches conj(s) C^H Y_lambda C  (into sector nu),
// i.e. the fused plan with H -> -H, K -> K^H, each branch lambda <- nu turned
// into nu <- lambda with C -> C^H, and conjugated scales.
static CSC conj_transpose(const CSC &A) {
    std::vector<cd> d = csc_to_dense(A), t((size_t)(A.m * A.n));
    for (int64_t c = 0; c < A.n; ++c)
        for (int64_t r = 0; r < A.m; ++r) t[(size_t)(c + r * A.n)] = std::conj(d[(size_t)(r + c * A.m)]);
    return dense_to_csc(t, A.n, A.m);
}

void adjoint_plan(const FusedPlan &P, FusedPlan &A) {
    A.ns = P.ns;
    A.n = P.n;
    A.dims = P.dims;
    A.offs = P.offs;
    A.has_h = P.has_h;
    A.has_q = P.has_q;
    A.H = P.H;
    for (CSC &h : A.H)
        for (cd &v : h.vals) v = -v;
    A.Q = P.Q;
    A.cg_scale.clear();
    for (cd c : P.cg_scale) A.cg_scale.push_back(std::conj(c));
    A.K.assign(P.K.size(), {});
    for (size_t g = 0; g < P.K.size(); ++g)
        for (const CSC &k : P.K[g]) A.K[g].push_back(conj_transpose(k));
    A.ob_scale.clear();
    for (cd c : P.ob_scale) A.ob_scale.push_back(std::conj(c));
    A.ob.assign(P.ob.size(), {});
    for (size_t g = 0; g < P.ob.size(); ++g)
        for (const Branch &b : P.ob[g]) {
            Branch c;
            c.out = b.in;
            c.in = b.out;
            c.r = b.r;
            const int64_t mo = P.dims[(size_t)b.out], mi = P.dims[(size_t)b.in];
            c.dense.assign((size_t)(mi * mo), cd(0, 0));
            for (int64_t j = 0; j < mi; ++j)
                for (int64_t i = 0; i < mo; ++i)
                    c.dense[(size_t)(j + i * mi)] = std::conj(b.dense[(size_t)(i + j * mo)]);
            A.ob[g].push_back(std::move(c));
        }
}

// ---------------------------------------------------------------------------
// stationary-solve preconditioner factors
// ---------------------------------------------------------------------------
void hermitian_eigh(int64_t m, std::vector<cd> A, std::vector<double> &d, std::vector<cd> &U) {
    // 1. Householder reduction A = Q T Q^H, T Hermitian tridiagonal.  With
    //    v the unit reflector vector, p = A v and beta = v^H A v (real),
    //    H A H = A - v w^H - w v^H for w = 2 p - 2 beta v.
    auto at = [m](std::vector<cd> &M, int64_t i, int64_t j) -> cd & { return M[(size_t)(i + j * m)]; };
    std::vector<cd> Q((size_t)(m * m), cd(0, 0));
    for (int64_t i = 0; i < m; ++i) at(Q, i, i) = 1.0;
    std::vector<cd> v((size_t)m), w((size_t)m);
    for (int64_t k = 0; k + 2 < m; ++k) {
        double xn = 0;
        for (int64_t i = k + 1; i < m; ++i) xn += std::norm(at(A, i, k));
        xn = std::sqrt(xn);
        if (xn == 0) continue;
        const cd x0 = at(A, k + 1, k);
        const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1, 0);
        const cd alpha = -ph * xn;
        for (int64_t i = k + 1; i < m; ++i) v[(size_t)i] = at(A, i, k);
        v[(size_t)(k + 1)] -= alpha;
        double vn = 0;
        for (int64_t i = k + 1; i < m; ++i) vn += std::norm(v[(size_t)i]);
        vn = std::sqrt(vn);
        if (vn == 0) continue;
        for (int64_t i = k + 1; i < m; ++i) v[(size_t)i] /= vn;
        // p = A v over the trailing block (rows k..m-1; column k handled exactly below)
        for (int64_t i = k + 1; i < m; ++i) w[(size_t)i] = 0;
        for (int64_t j = k + 1; j < m; ++j) {
            const cd vj = v[(size_t)j];
            if (vj == cd(0, 0)) continue;
            for (int64_t i = k + 1; i < m; ++i) w[(size_t)i] += at(A, i, j) * vj;
        }
        double beta = 0;
        for (int64_t i = k + 1; i < m; ++i) beta += (std::conj(v[(size_t)i]) * w[(size_t)i]).real();
        for (int64_t i = k + 1; i < m; ++i) w[(size_t)i] = 2.0 * w[(size_t)i] - 2.0 * beta * v[(size_t)i];
        for (int64_t j = k + 1; j < m; ++j) {
            const cd cvj = std::conj(v[(size_t)j]), cwj = std::conj(w[(size_t)j]);
            for (int64_t i = k + 1; i < m; ++i) at(A, i, j) -= v[(size_t)i] * cwj + w[(size_t)i] * cvj;
        }
        // column / row k: H x = alpha e1
        at(A, k + 1, k) = alpha;
        at(A, k, k + 1) = std::conj(alpha);
        for (int64_t i = k + 2; i < m; ++i) { at(A, i, k) = 0; at(A, k, i) = 0; }
        // Q <- Q H on columns k+1..m-1
        for (int64_t i = 0; i < m; ++i) {
            cd qv(0, 0);
            for (int64_t j = k + 1; j < m; ++j) qv += at(Q, i, j) * v[(size_t)j];
            qv *= 2.0;
            for (int64_t j = k + 1; j < m; ++j) at(Q, i, j) -= qv * std::conj(v[(size_t)j]);
        }
    }
    // 2. T = D T_r D^H with T_r real (subdiagonal |e_k|), D = diag of phases.
    std::vector<double> diag((size_t)m), sub((size_t)m, 0.0);
    std::vector<cd> ph((size_t)m, cd(1, 0));
    for (int64_t k = 0; k < m; ++k) diag[(size_t)k] = at(A, k, k).real();
    for (int64_t k = 0; k + 1 < m; ++k) {
        const cd e = at(A, k + 1, k);
        const double ae = std::abs(e);
        sub[(size_t)k] = ae;
        ph[(size_t)(k + 1)] = ae > 0 ? ph[(size_t)k] * (e / ae) : ph[(size_t)k];
    }
    // 3. T_r = Z diag(d) Z^T by the implicit QL iteration with Wilkinson shifts
    //    (Bowdler, Martin, Reinsch and Wilkinson 1968).
    std::vector<double> Z((size_t)(m * m), 0.0);
    for (int64_t i = 0; i < m; ++i) Z[(size_t)(i + i * m)] = 1.0;
    std::vector<double> &dd = diag, &e = sub;
    for (int64_t l = 0; l < m; ++l) {
        int iter = 0;
        while (true) {
            int64_t mm = l;
            for (; mm + 1 < m; ++mm) {
                const double t = std::fabs(dd[(size_t)mm]) + std::fabs(dd[(size_t)(mm + 1)]);
                if (std::fabs(e[(size_t)mm]) <= std::numeric_limits<double>::epsilon() * t) break;
            }
            if (mm == l) break;
            if (++iter > 60) throw Error(PIMI_EINVAL, "tridiagonal QL iteration did not converge");
            double g = (dd[(size_t)(l + 1)] - dd[(size_t)l]) / (2.0 * e[(size_t)l]);
            double r = std::hypot(g, 1.0);
            g = dd[(size_t)mm] - dd[(size_t)l] + e[(size_t)l] / (g + (g >= 0 ? r : -r));
            double sn = 1.0, c = 1.0, p = 0.0;
            int64_t i = mm - 1;
            bool early = false;
            for (; i >= l; --i) {
                double f = sn * e[(size_t)i], b = c * e[(size_t)i];
                r = std::hypot(f, g);
                e[(size_t)(i + 1)] = r;
                if (r == 0.0) {
                    dd[(size_t)(i + 1)] -= p;
                    e[(size_t)mm] = 0.0;
                    early = true;
                    break;
                }
                sn = f / r;
                c = g / r;
                g = dd[(size_t)(i + 1)] - p;
                r = (dd[(size_t)i] - g) * sn + 2.0 * c * b;
                p = sn * r;
                dd[(size_t)(i + 1)] = g + p;
                g = c * r - b;
                double *zi = &Z[(size_t)(i * m)], *zi1 = &Z[(size_t)((i + 1) * m)];
                for (int64_t k = 0; k < m; ++k) {
                    const double t = zi1[k];
                    zi1[k] = sn * zi[k] + c * t;
                    zi[k] = c * zi[k] - sn * t;
                }
            }
            if (early) continue;
            dd[(size_t)l] -= p;
            e[(size_t)l] = g;
            e[(size_t)mm] = 0.0;
        }
    }
    // 4. U = Q D Z
    U.assign((size_t)(m * m), cd(0, 0));
    for (int64_t k = 0; k < m; ++k) {
        const cd pk = ph[(size_t)k];
        for (int64_t i = 0; i < m; ++i) at(Q, i, k) *= pk;
    }
    for (int64_t j = 0; j < m; ++j)
        for (int64_t k = 0; k < m; ++k) {
            const double z = Z[(size_t)(k + j * m)];
            if (z == 0.0) continue;
            for (int64_t i = 0; i < m; ++i) at(U, i, j) += at(Q, i, k) * z;
        }
    d = dd;
}

void build_sector_preconditioner(const FusedPlan &P, std::vector<cd> &U, std::vector<cd> &inv_den) {
    U.assign((size_t)P.n, cd(0, 0));
    inv_den.assign((size_t)P.n, cd(0, 0));
    auto one_sector = [&](int s) {
        int64_t m = P.dims[s];
        std::vector<cd> Hd = P.has_h ? csc_to_dense(P.H[s]) : std::vector<cd>((size_t)(m * m), cd(0, 0));
        // Hermitian part of the fused block (it is Hermitian up to rounding)
        for (int64_t j = 0; j < m; ++j)
            for (int64_t i = 0; i <= j; ++i) {
                cd h = 0.5 * (Hd[(size_t)(i + j * m)] + std::conj(Hd[(size_t)(j + i * m)]));
                Hd[(size_t)(i + j * m)] = h;
                Hd[(size_t)(j + i * m)] = std::conj(h);
            }
        std::vector<double> d;
        std::vector<cd> Us;
        hermitian_eigh(m, Hd, d, Us);
        // q_a = (U^H Q U)_{aa}, Q in its sparse column form
        std::vector<double> q((size_t)m, 0.0);
        if (P.has_q) {
            const CSC &Qs = P.Q[s];
            for (int64_t a = 0; a < m; ++a) {
                cd acc(0, 0);
                for (int64_t b = 0; b < m; ++b) {
                    const cd uba = Us[(size_t)(b + a * m)];
                    for (int64_t k = Qs.colptr[(size_t)b]; k < Qs.colptr[(size_t)(b + 1)]; ++k)
                        acc += std::conj(Us[(size_t)(Qs.rows[(size_t)k] + a * m)]) * Qs.vals[(size_t)k] * uba;
                }
                q[(size_t)a] = acc.real();
            }
        }
        double dscale = 0;
        for (int64_t a = 0; a < m; ++a) dscale = std::max(dscale, std::fabs(d[(size_t)a]) + std::fabs(q[(size_t)a]));
        for (int64_t b = 0; b < m; ++b)
            for (int64_t a = 0; a < m; ++a) {
                cd den(-0.5 * (q[(size_t)a] + q[(size_t)b]), -(d[(size_t)a] - d[(size_t)b]));
                // an exactly undamped, degenerate pair has no inverse: project it out
                cd inv = std::abs(den) > 1e-13 * std::max(dscale, 1e-300) ? 1.0 / den : cd(0, 0);
                inv_den[(size_t)(P.offs[s] + a + b * m)] = inv;
                U[(size_t)(P.offs[s] + a + b * m)] = Us[(size_t)(a + b * m)];
            }
    };
    // sectors are independent: largest first over a small thread pool
    std::vector<int> order((size_t)P.ns);
    for (int s = 0; s < P.ns; ++s) order[(size_t)s] = s;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return P.dims[a] > P.dims[b]; });
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int nthreads = (int)std::min<unsigned>(std::min<unsigned>(hw, 16u), (unsigned)std::max(P.ns, 1));
    std::atomic<int> next{0};
    std::vector<std::string> errors((size_t)nthreads);
    auto worker = [&](int tid) {
        try {
            for (int k; (k = next.fetch_add(1)) < P.ns;) one_sector(order[(size_t)k]);
        } catch (const std::exception &ex) {
            errors[(size_t)tid] = ex.what();
            next.store(P.ns);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto &th : pool) th.join();
    for (auto &msg : errors)
        if (!msg.empty()) throw Error(PIMI_EINVAL, msg);
}

}  // namespace pimi
