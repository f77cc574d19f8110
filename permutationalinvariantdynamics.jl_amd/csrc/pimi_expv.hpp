// Host side of the adaptive Krylov exponential (src/krylov_extensions.jl:1346-1568):
// the small dense exponential, the factorization record, the projection
// coefficients and error estimate, and the step control.  Shared by the
// single-vector path (pimi_expv_series) and the batch engine's dynamics jobs.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "pimi_internal.hpp"

namespace pimi {

// ---------------------------------------------------------------------------
// small dense exponential (host): scaling and squaring with diagonal Pade
// approximants of degree 3..13, Higham (2005) thresholds.
// ---------------------------------------------------------------------------
using cdv = std::vector<cd>;
// C = A B on separate real and imaginary planes (pimi_dense.cpp: host code,
// built for AVX2 and for the baseline ISA, picked at load time)
void mm_planes(int q, const double *ar, const double *ai, const cd *B, double *cr, double *ci);
// C = A B for q x q column-major complex matrices (0.9 ms at q = 41 with
// std::complex arithmetic in the innermost loop, which stays scalar)
inline void mm(int q, const cdv &A, const cdv &B, cdv &C) {
    const size_t qq = (size_t)q * q;
    std::vector<double> pl(4 * qq);
    double *ar = pl.data(), *ai = ar + qq, *cr = ai + qq, *ci = cr + qq;
    for (size_t e = 0; e < qq; ++e) { ar[e] = A[e].real(); ai[e] = A[e].imag(); }
    mm_planes(q, ar, ai, B.data(), cr, ci);
    C.resize(qq);
    for (size_t e = 0; e < qq; ++e) C[e] = cd(cr[e], ci[e]);
}
// B <- A^-1 B: Gaussian elimination with partial pivoting (largest |a|^2),
// column-oriented so that every inner loop runs with unit stride.
inline void solve_inplace(int q, cdv A, cdv &B) {
    double *a = reinterpret_cast<double *>(A.data());
    double *b = reinterpret_cast<double *>(B.data());
    const size_t ld = 2 * (size_t)q;  // doubles per column
    for (int k = 0; k < q; ++k) {
        int p = k;
        double best = -1;
        for (int i = k; i < q; ++i) {
            const double m2 = a[ld * k + 2 * i] * a[ld * k + 2 * i] + a[ld * k + 2 * i + 1] * a[ld * k + 2 * i + 1];
            if (m2 > best) { best = m2; p = i; }
        }
        if (p != k)
            for (int j = 0; j < q; ++j) {
                std::swap(A[k + (size_t)j * q], A[p + (size_t)j * q]);
                std::swap(B[k + (size_t)j * q], B[p + (size_t)j * q]);
            }
        const cd piv = A[k + (size_t)k * q];
        if (piv == cd(0, 0)) throw Error(PIMI_EINVAL, "singular Pade denominator");
        // multipliers l_i = a_ik / piv, kept in column k below the diagonal
        // (one complex division per column; the rest multiply)
        const cd rp = 1.0 / piv;
        const double rr = rp.real(), ri = rp.imag();
        for (int i = k + 1; i < q; ++i) {
            const double xr = a[ld * k + 2 * i], xi = a[ld * k + 2 * i + 1];
            a[ld * k + 2 * i] = xr * rr - xi * ri;
            a[ld * k + 2 * i + 1] = xr * ri + xi * rr;
        }
        const double *l = a + ld * k;
        auto eliminate = [&](double *col) {  // col[i] -= l_i col[k], i > k
            const double xr = col[2 * k], xi = col[2 * k + 1];
            if (xr == 0.0 && xi == 0.0) return;
            for (int i = k + 1; i < q; ++i) {
                const double lr = l[2 * i], li = l[2 * i + 1];
                col[2 * i] -= lr * xr - li * xi;
                col[2 * i + 1] -= lr * xi + li * xr;
            }
        };
        for (int j = k + 1; j < q; ++j) eliminate(a + ld * j);
        for (int j = 0; j < q; ++j) eliminate(b + ld * j);
    }
    // back substitution, column sweep: x_i = b_i / u_ii, then b_r -= u_ri x_i (r < i)
    std::vector<cd> rdiag((size_t)q);
    for (int i = 0; i < q; ++i) rdiag[(size_t)i] = 1.0 / A[i + (size_t)i * q];
    for (int j = 0; j < q; ++j) {
        double *col = b + ld * j;
        for (int i = q - 1; i >= 0; --i) {
            const double br = col[2 * i], bi = col[2 * i + 1];
            const double dr = rdiag[(size_t)i].real(), di = rdiag[(size_t)i].imag();
            const double xr = br * dr - bi * di, xi = br * di + bi * dr;
            col[2 * i] = xr;
            col[2 * i + 1] = xi;
            const double *u = a + ld * i;
            for (int r = 0; r < i; ++r) {
                col[2 * r] -= u[2 * r] * xr - u[2 * r + 1] * xi;
                col[2 * r + 1] -= u[2 * r] * xi + u[2 * r + 1] * xr;
            }
        }
    }
}
inline cdv expm(int q, const cdv &Ain) {
    static const double c13[] = {64764752532480000., 32382376266240000., 7771770303897600., 1187353796428800.,
                                 129060195264000., 10559470521600., 670442572800., 33522128640., 1323241920.,
                                 40840800., 960960., 16380., 182., 1.};
    static const double c9[] = {17643225600., 8821612800., 2075673600., 302702400., 30270240.,
                                2162160., 110880., 3960., 90., 1.};
    static const double c7[] = {17297280., 8648640., 1995840., 277200., 25200., 1512., 56., 1.};
    static const double c5[] = {30240., 15120., 3360., 420., 30., 1.};
    static const double c3[] = {120., 60., 12., 1.};
    static const double th[] = {1.495585217958292e-2, 2.539398330063230e-1, 9.504178996162932e-1,
                                2.097847961257068e0, 5.371920351148152e0};
    cdv A = Ain;
    double nrm = 0;
    for (int j = 0; j < q; ++j) {
        double s = 0;
        for (int i = 0; i < q; ++i) s += std::abs(A[i + (size_t)j * q]);
        nrm = std::max(nrm, s);
    }
    int deg = 13, sq = 0;
    const double *c = c13;
    if (nrm <= th[0]) { deg = 3; c = c3; }
    else if (nrm <= th[1]) { deg = 5; c = c5; }
    else if (nrm <= th[2]) { deg = 7; c = c7; }
    else if (nrm <= th[3]) { deg = 9; c = c9; }
    else if (nrm > th[4]) {
        sq = std::max(0, (int)std::ceil(std::log2(nrm / th[4])));
        double f = std::ldexp(1.0, -sq);
        for (auto &z : A) z *= f;
    }
    const size_t qq = (size_t)q * q;
    cdv I(qq, cd(0, 0));
    for (int i = 0; i < q; ++i) I[i + (size_t)i * q] = 1.0;
    cdv A2, U(qq), Vv(qq), T(qq);
    mm(q, A, A, A2);
    if (deg == 13) {
        cdv A4, A6, W1, W2;
        mm(q, A2, A2, A4);
        mm(q, A4, A2, A6);
        for (size_t k = 0; k < qq; ++k) T[k] = c[13] * A6[k] + c[11] * A4[k] + c[9] * A2[k];
        mm(q, A6, T, W1);
        for (size_t k = 0; k < qq; ++k) W1[k] += c[7] * A6[k] + c[5] * A4[k] + c[3] * A2[k] + c[1] * I[k];
        mm(q, A, W1, U);
        for (size_t k = 0; k < qq; ++k) T[k] = c[12] * A6[k] + c[10] * A4[k] + c[8] * A2[k];
        mm(q, A6, T, W2);
        for (size_t k = 0; k < qq; ++k) Vv[k] = W2[k] + c[6] * A6[k] + c[4] * A4[k] + c[2] * A2[k] + c[0] * I[k];
    } else {
        cdv Pw = I, odd(qq, cd(0, 0)), nxt;
        std::fill(Vv.begin(), Vv.end(), cd(0, 0));
        for (int k = 0; 2 * k <= deg; ++k) {
            for (size_t e = 0; e < qq; ++e) { odd[e] += c[2 * k + 1] * Pw[e]; Vv[e] += c[2 * k] * Pw[e]; }
            if (2 * (k + 1) <= deg) { mm(q, Pw, A2, nxt); Pw.swap(nxt); }
        }
        mm(q, A, odd, U);
    }
    cdv Num(qq), Den(qq);
    for (size_t k = 0; k < qq; ++k) { Num[k] = Vv[k] + U[k]; Den[k] = Vv[k] - U[k]; }
    solve_inplace(q, Den, Num);
    for (int k = 0; k < sq; ++k) { mm(q, Num, Num, T); Num.swap(T); }
    return Num;
}


// One Arnoldi factorization as krylov_expv! uses it: S is the (m+1) x (m+1)
// (column stride m + 1) augmented Hessenberg, k its Arnoldi steps, q the
// projected dimension (k + 1 unless an exact breakdown), coeff_count the
// basis columns the projection uses.
struct Factorization {
    double beta = 0, next_remainder = 0;
    bool zero_state = false, completed = false, enlarged = false;
    int k = 0, used = 0, q = 0, coeff_count = 0;
    std::vector<cd> S;  // (m+1) x (m+1)
};


// The Arnoldi Hessenberg of a factorization from delayed-reorthogonalization
// steps (the batch engine's kb_dcgs_* kernels, also used by the
// single-vector path): step j reports s = Q^H p_j and t = Q^H w (j + 1
// entries each), rho_j and e_j; column j - 1 is final after step j.  The
// breakdown tests follow krylov_expv!'s Arnoldi (src/krylov_extensions.jl:
// 1346-1401): an exact breakdown at a zero subdiagonal or when the whole
// space is spanned, a near breakdown when a subdiagonal is at most
// sqrt(eps) |A v_j| (here the norm of the final column, which equals |A v_j|
// for an orthonormal basis); after a near breakdown the factorization takes
// one more step (the enlargement).
struct DelayedArnoldi {
    int m = 0, ld = 0;
    int64_t n = 0;
    std::vector<cd> Hraw, cprov;  // final columns ((m + 2) x (m + 1), stride m + 2); first-pass coefficients
    bool near = false, enlarging = false, exact = false;
    void reset(int m_, int64_t n_) {
        m = m_;
        n = n_;
        ld = m + 2;
        Hraw.assign((size_t)(m + 2) * (m + 1), cd(0, 0));
        cprov.assign((size_t)m + 3, cd(0, 0));
        near = enlarging = exact = false;
    }
    cd &H(int i, int c) { return Hraw[(size_t)i + (size_t)c * ld]; }
    // step jj's reductions; 1: the factorization is complete, 0: continue
    // with step jj + 1, -1: the start vector was zero or not finite
    int step(int jj, const cx<double> *sv, const cx<double> *tv, double rho, cd e) {
        auto S_ = [&](int i) { return cd(sv[i].x, sv[i].y); };
        auto T_ = [&](int i) { return cd(tv[i].x, tv[i].y); };
        bool complete = false;
        if (jj >= 1) {
            for (int i = 0; i < jj; ++i) H(i, jj - 1) = cprov[(size_t)i] + S_(i);
            H(jj, jj - 1) = rho;
            if (enlarging) {
                complete = true;
            } else {
                double img2 = 0;
                for (int i = 0; i <= jj; ++i) img2 += std::norm(H(i, jj - 1));
                const double breakfactor = std::sqrt(2.220446049250313e-16);
                if (rho == 0 || (int64_t)jj == n) {
                    exact = complete = true;
                } else {
                    near |= rho <= breakfactor * std::max(std::sqrt(img2), 2.2250738585072014e-308);
                    if (jj == m) {
                        if (near) enlarging = true;
                        else complete = true;
                    }
                }
            }
        } else if (!(rho > 0) || !std::isfinite(rho)) {
            return -1;
        }
        if (complete) return 1;
        for (int i = 0; i < jj; ++i) {
            cd acc = T_(i);
            for (int c = 0; c < jj; ++c) acc -= H(i, c) * S_(c);
            cprov[(size_t)i] = acc / rho;
        }
        cprov[(size_t)jj] = e - (jj >= 1 ? S_(jj - 1) : cd(0, 0));
        return 0;
    }
    // the factorization after the step jj that completed it, laid out as
    // the single-vector path's (S of stride m + 1)
    void finish(int jj, double beta, Factorization &F) {
        const bool enlarged = enlarging;
        const int k = enlarged ? jj - 1 : jj, fl = m + 1;
        F = Factorization();
        F.beta = beta;
        F.S.assign((size_t)fl * fl, cd(0, 0));
        for (int c = 0; c < k; ++c)
            for (int i = 0; i < k; ++i) F.S[(size_t)i + (size_t)c * fl] = H(i, c);
        const int q = k + (exact ? 0 : 1);
        if (!exact) F.S[(size_t)k + (size_t)(k - 1) * fl] = H(k, k - 1);
        F.used = k;
        if (enlarged) {
            F.enlarged = true;
            F.used += 1;
            for (int i = 0; i < q; ++i) F.S[(size_t)i + (size_t)(q - 1) * fl] += H(i, k);
            F.next_remainder = H(k + 1, k).real();
        }
        F.k = k;
        F.q = q;
        F.coeff_count = enlarged ? q : k;
        F.completed = exact || (enlarged && ((int64_t)q == n || F.next_remainder == 0));
    }
};

// The projection of one attempt with step h (signed): coeffs[j] = beta
// exp(h S)[j, 0] for j < coeff_count, and the local error estimate; `ld` is
// the column stride of F.S (m + 1).
inline double expv_coefficients(const Factorization &F, int ld, double h, std::vector<cx<double>> &coeffs) {
    const int q = F.q;
    cdv S((size_t)q * q);
    for (int j = 0; j < q; ++j)
        for (int i = 0; i < q; ++i) S[(size_t)i + (size_t)j * q] = h * F.S[(size_t)i + (size_t)j * ld];
    cdv E = expm(q, S);
    coeffs.resize((size_t)F.coeff_count);
    for (int j = 0; j < F.coeff_count; ++j) coeffs[(size_t)j] = cx<double>{E[(size_t)j].real(), E[(size_t)j].imag()};
    if (F.completed) return 0.0;
    if (F.enlarged) return F.beta * std::fabs(h) * F.next_remainder * std::abs(E[(size_t)(q - 1)]);
    return F.beta * std::abs(E[(size_t)F.k]);
}

// Step control of one interval of length |t|: attempts take h (clamped to
// what remains), an accepted step advances, the next h follows from the
// error estimate (factor clamped to [0.2, 5], safety factor, at most 0.5
// after a rejection).
struct ExpvControl {
    double total = 0, reached = 0, h = 0, hmax = 0, hmin = 0, direction = 1;
    // false when the interval is empty
    bool begin(double t) {
        total = std::fabs(t);
        reached = 0;
        hmax = total;
        hmin = 0;
        h = hmax;
        direction = t > 0 ? 1.0 : -1.0;
        return total != 0;
    }
    bool done() const { return reached >= total; }
    // the signed step of the next attempt
    double step() {
        h = std::min(h, total - reached);
        return direction * h;
    }
    // after an attempt with error estimate err and trial norm: 1 accepted,
    // 0 rejected (retry with the new h), -1 rejected with no step left.
    // cur_norm() gives |current state| when the trial is not finite.
    template <typename CurNorm>
    int decide(double err, double trial_norm, CurNorm cur_norm, const Factorization &F, const pimi_expv_opts &o) {
        const bool breakdown = F.completed;
        const bool finite = std::isfinite(err) && std::isfinite(trial_norm);
        if (!finite) trial_norm = cur_norm();
        const double localtol = (h / total) * (o.atol + o.rtol * trial_norm);
        if (!std::isfinite(localtol)) throw Error(PIMI_EINVAL, "exponential-action local tolerance overflowed");
        const bool accept = finite && (breakdown || err <= localtol);
        if (accept) {
            const double remaining = total - reached;
            reached = (h == remaining) ? total : reached + h;
        }
        double factor;
        if (!finite) factor = 0.2;
        else if (breakdown) factor = 5.0;
        else if (err == 0) factor = 2.0;
        else {
            factor = o.safety * std::pow(localtol / err, 1.0 / (double)std::max(F.k, 1));
            factor = std::min(5.0, std::max(0.2, factor));
        }
        if (!accept) factor = std::min(factor, 0.5);
        const double hnew = std::min(hmax, h * factor);
        if (!accept && (hnew < hmin || hnew == 0 || (reached > 0 && reached + hnew == reached))) return -1;
        h = std::max(hnew, hmin);
        return accept ? 1 : 0;
    }
};

}  // namespace pimi
