/*
 * Oracle: CPU restatement of the reference's PI Liouvillian apply and the
 * Krylov solvers that drive it.
 *
 * TEST INFRASTRUCTURE ONLY.  The product library never links this file; only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, and
 * only as the checker.  Parity of each function is pinned by
 * tests/test_oracle_pins.py (brute-force Hilbert-space Lindbladians, closed
 * forms, dense norms and inverses, invariants) and the fixtures under
 * tests/golden/.
 *
 * Restated reference functions (paths relative to the reference root):
 *   fused apply            src/liouvillian.jl:3070-3105, 3018-3028, 2461-2561
 *   threaded apply         src/threaded_apply.jl:58-84, 506-549, 564-609
 *   GMRES(m) with MGS2     src/krylov.jl:710-786
 *   krylov_steady_state    src/krylov.jl:156-169, 797-891
 *   (plus the product's sector preconditioner and inner-outer flexible GMRES,
 *   which have no counterpart in the reference)
 *   krylov_expv!           src/krylov_extensions.jl:1346-1568
 *   solve_dynamics (expv)  src/highlevel.jl:489-525, 614-733
 *
 * The plan is the reference's FusedStaticPIKernel as flat arrays: per-sector
 * CSC blocks for the fused Hamiltonian and loss, per-gain CSC blocks for
 * collective gains, and one-body gain branches with either an exact support
 * list or a dense rectangular contraction (src/liouvillian.jl:270-292).
 *
 * Operator-scale probe: the reference draws its three probes from Julia's
 * MersenneTwister(0x51ca1e); that stream is not reproducible here, so the
 * probes come from splitmix64(0x51ca1e) + Box-Muller.  The converged steady
 * state does not depend on the probe beyond the solver tolerance.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* A set of CSC blocks stored back to back. */
typedef struct {
    const int64_t *cp_base;  /* per block: start index into colptr */
    const int64_t *colptr;   /* concatenated colptr arrays (m+1 each), values index into rows/vals */
    const int32_t *rows;
    const cplx *vals;
} csc_set;

typedef struct {
    int32_t ns;
    const int64_t *dims;   /* m per sector */
    const int64_t *offs;   /* zero-based coordinate offset per sector */
    int64_t n;
    int32_t has_h, has_q;
    csc_set H, Q;          /* ns blocks each */
    int32_t ncg;
    const cplx *cg_scale;  /* ncg */
    csc_set K;             /* ncg*ns blocks, gain-major */
    int32_t nob;
    const cplx *ob_scale;  /* nob */
    const int64_t *ob_br_base; /* nob+1: branch range per gain */
    const int32_t *br_out, *br_in;
    const double *br_r;
    const int32_t *br_use_support;
    const int64_t *br_sup_base; /* nbr+1 into sup arrays */
    const int32_t *sup_rows, *sup_cols;
    const cplx *sup_vals;
    const int64_t *br_dense_base; /* nbr: offset into dense array (column-major m_out x m_in) */
    const cplx *dense;
    int64_t max_dim;
} oplan;

/* ---------------- sparse x dense helpers (SparseArrays loop orders) ---------------- */

/* C = A*X, A sparse m x m (CSC), X dense m x m column-major */
static void sp_mul_left(cplx *C, const csc_set *S, int64_t blk, const cplx *X, int64_t m) {
    memset(C, 0, sizeof(cplx) * m * m);
    const int64_t *cp = S->colptr + S->cp_base[blk];
    for (int64_t col = 0; col < m; ++col)
        for (int64_t k = 0; k < m; ++k) {
            cplx b = X[k + col * m];
            for (int64_t p = cp[k]; p < cp[k + 1]; ++p) C[S->rows[p] + col * m] += S->vals[p] * b;
        }
}

/* C = X*A, A sparse (CSC) */
static void sp_mul_right(cplx *C, const csc_set *S, int64_t blk, const cplx *X, int64_t m) {
    memset(C, 0, sizeof(cplx) * m * m);
    const int64_t *cp = S->colptr + S->cp_base[blk];
    for (int64_t j = 0; j < m; ++j)
        for (int64_t p = cp[j]; p < cp[j + 1]; ++p) {
            int64_t k = S->rows[p];
            cplx v = S->vals[p];
            for (int64_t i = 0; i < m; ++i) C[i + j * m] += X[i + k * m] * v;
        }
}

/* C = X*A', A sparse (CSC) */
static void sp_mul_right_adj(cplx *C, const csc_set *S, int64_t blk, const cplx *X, int64_t m) {
    memset(C, 0, sizeof(cplx) * m * m);
    const int64_t *cp = S->colptr + S->cp_base[blk];
    for (int64_t k = 0; k < m; ++k)
        for (int64_t p = cp[k]; p < cp[k + 1]; ++p) {
            int64_t j = S->rows[p];
            cplx v = conj(S->vals[p]);
            for (int64_t i = 0; i < m; ++i) C[i + j * m] += X[i + k * m] * v;
        }
}

/* ---------------- per-output-sector pieces ---------------- */

typedef struct { cplx *left, *right, *scratch, *target; } sector_scratch;

static void alloc_scratch(sector_scratch *w, int64_t md) {
    size_t sz = sizeof(cplx) * (size_t)md * (size_t)md;
    w->left = malloc(sz); w->right = malloc(sz); w->scratch = malloc(sz); w->target = malloc(sz);
}
static void free_scratch(sector_scratch *w) { free(w->left); free(w->right); free(w->scratch); free(w->target); }

static void apply_h_sector(const oplan *P, int s, const cplx *x, cplx *y, sector_scratch *w) {
    int64_t m = P->dims[s], off = P->offs[s];
    const cplx *X = x + off;
    sp_mul_left(w->left, &P->H, s, X, m);
    sp_mul_right(w->right, &P->H, s, X, m);
    for (int64_t i = 0; i < m * m; ++i) y[off + i] += -I * (w->left[i] - w->right[i]);
}

static void apply_cg_sector(const oplan *P, int g, int s, const cplx *x, cplx *y, sector_scratch *w) {
    int64_t m = P->dims[s], off = P->offs[s];
    const cplx *X = x + off;
    int64_t blk = (int64_t)g * P->ns + s;
    sp_mul_left(w->left, &P->K, blk, X, m);
    sp_mul_right_adj(w->right, &P->K, blk, w->left, m);
    for (int64_t i = 0; i < m * m; ++i) y[off + i] += P->cg_scale[g] * w->right[i];
}

static void apply_q_sector(const oplan *P, int s, const cplx *x, cplx *y, sector_scratch *w) {
    int64_t m = P->dims[s], off = P->offs[s];
    const cplx *X = x + off;
    sp_mul_left(w->left, &P->Q, s, X, m);
    sp_mul_right(w->right, &P->Q, s, X, m);
    for (int64_t i = 0; i < m * m; ++i) y[off + i] -= 0.5 * (w->left[i] + w->right[i]);
}

/* One branch of a one-body gain: y_out += scale * C X_in C'
 * (src/liouvillian.jl:2461-2537). */
static void apply_branch(const oplan *P, int64_t br, cplx scale, const cplx *x, cplx *y, sector_scratch *w) {
    int li = P->br_out[br], ni = P->br_in[br];
    int64_t ml = P->dims[li], mn = P->dims[ni];
    const cplx *X = x + P->offs[ni];
    int64_t oo = P->offs[li];
    if (P->br_use_support[br]) {
        int64_t b0 = P->br_sup_base[br], b1 = P->br_sup_base[br + 1];
        for (int64_t r = b0; r < b1; ++r) {
            int64_t ocol = P->sup_rows[r], icol = P->sup_cols[r];
            cplx rv = P->sup_vals[r];
            for (int64_t l = b0; l < b1; ++l) {
                int64_t orow = P->sup_rows[l], irow = P->sup_cols[l];
                y[oo + orow + ocol * ml] += scale * P->sup_vals[l] * conj(rv) * X[irow + icol * mn];
            }
        }
        return;
    }
    const cplx *C = P->dense + P->br_dense_base[br];
    cplx *S = w->scratch, *T = w->target;
    /* S = C X (ml x mn), T = S C' (ml x ml) */
    for (int64_t col = 0; col < mn; ++col)
        for (int64_t row = 0; row < ml; ++row) {
            cplx v = 0;
            for (int64_t k = 0; k < mn; ++k) v += C[row + k * ml] * X[k + col * mn];
            S[row + col * ml] = v;
        }
    for (int64_t col = 0; col < ml; ++col)
        for (int64_t row = 0; row < ml; ++row) {
            cplx v = 0;
            for (int64_t k = 0; k < mn; ++k) v += S[row + k * ml] * conj(C[col + k * ml]);
            T[row + col * ml] = v;
        }
    for (int64_t i = 0; i < ml * ml; ++i) y[oo + i] += scale * T[i];
}

/* ---------------- serial apply (reference order) ---------------- */

void oracle_apply(const oplan *P, const cplx *x, cplx *y) {
    sector_scratch w;
    alloc_scratch(&w, P->max_dim);
    memset(y, 0, sizeof(cplx) * P->n);
    if (P->has_h)
        for (int s = 0; s < P->ns; ++s) apply_h_sector(P, s, x, y, &w);
    for (int g = 0; g < P->ncg; ++g)
        for (int s = 0; s < P->ns; ++s) apply_cg_sector(P, g, s, x, y, &w);
    if (P->has_q)
        for (int s = 0; s < P->ns; ++s) apply_q_sector(P, s, x, y, &w);
    for (int g = 0; g < P->nob; ++g)
        for (int64_t br = P->ob_br_base[g]; br < P->ob_br_base[g + 1]; ++br)
            apply_branch(P, br, P->ob_scale[g] * P->br_r[br], x, y, &w);
    free_scratch(&w);
}

/* ---------------- threaded apply: output-sector ownership ----------------
 * LPT assignment on cost m^3+m with ties broken by sector index
 * (src/threaded_apply.jl:58-84); each worker computes all contributions to its
 * own sectors in the serial order, so the result is bitwise equal to
 * oracle_apply. */

static void lpt_assign(const oplan *P, int workers, int *owner) {
    int ns = P->ns;
    int *order = malloc(sizeof(int) * ns);
    double *load = calloc(workers, sizeof(double));
    for (int s = 0; s < ns; ++s) order[s] = s;
    for (int a = 0; a < ns; ++a)
        for (int b = a + 1; b < ns; ++b) {
            double ca = (double)P->dims[order[a]], cb = (double)P->dims[order[b]];
            ca = ca * ca * ca + ca; cb = cb * cb * cb + cb;
            if (cb > ca || (cb == ca && order[b] < order[a])) { int t = order[a]; order[a] = order[b]; order[b] = t; }
        }
    for (int k = 0; k < ns; ++k) {
        int s = order[k], best = 0;
        for (int wi = 1; wi < workers; ++wi) if (load[wi] < load[best]) best = wi;
        owner[s] = best;
        double c = (double)P->dims[s];
        load[best] += c * c * c + c;
    }
    free(order); free(load);
}

void oracle_apply_threaded(const oplan *P, const cplx *x, cplx *y, int nthreads) {
    int workers = nthreads < P->ns ? nthreads : P->ns;
    if (workers <= 1) { oracle_apply(P, x, y); return; }
    int *owner = malloc(sizeof(int) * P->ns);
    lpt_assign(P, workers, owner);
    memset(y, 0, sizeof(cplx) * P->n);
#pragma omp parallel num_threads(workers)
    {
        int wid = 0;
#ifdef _OPENMP
        wid = omp_get_thread_num();
#endif
        sector_scratch w;
        alloc_scratch(&w, P->max_dim);
        for (int s = 0; s < P->ns; ++s) {
            if (owner[s] != wid) continue;
            if (P->has_h) apply_h_sector(P, s, x, y, &w);
            for (int g = 0; g < P->ncg; ++g) apply_cg_sector(P, g, s, x, y, &w);
            if (P->has_q) apply_q_sector(P, s, x, y, &w);
            for (int g = 0; g < P->nob; ++g)
                for (int64_t br = P->ob_br_base[g]; br < P->ob_br_base[g + 1]; ++br)
                    if (P->br_out[br] == s) apply_branch(P, br, P->ob_scale[g] * P->br_r[br], x, y, &w);
        }
        free_scratch(&w);
    }
    free(owner);
}

static void do_apply(const oplan *P, const cplx *x, cplx *y, int nthreads) {
    if (nthreads > 1) oracle_apply_threaded(P, x, y, nthreads);
    else oracle_apply(P, x, y);
}

/* y = D x, D the sector-diagonal part of L: every term whose input and output
 * sector coincide (Hamiltonian, loss, collective gains, and the one-body
 * branches with br_in == br_out), in the order of oracle_apply. */
void oracle_apply_diag(const oplan *P, const cplx *x, cplx *y, int nthreads) {
    int workers = nthreads < P->ns ? nthreads : P->ns;
    if (workers < 1) workers = 1;
    int *owner = malloc(sizeof(int) * P->ns);
    if (workers > 1) lpt_assign(P, workers, owner);
    else for (int s = 0; s < P->ns; ++s) owner[s] = 0;
    memset(y, 0, sizeof(cplx) * P->n);
#pragma omp parallel num_threads(workers)
    {
        int wid = 0;
#ifdef _OPENMP
        wid = omp_get_thread_num();
#endif
        sector_scratch w;
        alloc_scratch(&w, P->max_dim);
        for (int s = 0; s < P->ns; ++s) {
            if (owner[s] != wid) continue;
            if (P->has_h) apply_h_sector(P, s, x, y, &w);
            for (int g = 0; g < P->ncg; ++g) apply_cg_sector(P, g, s, x, y, &w);
            if (P->has_q) apply_q_sector(P, s, x, y, &w);
            for (int g = 0; g < P->nob; ++g)
                for (int64_t br = P->ob_br_base[g]; br < P->ob_br_base[g + 1]; ++br)
                    if (P->br_out[br] == s && P->br_in[br] == s)
                        apply_branch(P, br, P->ob_scale[g] * P->br_r[br], x, y, &w);
        }
        free_scratch(&w);
    }
    free(owner);
}

/* ---------------- BLAS-1 ----------------
 * The solvers' vector operations run on g_blas_threads OpenMP threads (set by
 * each solver entry point from its nthreads), as the reference's threaded
 * CPU path runs BLAS-1 over multi-threaded BLAS.  Dots are summed over a
 * fixed chunking and the partial sums added in chunk order, so a result
 * depends only on the thread count. */
static int g_blas_threads = 1;

static cplx zdot(const cplx *a, const cplx *b, int64_t n) {  /* sum conj(a) b */
    int nt = g_blas_threads;
    if (nt <= 1 || n < 65536) {
        cplx s = 0;
        for (int64_t i = 0; i < n; ++i) s += conj(a[i]) * b[i];
        return s;
    }
    double re[256], im[256];
    if (nt > 256) nt = 256;
    int64_t chunk = (n + nt - 1) / nt;
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int c = 0; c < nt; ++c) {
        cplx s = 0;
        int64_t i1 = (int64_t)(c + 1) * chunk < n ? (int64_t)(c + 1) * chunk : n;
        for (int64_t i = (int64_t)c * chunk; i < i1; ++i) s += conj(a[i]) * b[i];
        re[c] = creal(s);
        im[c] = cimag(s);
    }
    cplx s = 0;
    for (int c = 0; c < nt; ++c) s += re[c] + I * im[c];
    return s;
}

/* w -= h v */
static void zaxpy_neg(cplx *w, cplx h, const cplx *v, int64_t n) {
#pragma omp parallel for schedule(static) num_threads(g_blas_threads) if (n >= 65536)
    for (int64_t k = 0; k < n; ++k) w[k] = w[k] - h * v[k];
}

/* out = V[:, 0:k] y (columns of length n) */
static void zgemv_cols(cplx *out, const cplx *V, int64_t n, int k, const cplx *y) {
#pragma omp parallel for schedule(static) num_threads(g_blas_threads) if (n >= 65536)
    for (int64_t t = 0; t < n; ++t) {
        cplx zz = 0;
        for (int jj = 0; jj < k; ++jj) zz += V[(int64_t)jj * n + t] * y[jj];
        out[t] = zz;
    }
}
static double dznrm2(const cplx *a, int64_t n) {
    /* scaled 2-norm (LAPACK-style) so that tiny or huge vectors do not
       underflow/overflow */
    double scale = 0.0, ssq = 1.0;
    for (int64_t i = 0; i < n; ++i) {
        double v[2] = {creal(a[i]), cimag(a[i])};
        for (int k = 0; k < 2; ++k) {
            if (v[k] != 0.0) {
                double t = fabs(v[k]);
                if (scale < t) { ssq = 1.0 + ssq * (scale / t) * (scale / t); scale = t; }
                else ssq += (t / scale) * (t / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

/* ---------------- deterministic probe generator ---------------- */

static uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double unif(uint64_t *s) { return ((splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

void oracle_probe_vector(uint64_t *state, cplx *x, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        double u = unif(state), v = unif(state);
        double r = sqrt(-2.0 * log(u));
        double a = 6.283185307179586 * v;
        x[i] = (r * cos(a) + I * r * sin(a)) * 0.7071067811865476;
    }
}

double oracle_operator_scale(const oplan *P, int nthreads) {
    uint64_t st = 0x51ca1eULL;
    cplx *x = malloc(sizeof(cplx) * P->n), *y = malloc(sizeof(cplx) * P->n);
    double scale = 0.0;
    for (int k = 0; k < 3; ++k) {
        oracle_probe_vector(&st, x, P->n);
        double nx = dznrm2(x, P->n);
        do_apply(P, x, y, nthreads);
        double r = dznrm2(y, P->n) / nx;
        if (r > scale) scale = r;
    }
    free(x); free(y);
    return scale;
}

/* ---------------- sector preconditioner ----------------
 * Right preconditioner of the stationary solve (an addition of this build;
 * the reference's GMRES accepts a preconditioner but has no such one):
 * per sector, R -> scale * U ((U^H R U) .* inv_den) U^H with U the
 * eigenvectors of the fused Hamiltonian block and inv_den[a,b] =
 * 1/(-i(d_a-d_b) - (q_a+q_b)/2), q = diag(U^H Q U).  U and inv_den come in
 * the PI block layout. */
typedef struct { const cplx *U, *inv_den; double scale; } sector_pc;

static void pc_apply(const oplan *P, const sector_pc *pc, const cplx *r, cplx *out, int nthreads) {
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
    for (int s = 0; s < P->ns; ++s) {
        int64_t m = P->dims[s], o = P->offs[s];
        const cplx *U = pc->U + o, *R = r + o, *Dn = pc->inv_den + o;
        cplx *T1 = malloc(sizeof(cplx) * m * m), *T2 = malloc(sizeof(cplx) * m * m);
        /* T1 = U^H R */
        for (int64_t j = 0; j < m; ++j)
            for (int64_t i = 0; i < m; ++i) {
                cplx acc = 0;
                for (int64_t k = 0; k < m; ++k) acc += conj(U[k + i * m]) * R[k + j * m];
                T1[i + j * m] = acc;
            }
        /* T2 = (T1 U) .* inv_den * scale */
        for (int64_t j = 0; j < m; ++j)
            for (int64_t i = 0; i < m; ++i) {
                cplx acc = 0;
                for (int64_t k = 0; k < m; ++k) acc += T1[i + k * m] * U[k + j * m];
                T2[i + j * m] = acc * Dn[i + j * m] * pc->scale;
            }
        /* T1 = U T2 */
        for (int64_t j = 0; j < m; ++j)
            for (int64_t i = 0; i < m; ++i) {
                cplx acc = 0;
                for (int64_t k = 0; k < m; ++k) acc += U[i + k * m] * T2[k + j * m];
                T1[i + j * m] = acc;
            }
        /* out = T1 U^H */
        for (int64_t j = 0; j < m; ++j)
            for (int64_t i = 0; i < m; ++i) {
                cplx acc = 0;
                for (int64_t k = 0; k < m; ++k) acc += T1[i + k * m] * conj(U[j + k * m]);
                out[o + i + j * m] = acc;
            }
        free(T1); free(T2);
    }
}

/* ---------------- GMRES(m) with MGS2 (src/krylov.jl:732-786) ---------------- */

typedef struct {
    int64_t iterations, restarts, applications;
    int32_t converged;
    double residual, normalized_residual, trace_error, operator_scale, linear_residual;
    double weighted_relative_residual;
} steady_info;

typedef struct {
    const oplan *P;
    int nthreads;
    double invscale;
    const int64_t *tidx;
    const cplx *tval;   /* trace functional (converted) */
    const cplx *vval;   /* normalized trace functional t/(t.t) */
    int64_t nt;
    int64_t applications;
    const double *sigma; /* sector weighting: per-coordinate sqrt(f) (NULL: none) */
    cplx *scratch;
} steady_op;

/* y = A z, A = L/s + v t^H, or with sector weighting S A S^{-1}: the input is
 * unscaled, L applied, the output scaled back (t and v are passed already
 * transformed: t/sigma and sigma v on the trace support). */
static void steady_apply(steady_op *op, const cplx *z, cplx *y) {
    const cplx *in = z;
    if (op->sigma) {
        for (int64_t i = 0; i < op->P->n; ++i) op->scratch[i] = z[i] / op->sigma[i];
        in = op->scratch;
    }
    do_apply(op->P, in, y, op->nthreads);
    op->applications++;
    cplx alpha = 0;
    for (int64_t k = 0; k < op->nt; ++k) alpha += conj(op->tval[k]) * z[op->tidx[k]];
    if (op->sigma)
        for (int64_t i = 0; i < op->P->n; ++i) y[i] = op->invscale * (op->sigma[i] * y[i]);
    else
        for (int64_t i = 0; i < op->P->n; ++i) y[i] = op->invscale * y[i];
    for (int64_t k = 0; k < op->nt; ++k) y[op->tidx[k]] += alpha * op->vval[k];
}

static void complex_givens(cplx a, cplx b, double *c, cplx *s) {
    if (b == 0) { *c = 1.0; *s = 0; return; }
    if (a == 0) { *c = 0.0; *s = 1.0; return; }
    double scale = cabs(a) + cabs(b);
    cplx as = a / scale, bs = b / scale;
    double normab = scale * sqrt(creal(as) * creal(as) + cimag(as) * cimag(as) +
                                 creal(bs) * creal(bs) + cimag(bs) * cimag(bs));
    cplx alpha = a / cabs(a);
    *c = cabs(a) / normab;
    *s = alpha * conj(b) / normab;
}

/* Solve (apply) x = b by restarted GMRES; returns converged flag. */
/* norm of the current iterate x + M^{-1} V y (or x + V y), y from the rotated
 * least-squares problem of the first k Arnoldi steps */
static double iterate_norm(steady_op *op, const cplx *x, const cplx *V, const cplx *H, const cplx *g, int m, int k,
                           const sector_pc *pc, cplx *tmp, cplx *z) {
    int64_t n = op->P->n;
    cplx *yv = malloc(sizeof(cplx) * k);
    for (int i = k - 1; i >= 0; --i) {
        cplx acc = g[i];
        for (int jj = i + 1; jj < k; ++jj) acc -= H[i + jj * (m + 1)] * yv[jj];
        yv[i] = acc / H[i + i * (m + 1)];
    }
    zgemv_cols(tmp, V, n, k, yv);
    if (pc) pc_apply(op->P, pc, tmp, z, op->nthreads);
    else memcpy(z, tmp, sizeof(cplx) * n);
    for (int64_t t = 0; t < n; ++t) z[t] += x[t];
    double r = dznrm2(z, n);
    free(yv);
    return r;
}

static int gmres(steady_op *op, cplx *x, const cplx *b, int m, int64_t maxiter, double atol, double rtol,
                 int64_t *iters_out, int64_t *restarts_out, double *res_out, const sector_pc *pc, int backward) {
    int64_t n = op->P->n;
    cplx *V = malloc(sizeof(cplx) * n * (m + 1));
    cplx *H = malloc(sizeof(cplx) * (m + 1) * m);
    cplx *g = malloc(sizeof(cplx) * (m + 1));
    cplx *sn = malloc(sizeof(cplx) * m), *yv = malloc(sizeof(cplx) * m);
    double *cs = malloc(sizeof(double) * m);
    cplx *w = malloc(sizeof(cplx) * n), *r = malloc(sizeof(cplx) * n);
    cplx *z = malloc(sizeof(cplx) * n), *tmp = malloc(sizeof(cplx) * n);
#define Hij(i, j) H[(i) + (j) * (m + 1)]
    double bnorm = dznrm2(b, n);
    double xnorm = dznrm2(x, n);
    /* backward: t^H L = 0 and t.v = 1 give |L x|/s <= 2 |r|, so |r| <= (atol + rtol |x|) / 2
       guarantees |L x| / s <= atol + rtol |x| */
    double tol = backward ? 0.5 * (atol + rtol * xnorm) : atol + rtol * bnorm;
    int64_t iterations = 0, restarts = 0;
    steady_apply(op, x, w);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - w[i];
    double beta = dznrm2(r, n);
    int conv = 0;
    if (beta <= tol) { conv = 1; goto done; }
    while (iterations < maxiter) {
        memset(H, 0, sizeof(cplx) * (m + 1) * m);
        memset(g, 0, sizeof(cplx) * (m + 1));
        for (int64_t i = 0; i < n; ++i) V[i] = r[i] / beta;
        g[0] = beta;
        int kdone = 0;
        int jmax = (int)((maxiter - iterations) < m ? (maxiter - iterations) : m);
        for (int j = 0; j < jmax; ++j) {
            if (pc) {
                pc_apply(op->P, pc, V + (int64_t)j * n, z, op->nthreads);
                steady_apply(op, z, w);
            } else
                steady_apply(op, V + (int64_t)j * n, w);
            for (int pass = 0; pass < 2; ++pass)
                for (int i = 0; i <= j; ++i) {
                    cplx hij = zdot(V + (int64_t)i * n, w, n);
                    Hij(i, j) += hij;
                    zaxpy_neg(w, hij, V + (int64_t)i * n, n);
                }
            Hij(j + 1, j) = dznrm2(w, n);
            if (Hij(j + 1, j) != 0)
                for (int64_t k = 0; k < n; ++k) V[(int64_t)(j + 1) * n + k] = w[k] / Hij(j + 1, j);
            for (int i = 0; i < j; ++i) {
                cplx h1 = Hij(i, j), h2 = Hij(i + 1, j);
                double c = cs[i]; cplx s = sn[i];
                Hij(i, j) = c * h1 + s * h2;
                Hij(i + 1, j) = -conj(s) * h1 + c * h2;
            }
            double c; cplx s;
            complex_givens(Hij(j, j), Hij(j + 1, j), &c, &s);
            cs[j] = c; sn[j] = s;
            cplx h1 = Hij(j, j), h2 = Hij(j + 1, j);
            Hij(j, j) = c * h1 + s * h2;
            Hij(j + 1, j) = 0;
            cplx g1 = g[j], g2 = g[j + 1];
            g[j] = c * g1 + s * g2;
            g[j + 1] = -conj(s) * g1 + c * g2;
            iterations++;
            kdone = j + 1;
            if (backward && (kdone % 25 == 0 || cabs(g[j + 1]) <= tol)) {
                xnorm = iterate_norm(op, x, V, H, g, m, kdone, pc, tmp, z);
                tol = 0.5 * (atol + rtol * xnorm);
            }
            if (cabs(g[j + 1]) <= tol) break;
        }
        for (int i = kdone - 1; i >= 0; --i) {
            cplx sacc = g[i];
            for (int jj = i + 1; jj < kdone; ++jj) sacc -= Hij(i, jj) * yv[jj];
            yv[i] = sacc / Hij(i, i);
        }
        zgemv_cols(w, V, n, kdone, yv);
        if (pc) {
            pc_apply(op->P, pc, w, z, op->nthreads);
            for (int64_t k = 0; k < n; ++k) x[k] += z[k];
        } else
            for (int64_t k = 0; k < n; ++k) x[k] += w[k];
        restarts++;
        steady_apply(op, x, w);
        for (int64_t i = 0; i < n; ++i) r[i] = b[i] - w[i];
        beta = dznrm2(r, n);
        if (backward) { xnorm = dznrm2(x, n); tol = 0.5 * (atol + rtol * xnorm); }
        if (beta <= tol) { conv = 1; break; }
    }
done:
#undef Hij
    *iters_out = iterations; *restarts_out = restarts; *res_out = beta;
    free(V); free(H); free(g); free(sn); free(yv); free(cs); free(w); free(r); free(z); free(tmp);
    return conv;
}

/* ---------------- inner-outer flexible GMRES ----------------
 * The product's PIMI_PRECOND_SECTOR_INNER solver (no counterpart in the
 * reference, whose GMRES takes a fixed preconditioner): flexible GMRES on the
 * bordered A = L/s + v t^H (modified Gram-Schmidt twice here; classical on the
 * device) whose direction z_j = inner(v_j) is `kin` steps of
 * sector-preconditioned GMRES on D/s from zero, D the sector-diagonal part of
 * L (oracle_apply_diag), with one classical Gram-Schmidt pass per inner step
 * and the inner least-squares problem solved by Givens rotations:
 * inner(r) = M^{-1} V_i y, y = argmin |beta e1 - H_i y|, beta = |r|.  The
 * stopping rules are those of gmres() above (the iterate norm refreshed every
 * 5 outer steps under the backward rule, as the device does). */
static void inner_solve(const steady_op *op, const sector_pc *pc, int kin, const cplx *r, cplx *z, cplx *Vi,
                        cplx *t1, cplx *t2) {
    const oplan *P = op->P;
    int64_t n = P->n;
    double beta = dznrm2(r, n);
    if (beta == 0) { memset(z, 0, sizeof(cplx) * n); return; }
    cplx *Hm = calloc((size_t)(kin + 1) * kin, sizeof(cplx));
    cplx *h = malloc(sizeof(cplx) * (kin + 1));
    for (int64_t t = 0; t < n; ++t) Vi[t] = r[t] / beta;
    int kk = 0;
    for (int i = 0; i < kin; ++i) {
        pc_apply(P, pc, Vi + (int64_t)i * n, t1, op->nthreads);
        oracle_apply_diag(P, t1, t2, op->nthreads);
        for (int64_t t = 0; t < n; ++t) t2[t] *= op->invscale;
        for (int k = 0; k <= i; ++k) h[k] = zdot(Vi + (int64_t)k * n, t2, n);   /* one classical pass */
        for (int k = 0; k <= i; ++k) {
            Hm[k + i * (kin + 1)] = h[k];
            zaxpy_neg(t2, h[k], Vi + (int64_t)k * n, n);
        }
        double hn = dznrm2(t2, n);
        Hm[i + 1 + i * (kin + 1)] = hn;
        kk = i + 1;
        if (hn == 0) break;
        for (int64_t t = 0; t < n; ++t) Vi[(int64_t)(i + 1) * n + t] = t2[t] / hn;
    }
    /* least squares min |beta e1 - H y| by Givens rotations */
    cplx *g = calloc(kk + 1, sizeof(cplx)), *y = malloc(sizeof(cplx) * (kk > 0 ? kk : 1));
    g[0] = beta;
    {
        double *cs = malloc(sizeof(double) * (kk > 0 ? kk : 1));
        cplx *sn = malloc(sizeof(cplx) * (kk > 0 ? kk : 1));
        for (int j = 0; j < kk; ++j) {
#define HM(i, jj) Hm[(i) + (jj) * (kin + 1)]
            for (int i = 0; i < j; ++i) {
                cplx h1 = HM(i, j), h2 = HM(i + 1, j);
                HM(i, j) = cs[i] * h1 + sn[i] * h2;
                HM(i + 1, j) = -conj(sn[i]) * h1 + cs[i] * h2;
            }
            double c; cplx sg;
            complex_givens(HM(j, j), HM(j + 1, j), &c, &sg);
            cs[j] = c; sn[j] = sg;
            cplx h1 = HM(j, j), h2 = HM(j + 1, j);
            HM(j, j) = c * h1 + sg * h2;
            HM(j + 1, j) = 0;
            cplx g1 = g[j], g2 = g[j + 1];
            g[j] = c * g1 + sg * g2;
            g[j + 1] = -conj(sg) * g1 + c * g2;
        }
        for (int i = kk - 1; i >= 0; --i) {
            cplx acc = g[i];
            for (int jj = i + 1; jj < kk; ++jj) acc -= HM(i, jj) * y[jj];
            y[i] = acc / HM(i, i);
        }
#undef HM
        free(cs); free(sn);
    }
    zgemv_cols(t2, Vi, n, kk, y);
    pc_apply(P, pc, t2, z, op->nthreads);
    free(Hm); free(h); free(g); free(y);
}

static int fgmres(steady_op *op, cplx *x, const cplx *b, int m, int64_t maxiter, double atol, double rtol,
                  int64_t *iters_out, int64_t *restarts_out, double *res_out, const sector_pc *pc, int backward,
                  int kin) {
    int64_t n = op->P->n;
    cplx *V = malloc(sizeof(cplx) * n * (m + 1)), *Z = malloc(sizeof(cplx) * n * m);
    cplx *Vi = malloc(sizeof(cplx) * n * (kin + 1));
    cplx *H = malloc(sizeof(cplx) * (m + 1) * m);
    cplx *g = malloc(sizeof(cplx) * (m + 1));
    cplx *sn = malloc(sizeof(cplx) * m), *yv = malloc(sizeof(cplx) * m);
    double *cs = malloc(sizeof(double) * m);
    cplx *w = malloc(sizeof(cplx) * n), *r = malloc(sizeof(cplx) * n);
    cplx *t1 = malloc(sizeof(cplx) * n), *t2 = malloc(sizeof(cplx) * n);
#define Hij(i, j) H[(i) + (j) * (m + 1)]
    double bnorm = dznrm2(b, n);
    double xnorm = dznrm2(x, n);
    double tol = backward ? 0.5 * (atol + rtol * xnorm) : atol + rtol * bnorm;
    int64_t iterations = 0, restarts = 0;
    steady_apply(op, x, w);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - w[i];
    double beta = dznrm2(r, n);
    int conv = 0;
    if (beta <= tol) { conv = 1; goto done; }
    while (iterations < maxiter) {
        memset(H, 0, sizeof(cplx) * (m + 1) * m);
        memset(g, 0, sizeof(cplx) * (m + 1));
        for (int64_t i = 0; i < n; ++i) V[i] = r[i] / beta;
        g[0] = beta;
        int kdone = 0;
        int jmax = (int)((maxiter - iterations) < m ? (maxiter - iterations) : m);
        for (int j = 0; j < jmax; ++j) {
            inner_solve(op, pc, kin, V + (int64_t)j * n, Z + (int64_t)j * n, Vi, t1, t2);
            op->applications += kin;
            steady_apply(op, Z + (int64_t)j * n, w);
            for (int pass = 0; pass < 2; ++pass)
                for (int i = 0; i <= j; ++i) {
                    cplx hij = zdot(V + (int64_t)i * n, w, n);
                    Hij(i, j) += hij;
                    zaxpy_neg(w, hij, V + (int64_t)i * n, n);
                }
            Hij(j + 1, j) = dznrm2(w, n);
            if (Hij(j + 1, j) != 0)
                for (int64_t k = 0; k < n; ++k) V[(int64_t)(j + 1) * n + k] = w[k] / Hij(j + 1, j);
            for (int i = 0; i < j; ++i) {
                cplx h1 = Hij(i, j), h2 = Hij(i + 1, j);
                double c = cs[i]; cplx s = sn[i];
                Hij(i, j) = c * h1 + s * h2;
                Hij(i + 1, j) = -conj(s) * h1 + c * h2;
            }
            double c; cplx s;
            complex_givens(Hij(j, j), Hij(j + 1, j), &c, &s);
            cs[j] = c; sn[j] = s;
            cplx h1 = Hij(j, j), h2 = Hij(j + 1, j);
            Hij(j, j) = c * h1 + s * h2;
            Hij(j + 1, j) = 0;
            cplx g1 = g[j], g2 = g[j + 1];
            g[j] = c * g1 + s * g2;
            g[j + 1] = -conj(s) * g1 + c * g2;
            iterations++;
            kdone = j + 1;
            if (backward && (kdone % 5 == 0 || cabs(g[j + 1]) <= tol) && Hij(j, j) != 0) {
                for (int i = kdone - 1; i >= 0; --i) {
                    cplx sacc = g[i];
                    for (int jj = i + 1; jj < kdone; ++jj) sacc -= Hij(i, jj) * yv[jj];
                    yv[i] = sacc / Hij(i, i);
                }
                zgemv_cols(t1, Z, n, kdone, yv);
                for (int64_t k = 0; k < n; ++k) t1[k] += x[k];
                xnorm = dznrm2(t1, n);
                tol = 0.5 * (atol + rtol * xnorm);
            }
            if (cabs(g[j + 1]) <= tol) break;
        }
        for (int i = kdone - 1; i >= 0; --i) {
            cplx sacc = g[i];
            for (int jj = i + 1; jj < kdone; ++jj) sacc -= Hij(i, jj) * yv[jj];
            yv[i] = sacc / Hij(i, i);
        }
        zgemv_cols(w, Z, n, kdone, yv);
        for (int64_t k = 0; k < n; ++k) x[k] += w[k];
        restarts++;
        steady_apply(op, x, w);
        for (int64_t i = 0; i < n; ++i) r[i] = b[i] - w[i];
        beta = dznrm2(r, n);
        if (backward) { xnorm = dznrm2(x, n); tol = 0.5 * (atol + rtol * xnorm); }
        if (beta <= tol) { conv = 1; break; }
    }
done:
#undef Hij
    *iters_out = iterations; *restarts_out = restarts; *res_out = beta;
    free(V); free(Z); free(Vi); free(H); free(g); free(sn); free(yv); free(cs); free(w); free(r); free(t1); free(t2);
    return conv;
}

/* krylov_steady_state (src/krylov.jl:797-891) with x0 = v = t/(t.t), or the
 * given initial state.
 * operator_scale <= 0 selects the deterministic probe estimate. */
int oracle_steady_state(const oplan *P, int nthreads, const int64_t *tidx, const double *tvals, int64_t nt,
                        int krylovdim, int64_t maxiter, double atol, double rtol, double operator_scale,
                        cplx *x_out, steady_info *info, const cplx *pc_U, const cplx *pc_inv, int backward,
                        const double *sigma, int inner_steps, const cplx *x0) {
    int64_t n = P->n;
    g_blas_threads = nthreads > 0 ? nthreads : 1;
    cplx *tv = malloc(sizeof(cplx) * nt), *vv = malloc(sizeof(cplx) * nt);
    cplx *tw = malloc(sizeof(cplx) * nt), *vw = malloc(sizeof(cplx) * nt);
    cplx tt = 0;
    for (int64_t k = 0; k < nt; ++k) { tv[k] = tvals[k]; tt += conj(tv[k]) * tv[k]; }
    for (int64_t k = 0; k < nt; ++k) vv[k] = tv[k] / tt;
    /* sector weighting: iterate on u = S x; t -> t / sigma, v -> sigma v */
    for (int64_t k = 0; k < nt; ++k) {
        tw[k] = sigma ? tv[k] / sigma[tidx[k]] : tv[k];
        vw[k] = sigma ? vv[k] * sigma[tidx[k]] : vv[k];
    }
    double scale = operator_scale > 0 ? operator_scale : oracle_operator_scale(P, nthreads);
    cplx *scr = sigma ? malloc(sizeof(cplx) * n) : NULL;
    steady_op op = {P, nthreads, 1.0 / scale, tidx, tw, vw, nt, 0, sigma, scr};
    cplx *b = calloc(n, sizeof(cplx));
    for (int64_t k = 0; k < nt; ++k) b[tidx[k]] += vw[k];
    if (x0)  /* initial_state (src/krylov.jl:818-821), in the iteration's coordinates */
        for (int64_t i = 0; i < n; ++i) x_out[i] = sigma ? x0[i] * sigma[i] : x0[i];
    else
        memcpy(x_out, b, sizeof(cplx) * n);
    int64_t it = 0, rs = 0; double res = 0;
    sector_pc pc = {pc_U, pc_inv, scale};
    /* with weighting the iteration controls the weighted residual; if
       |L x| / s <= atol + rtol |x| does not follow on x = u / sigma, continue
       with a tenfold tighter threshold (at most four times) */
    int conv = 0;
    double tight = 1.0;
    cplx *lw = sigma ? malloc(sizeof(cplx) * n) : NULL;
    for (int pass = 0;; ++pass) {
        int64_t it1, rs1;
        if (inner_steps > 0 && pc_U)
            conv = fgmres(&op, x_out, b, krylovdim, maxiter - it, atol * tight, rtol * tight, &it1, &rs1, &res,
                          &pc, backward, inner_steps);
        else
            conv = gmres(&op, x_out, b, krylovdim, maxiter - it, atol * tight, rtol * tight, &it1, &rs1, &res,
                         pc_U ? &pc : NULL, backward);
        it += it1; rs += rs1;
        /* backward == 2 (weighted criterion): accept on the weighted residual, no tightening */
        if (!sigma || backward == 2 || !conv || pass >= 4 || it >= maxiter) break;
        for (int64_t i = 0; i < n; ++i) scr[i] = x_out[i] / sigma[i];
        do_apply(P, scr, lw, nthreads);
        op.applications++;
        if (dznrm2(lw, n) / scale <= atol + rtol * dznrm2(scr, n)) break;
        tight *= 0.1;
    }
    free(lw);
    if (sigma)
        for (int64_t i = 0; i < n; ++i) x_out[i] /= sigma[i];
    free(scr); free(tw); free(vw);
    if (backward) {  /* restore the trace row, weighted by |v| in the residual */
        cplx tr0 = 0;
        for (int64_t k = 0; k < nt; ++k) tr0 += conj(tv[k]) * x_out[tidx[k]];
        for (int64_t i = 0; i < n; ++i) x_out[i] /= tr0;
    }
    cplx *w = malloc(sizeof(cplx) * n);
    do_apply(P, x_out, w, nthreads);
    double resid = dznrm2(w, n);
    double nres = resid / scale;
    cplx tr = 0;
    for (int64_t k = 0; k < nt; ++k) tr += conj(tv[k]) * x_out[tidx[k]];
    double terr = cabs(tr - 1.0);
    double xn = dznrm2(x_out, n);
    double tol = backward ? atol + rtol * xn : atol + rtol * (xn > 1.0 ? xn : 1.0);
    /* weighted relative residual |S L x| / (s |S x|) (when the weights are given) */
    double wres = NAN;
    if (sigma) {
        cplx *sx = malloc(sizeof(cplx) * n);
        for (int64_t i = 0; i < n; ++i) sx[i] = sigma[i] * w[i];
        double a = dznrm2(sx, n);
        for (int64_t i = 0; i < n; ++i) sx[i] = sigma[i] * x_out[i];
        double b2 = dznrm2(sx, n);
        wres = b2 > 0 ? a / (scale * b2) : INFINITY;
        free(sx);
    }
    info->weighted_relative_residual = wres;
    info->converged = conv && terr <= atol + rtol && (backward == 2 ? wres <= atol + rtol : nres <= tol);
    info->iterations = it; info->restarts = rs; info->applications = op.applications;
    info->residual = resid; info->normalized_residual = nres; info->trace_error = terr;
    info->operator_scale = scale; info->linear_residual = res;
    free(tv); free(vv); free(b); free(w);
    return info->converged ? 0 : 1;
}

/* ---------------- small dense matrix exponential ----------------
 * Scaling and squaring with Pade approximants (Higham, SIAM J. Matrix Anal.
 * Appl. 26 (2005) 1179-1193): degree 3/5/7/9/13 chosen by the 1-norm. */

static void mat_mul(int q, const cplx *A, const cplx *B, cplx *C) {
    for (int j = 0; j < q; ++j)
        for (int i = 0; i < q; ++i) {
            cplx s = 0;
            for (int k = 0; k < q; ++k) s += A[i + k * q] * B[k + j * q];
            C[i + j * q] = s;
        }
}

static int lu_solve(int q, cplx *A, cplx *B) { /* solves A X = B in place, B q x q */
    int *piv = malloc(sizeof(int) * q);
    for (int k = 0; k < q; ++k) {
        int p = k; double best = cabs(A[k + k * q]);
        for (int i = k + 1; i < q; ++i) if (cabs(A[i + k * q]) > best) { best = cabs(A[i + k * q]); p = i; }
        piv[k] = p;
        if (p != k) {
            for (int j = 0; j < q; ++j) { cplx t = A[k + j * q]; A[k + j * q] = A[p + j * q]; A[p + j * q] = t; }
            for (int j = 0; j < q; ++j) { cplx t = B[k + j * q]; B[k + j * q] = B[p + j * q]; B[p + j * q] = t; }
        }
        if (A[k + k * q] == 0) { free(piv); return 1; }
        for (int i = k + 1; i < q; ++i) {
            cplx f = A[i + k * q] / A[k + k * q];
            A[i + k * q] = f;
            for (int j = k + 1; j < q; ++j) A[i + j * q] -= f * A[k + j * q];
            for (int j = 0; j < q; ++j) B[i + j * q] -= f * B[k + j * q];
        }
    }
    for (int j = 0; j < q; ++j)
        for (int i = q - 1; i >= 0; --i) {
            cplx s = B[i + j * q];
            for (int k = i + 1; k < q; ++k) s -= A[i + k * q] * B[k + j * q];
            B[i + j * q] = s / A[i + i * q];
        }
    free(piv);
    return 0;
}

void oracle_expm(int q, const cplx *Ain, cplx *E) {
    static const double b13[] = {64764752532480000., 32382376266240000., 7771770303897600.,
        1187353796428800., 129060195264000., 10559470521600., 670442572800.,
        33522128640., 1323241920., 40840800., 960960., 16380., 182., 1.};
    static const double b9[] = {17643225600., 8821612800., 2075673600., 302702400., 30270240.,
        2162160., 110880., 3960., 90., 1.};
    static const double b7[] = {17297280., 8648640., 1995840., 277200., 25200., 1512., 56., 1.};
    static const double b5[] = {30240., 15120., 3360., 420., 30., 1.};
    static const double b3[] = {120., 60., 12., 1.};
    static const double theta[] = {1.495585217958292e-2, 2.539398330063230e-1, 9.504178996162932e-1,
        2.097847961257068e0, 5.371920351148152e0};
    size_t sz = sizeof(cplx) * q * q;
    cplx *A = malloc(sz), *A2 = malloc(sz), *A4 = malloc(sz), *A6 = malloc(sz), *U = malloc(sz),
         *Vm = malloc(sz), *T = malloc(sz), *P = malloc(sz), *Qm = malloc(sz);
    memcpy(A, Ain, sz);
    double nrm = 0;
    for (int j = 0; j < q; ++j) {
        double c = 0;
        for (int i = 0; i < q; ++i) c += cabs(A[i + j * q]);
        if (c > nrm) nrm = c;
    }
    int s = 0, deg = 13;
    const double *b = b13;
    if (nrm <= theta[0]) { deg = 3; b = b3; }
    else if (nrm <= theta[1]) { deg = 5; b = b5; }
    else if (nrm <= theta[2]) { deg = 7; b = b7; }
    else if (nrm <= theta[3]) { deg = 9; b = b9; }
    else {
        if (nrm > theta[4]) s = (int)ceil(log2(nrm / theta[4]));
        if (s < 0) s = 0;
        double f = ldexp(1.0, -s);
        for (int i = 0; i < q * q; ++i) A[i] *= f;
    }
    mat_mul(q, A, A, A2);
    memset(U, 0, sz); memset(Vm, 0, sz);
    if (deg == 13) {
        mat_mul(q, A2, A2, A4);
        mat_mul(q, A4, A2, A6);
        /* U = A (A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I) */
        for (int i = 0; i < q * q; ++i) T[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
        mat_mul(q, A6, T, P);
        for (int i = 0; i < q * q; ++i) P[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i];
        for (int i = 0; i < q; ++i) P[i + i * q] += b[1];
        mat_mul(q, A, P, U);
        for (int i = 0; i < q * q; ++i) T[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
        mat_mul(q, A6, T, Vm);
        for (int i = 0; i < q * q; ++i) Vm[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i];
        for (int i = 0; i < q; ++i) Vm[i + i * q] += b[0];
    } else {
        /* powers A2^k, odd part into U, even part into V */
        memset(P, 0, sz);
        for (int i = 0; i < q; ++i) P[i + i * q] = 1.0; /* P = A2^0 */
        memset(T, 0, sz);
        for (int k = 0; 2 * k <= deg; ++k) {
            for (int i = 0; i < q * q; ++i) { T[i] += b[2 * k + 1] * P[i]; Vm[i] += b[2 * k] * P[i]; }
            if (2 * (k + 1) <= deg) { mat_mul(q, P, A2, Qm); memcpy(P, Qm, sz); }
        }
        mat_mul(q, A, T, U);
    }
    for (int i = 0; i < q * q; ++i) { P[i] = Vm[i] + U[i]; Qm[i] = Vm[i] - U[i]; }
    lu_solve(q, Qm, P); /* P = (V-U)^{-1} (V+U) */
    for (int k = 0; k < s; ++k) { mat_mul(q, P, P, T); memcpy(P, T, sz); }
    memcpy(E, P, sz);
    free(A); free(A2); free(A4); free(A6); free(U); free(Vm); free(T); free(P); free(Qm);
}

/* ---------------- krylov_expv! (src/krylov_extensions.jl:1346-1568) ---------------- */

typedef struct {
    int32_t converged;
    int64_t accepted, rejected, applications, factorizations, attempts;
    double reached, estimated_error;
} expv_info;

typedef struct {
    int64_t n; int m;
    cplx *V, *H, *small, *w, *current, *trial;
    /* factorization summary */
    double beta, next_remainder;
    int zero_state, k, used, completed, enlarged, q, coeff_count;
} expv_ws;

static void arnoldi_factorization(expv_ws *W, const oplan *P, int nthreads) {
    int64_t n = W->n; int m = W->m;
    cplx *V = W->V;
#define HH(i, j) W->H[(i) + (j) * (m + 1)]
#define SS(i, j) W->small[(i) + (j) * (m + 1)]
    W->beta = dznrm2(W->current, n);
    W->zero_state = 0; W->k = 0; W->used = 0; W->completed = 0; W->enlarged = 0; W->q = 0;
    W->coeff_count = 0; W->next_remainder = 0;
    if (W->beta == 0) { W->zero_state = 1; W->completed = 1; return; }
    memset(W->H, 0, sizeof(cplx) * (m + 1) * m);
    for (int64_t i = 0; i < n; ++i) V[i] = W->current[i] / W->beta;
    int k = 0, exact_breakdown = 0, near_breakdown = 0;
    double breakfactor = sqrt(2.220446049250313e-16);
    for (int j = 0; j < m; ++j) {
        do_apply(P, V + (int64_t)j * n, W->w, nthreads);
        double image_norm = dznrm2(W->w, n);
        for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i <= j; ++i) {
                cplx a = zdot(V + (int64_t)i * n, W->w, n);
                HH(i, j) += a;
                zaxpy_neg(W->w, a, V + (int64_t)i * n, n);
            }
        HH(j + 1, j) = dznrm2(W->w, n);
        k = j + 1;
        if (HH(j + 1, j) == 0 || (int64_t)(j + 1) == n) { exact_breakdown = 1; break; }
        double floor_ = image_norm > 2.2250738585072014e-308 ? image_norm : 2.2250738585072014e-308;
        near_breakdown |= cabs(HH(j + 1, j)) <= breakfactor * floor_;
        for (int64_t t = 0; t < n; ++t) V[(int64_t)(j + 1) * n + t] = W->w[t] / HH(j + 1, j);
    }
    int q = k + (exact_breakdown ? 0 : 1);
    memset(W->small, 0, sizeof(cplx) * (m + 1) * (m + 1));
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) SS(i, j) = HH(i, j);
    if (!exact_breakdown) SS(k, k - 1) = HH(k, k - 1);
    int used = k, enlarged = 0;
    double next_remainder = 0;
    if (!exact_breakdown && near_breakdown) {
        enlarged = 1; used += 1;
        do_apply(P, V + (int64_t)(q - 1) * n, W->w, nthreads);
        for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i < q; ++i) {
                cplx a = zdot(V + (int64_t)i * n, W->w, n);
                SS(i, q - 1) += a;
                zaxpy_neg(W->w, a, V + (int64_t)i * n, n);
            }
        next_remainder = dznrm2(W->w, n);
    }
    W->k = k; W->q = q; W->used = used; W->enlarged = enlarged; W->next_remainder = next_remainder;
    W->coeff_count = enlarged ? q : k;
    W->completed = exact_breakdown || (enlarged && ((int64_t)q == n || next_remainder == 0));
#undef HH
#undef SS
}

/* returns error estimate; fills trial */
static double project_factorization(expv_ws *W, double h) {
    int64_t n = W->n; int m = W->m;
    if (W->zero_state) { memset(W->trial, 0, sizeof(cplx) * n); return 0.0; }
    int q = W->q;
    cplx *S = malloc(sizeof(cplx) * q * q), *E = malloc(sizeof(cplx) * q * q);
    for (int j = 0; j < q; ++j)
        for (int i = 0; i < q; ++i) S[i + j * q] = h * W->small[i + j * (m + 1)];
    oracle_expm(q, S, E);
    int cc = W->coeff_count;
#pragma omp parallel for schedule(static) num_threads(g_blas_threads) if (n >= 65536)
    for (int64_t t = 0; t < n; ++t) {
        cplx z = 0;
        for (int j = 0; j < cc; ++j) z += W->V[(int64_t)j * n + t] * E[j];
        W->trial[t] = z * W->beta;
    }
    double err;
    if (W->completed) err = 0.0;
    else if (W->enlarged) err = W->beta * fabs(h) * W->next_remainder * cabs(E[(q - 1)]);
    else err = W->beta * cabs(E[W->k]);
    free(S); free(E);
    return err;
}

int oracle_expv(const oplan *P, int nthreads, cplx *y, const cplx *b, double t, int krylovdim,
                double atol, double rtol, double safety, int64_t max_steps, expv_info *info) {
    g_blas_threads = nthreads > 0 ? nthreads : 1;
    int64_t n = P->n;
    int m = krylovdim < n ? krylovdim : (int)n;
    expv_ws W;
    W.n = n; W.m = m;
    W.V = malloc(sizeof(cplx) * n * (m + 1));
    W.H = malloc(sizeof(cplx) * (m + 1) * m);
    W.small = malloc(sizeof(cplx) * (m + 1) * (m + 1));
    W.w = malloc(sizeof(cplx) * n); W.current = malloc(sizeof(cplx) * n); W.trial = malloc(sizeof(cplx) * n);
    memset(info, 0, sizeof(*info));
    double total = fabs(t);
    if (total == 0) { memmove(y, b, sizeof(cplx) * n); info->converged = 1; goto out; }
    double hmax = total, hmin = 0.0, h0 = hmax;
    double h = h0 < hmax ? h0 : hmax;
    memmove(W.current, b, sizeof(cplx) * n);
    double reached = 0.0, error_sum = 0.0, direction = t > 0 ? 1.0 : -1.0;
    int64_t accepted = 0, rejected = 0, applications = 0, factorizations = 0, attempts = 0;
    int have_fact = 0;
    while (reached < total && attempts < max_steps) {
        attempts++;
        if (h > total - reached) h = total - reached;
        if (!have_fact) {
            arnoldi_factorization(&W, P, nthreads);
            applications += W.used; factorizations++; have_fact = 1;
        }
        double err = project_factorization(&W, direction * h);
        int breakdown = W.completed;
        int finite = isfinite(err);
        for (int64_t i = 0; finite && i < n; ++i) finite = isfinite(creal(W.trial[i])) && isfinite(cimag(W.trial[i]));
        double trial_norm = finite ? dznrm2(W.trial, n) : dznrm2(W.current, n);
        double fraction = h / total;
        double localtol = fraction * (atol + rtol * trial_norm);
        int accept = finite && (breakdown || err <= localtol);
        if (accept) {
            memcpy(W.current, W.trial, sizeof(cplx) * n);
            double remaining = total - reached;
            double next = (h == remaining) ? total : reached + h;
            reached = next; accepted++; error_sum += err;
            have_fact = 0;
        } else rejected++;
        double factor;
        if (!finite) factor = 0.2;
        else if (breakdown) factor = 5.0;
        else if (err == 0) factor = 2.0;
        else {
            double ex = 1.0 / (double)(W.k > 1 ? W.k : 1);
            factor = safety * pow(localtol / err, ex);
            if (factor < 0.2) factor = 0.2;
            if (factor > 5.0) factor = 5.0;
        }
        if (!accept && factor > 0.5) factor = 0.5;
        double hnew = h * factor;
        if (hnew > hmax) hnew = hmax;
        if (!accept && (hnew < hmin || hnew == 0 || (reached > 0 && reached + hnew == reached))) break;
        h = hnew > hmin ? hnew : hmin;
    }
    info->converged = reached == total;
    memmove(y, W.current, sizeof(cplx) * n);
    info->accepted = accepted; info->rejected = rejected; info->applications = applications;
    info->factorizations = factorizations; info->attempts = attempts;
    info->reached = direction * reached; info->estimated_error = error_sum;
out:
    free(W.V); free(W.H); free(W.small); free(W.w); free(W.current); free(W.trial);
    return info->converged ? 0 : 1;
}

/* solve_dynamics with ExpvAlgorithm and streamed observables
 * (src/highlevel.jl:614-733): obs_out[k*nobs + o] = dot(obs_o, rho(t_k)). */
int oracle_solve_dynamics(const oplan *P, int nthreads, cplx *rho, const double *ts, int64_t nts,
                          const cplx *obs, int32_t nobs, cplx *obs_out, int krylovdim, double atol,
                          double rtol, double safety, int64_t max_steps, expv_info *total_info) {
    int64_t n = P->n;
    memset(total_info, 0, sizeof(*total_info));
    total_info->converged = 1;
    for (int64_t k = 0; k < nts; ++k) {
        if (k > 0) {
            double dt = ts[k] - ts[k - 1];
            if (dt != 0) {
                expv_info inf;
                int rc = oracle_expv(P, nthreads, rho, rho, dt, krylovdim, atol, rtol, safety, max_steps, &inf);
                total_info->accepted += inf.accepted; total_info->rejected += inf.rejected;
                total_info->applications += inf.applications; total_info->factorizations += inf.factorizations;
                total_info->attempts += inf.attempts; total_info->estimated_error += inf.estimated_error;
                total_info->reached = ts[k];
                if (rc) { total_info->converged = 0; return 1; }
            }
        }
        for (int32_t o = 0; o < nobs; ++o) obs_out[k * nobs + o] = zdot(obs + (int64_t)o * n, rho, n);
    }
    return 0;
}
