/*
 * pimi.h -- C ABI of the MI355X backend for the permutation-invariant (PI)
 * Liouvillian hot path: Schur-block Liouvillian apply inside GMRES
 * stationary-state solves and adaptive Krylov exponential dynamics.
 *
 * Plain pointers and sizes only; complex scalars are interleaved
 * (re, im) pairs of double or float.  Every function returns a pimi_status;
 * pimi_last_error() gives a thread-local message for the last failure.
 *
 * Reference interfaces each entry point replaces (PermutationalInvariantDynamics.jl):
 *   pimi_compile              compile(model; backend=...)        src/liouvillian.jl:4040-4098
 *                              with LiouvillianPlan lowering      src/liouvillian.jl:1290-1371,1700-1851
 *   pimi_plan_create_fused    upload of an already lowered FusedStaticPIKernel
 *                              (accelerate hook)                  src/accelerators.jl:348-386
 *   pimi_apply                apply!(y, compiled, x, t, p, work) src/liouvillian.jl:3215-3251
 *                              threaded_apply!                    src/threaded_apply.jl:564-623
 *   pimi_steady_gmres         krylov_steady_state / _gmres!      src/krylov.jl:732-891
 *   pimi_expv_series          solve_dynamics(...; ExpvAlgorithm) src/highlevel.jl:559-733
 *                              with krylov_expv!                  src/krylov_extensions.jl:1452-1568
 *   pimi_trace_functional     _trace_functional                  src/source_protocol.jl:15-34
 *   pimi_collective_operator  collective_operator                src/geometry.jl:1282-1308
 */
#ifndef PIMI_H
#define PIMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pimi_status;
enum {
    PIMI_OK = 0,
    PIMI_EINVAL = 1,         /* ArgumentError */
    PIMI_EDIM = 2,           /* DimensionMismatch */
    PIMI_ENOMEM = 3,         /* memory budget or allocation failure (checked before allocating) */
    PIMI_EDEVICE = 4,        /* HIP runtime error */
    PIMI_EBUSY = 5,          /* workspace already in use by another caller */
    PIMI_ENOTCONVERGED = 6,  /* solver did not converge; info is filled */
    PIMI_ERANGE = 7          /* exact coefficient outside the supported exact range */
};

typedef enum { PIMI_C64 = 0, PIMI_C32 = 1 } pimi_precision;

typedef struct pimi_context pimi_context;
typedef struct pimi_plan pimi_plan;
typedef struct pimi_workspace pimi_workspace;

/* Model terms (src/terms.jl:118-160).  op/op2 are d*d complex matrices,
 * column-major, interleaved doubles. */
typedef enum {
    PIMI_TERM_LOCAL_H = 0,          /* LocalHamiltonian(op; rate, hbar) */
    PIMI_TERM_COLLECTIVE_H = 1,     /* CollectiveHamiltonian(op; rate, hbar) */
    PIMI_TERM_LOCAL_JUMP = 2,       /* LocalJump(op; rate) */
    PIMI_TERM_COLLECTIVE_JUMP = 3,  /* CollectiveJump(op; rate) */
    PIMI_TERM_DIRECT_PRODUCT_H = 4  /* DirectPIHamiltonian(collective(op)*collective(op2); rate, hbar) */
} pimi_term_kind;

typedef struct {
    int32_t kind;
    int32_t reserved;
    const double *op;
    const double *op2;
    double rate;
    double hbar;
} pimi_term;

typedef struct {
    int32_t N;
    int32_t d;
    int32_t nterms;
    int32_t reserved;
    const pimi_term *terms;
} pimi_model_desc;

/* A lowered FusedStaticPIKernel (src/liouvillian.jl:539-545) as flat arrays.
 * CSC block k occupies colptr[cp_base[k] .. cp_base[k]+m], whose entries index
 * rows/vals directly.  Collective-gain blocks are gain-major (g*ns + s).
 * One-body branch b of gain g (ob_br_base[g] <= b < ob_br_base[g+1]) maps
 * input sector br_in[b] into output sector br_out[b] with scale
 * ob_scale[g]*br_r[b]; its contraction C (m_out x m_in) is given either as an
 * exact support list (br_use_support[b] = 1, entries sup_base[b]..sup_base[b+1])
 * or densely, column-major, at dense[br_dense_base[b]]. */
typedef struct {
    const int64_t *cp_base;
    const int64_t *colptr;
    const int32_t *rows;
    const double *vals;   /* complex */
} pimi_csc_set;

typedef struct {
    int32_t ns;
    int32_t has_h;
    int32_t has_q;
    int32_t ncg;
    int32_t nob;
    int32_t reserved;
    const int64_t *dims;
    pimi_csc_set H, Q, K;
    const double *cg_scale;       /* complex, ncg */
    const double *ob_scale;       /* complex, nob */
    const int64_t *ob_br_base;    /* nob+1 */
    const int32_t *br_out, *br_in;
    const double *br_r;
    const int32_t *br_use_support;
    const int64_t *br_sup_base;   /* nbr+1 */
    const int32_t *sup_rows, *sup_cols;
    const double *sup_vals;       /* complex */
    const int64_t *br_dense_base; /* nbr */
    const double *dense;          /* complex */
} pimi_fused_desc;

typedef struct {
    int64_t n;              /* PI dimension */
    int32_t nsectors;
    int32_t max_block;
    int64_t nterms;         /* stencil terms over all sectors */
    int64_t stencil_entries;/* sum over outputs of A-row x B-row widths */
    int64_t table_bytes_c64;/* device stencil table bytes (complex FP64 tables) */
    int64_t table_bytes_c32;
    int32_t apply_form;     /* 1: banded (diagonal) apply tables, 0: ELL tables (see PIMI_APPLY_FORM) */
    int32_t banded_groups;  /* banded form: chains (one strip load per tile each) over all sectors */
} pimi_plan_info;

typedef enum {
    PIMI_PRECOND_NONE = 0,   /* plain restarted GMRES, as the reference's default */
    PIMI_PRECOND_SECTOR = 1, /* right preconditioning by the per-sector inverse of X -> G X + X G^H,
                                G = -iH - Q/2 with Q replaced by its diagonal in the H eigenbasis */
    PIMI_PRECOND_SECTOR_INNER = 2 /* flexible GMRES whose directions come from `inner_steps` steps of
                                     PIMI_PRECOND_SECTOR-preconditioned GMRES on the sector-block-diagonal
                                     part of L, in the workspace precision (complex FP32 workspaces: FP32
                                     inner solves, FP64 outer iteration and residuals) */
} pimi_preconditioner;

typedef struct {
    int32_t krylovdim;      /* GMRESAlgorithm(krylovdim=30, ...) */
    int32_t preconditioner; /* pimi_preconditioner */
    int64_t maxiter;        /* GMRESAlgorithm(maxiter=500) */
    double atol, rtol;      /* convergence: |r| <= atol + rtol |b| */
    double operator_scale;  /* <= 0: deterministic 3-probe estimate */
    /* mixed precision (pimi_precision of the workspace = PIMI_C32):
       inner tolerance and maximum outer refinement sweeps */
    double inner_rtol;
    int64_t max_refinements;
    /* PIMI_CRITERION_REFERENCE: |r| <= atol + rtol |b| and |L x|/s <= atol + rtol max(|x|, 1)
         (src/krylov.jl:743-746,868-873);
       PIMI_CRITERION_BACKWARD:  |r| <= (atol + rtol |x|) / 2 and |L x|/s <= atol + rtol |x|, i.e. for
         atol = 0 the scale-invariant relative residual |L x|/(s |x|) <= rtol (t^H L = 0 bounds
         |L x|/s by 2 |r|); at large N |b| = 2^(-N/2) is far below eps |x|, where the
         reference rule cannot be met in floating point.
       PIMI_CRITERION_WEIGHTED (with PIMI_WEIGHT_SECTOR only): the backward rule on the weighted
         iterate u = S x, accepted when the weighted relative residual |S L x|/(s |S x|) <= atol + rtol.
         The coefficient-space residual |L x|/(s |x|) weighs a sector by its coefficients, which
         scale as 1/sqrt(f^nu) of its population: at large N (f^nu up to 1e34 at N = 120) it leaves
         the populated sectors, and every observable, unresolved, while the weighted residual
         weighs sectors as traces and observables do.  No counterpart in the reference. */
    int32_t criterion;
    int32_t inner_steps;    /* PIMI_PRECOND_SECTOR_INNER: inner GMRES steps per outer step */
    /* PIMI_WEIGHT_NONE: iterate on the PI coefficients x, as the reference does.
       PIMI_WEIGHT_SECTOR: iterate on u = S x, S = sqrt(f^nu) on sector nu (u_nu = f^nu times
         the physical block), i.e. on S (L/s + v t^H) S^{-1}; the trace functional becomes 1 on
         the block diagonals.  Residual norms then weigh each sector by its multiplicity, as
         traces and observables do; with PIMI_WEIGHT_NONE an error that is small against |x|
         but lies in a sector of large f^nu (up to 1e17 at N = 120) moves observables by up to
         sqrt(f^nu) times as much.  No counterpart in the reference.  Needs the backward
         criterion; the reported relative residual is still |L x| / (s |x|) on x. */
    int32_t weighting;
    int32_t reserved;
} pimi_gmres_opts;

enum { PIMI_CRITERION_REFERENCE = 0, PIMI_CRITERION_BACKWARD = 1, PIMI_CRITERION_WEIGHTED = 2 };
enum { PIMI_WEIGHT_NONE = 0, PIMI_WEIGHT_SECTOR = 1 };

typedef struct {
    int32_t krylovdim;      /* ExpvAlgorithm(krylovdim=30, ...) */
    int32_t reserved;
    double atol, rtol, safety;
    int64_t max_steps;
} pimi_expv_opts;

typedef struct {
    int32_t converged;
    int32_t reserved;
    int64_t iterations, restarts, applications;
    int64_t accepted, rejected, factorizations, refinements;
    double residual;            /* |L x|_2 */
    double normalized_residual; /* |L x|_2 / operator_scale */
    double trace_error;         /* |t.x - 1| */
    double operator_scale;
    double linear_residual;
    double estimated_error;
    double seconds;
    double relative_residual;   /* |L x|_2 / (operator_scale |x|_2) */
    double weighted_relative_residual; /* |S L x|_2 / (operator_scale |S x|_2), S = sqrt(f^nu) per sector
                                          (stationary single solves; NaN where not formed) */
} pimi_solve_info;

const char *pimi_last_error(void);
const char *pimi_version(void);

pimi_status pimi_context_create(int32_t device_ordinal, pimi_context **out);
void pimi_context_destroy(pimi_context *ctx);

/* compile(model; backend=:mi355x): lower the model on the host (exact
 * rationals) and upload the stencil tables once.  ctx == NULL gives a
 * host-only plan (lowering, export, trace functional, collective operators;
 * no workspaces). */
pimi_status pimi_compile(pimi_context *ctx, const pimi_model_desc *model, pimi_plan **out);
/* Upload an already lowered fused plan (e.g. the reference's own LiouvillianPlan). */
pimi_status pimi_plan_create_fused(pimi_context *ctx, const pimi_fused_desc *desc, pimi_plan **out);
pimi_status pimi_plan_info_get(const pimi_plan *plan, pimi_plan_info *info);
/* dims[ns] per-sector block dimensions */
pimi_status pimi_plan_dims(const pimi_plan *plan, int64_t *dims);
/* trace functional: nnz = sum of dims; idx/vals host buffers of that length */
pimi_status pimi_trace_functional(const pimi_plan *plan, int64_t *idx, double *vals);
/* PI coefficient vector of sum_i X_i for a d x d complex X (host in, host out) */
pimi_status pimi_collective_operator(const pimi_plan *plan, const double *X, double *out);
/* iid_state (src/operators.jl:146-250): PI coefficient vector (n complex,
 * host) of rho^{otimes N} for a d x d one-particle matrix rho (complex,
 * column-major, host), by the one-box Clebsch-Gordan recurrence
 * D_lambda = sum_{i,o} rho[i,o] U_i D_mu U_o^T over particle numbers 1..N.
 * rho is used as given (the Python and Julia callers validate Hermiticity,
 * positivity and trace first, as the reference does). */
pimi_status pimi_iid_state(const pimi_plan *plan, const double *rho, double *out);
/* The banded apply tables of a qubit plan (host copies, for checking the
 * table construction against the operator on the CPU): counts[4] = chain
 * base entries (ns + 1), chains, entries, coefficient slots (complex); each
 * array argument may be NULL.  chains: 7 int64 per chain (input block
 * offset, input block size, row shift of the loaded strip, first strip
 * column, strip span, entry range [e0, e1)); ents: 7 int32 per entry (kind,
 * strip column, qa, qb, qa2, qb2, row shift beyond the chain's); coef:
 * complex doubles.  counts are zero for plans on ELL tables.  No
 * counterpart in the reference (a device table layout). */
pimi_status pimi_plan_banded_export(const pimi_plan *plan, int64_t *counts, int32_t *cbase, int64_t *chains,
                                    int32_t *ents, double *coef);
/* The compiled banded programs of a qubit plan (straight-line device code
 * per sector structure, built with hiprtc when a device plan is created):
 * the module source (copied into source when non-NULL, at most cap - 1
 * characters and a terminating NUL; *len = its full length), the number of
 * distinct programs and each output sector's program id (pid: ns int32,
 * -1 = the sector's tables are interpreted; may be NULL).  compile = 1 also
 * compiles the source for gfx950 (no device needed; the FP32 kernels of tile
 * width 8) and reports a failure as PIMI_EDEVICE.  A device plan compiles
 * each kernel on its first launch outside graph capture (once per process
 * for every plan of the same structure); PIMI_DIA_PROGRAMS=0 keeps every
 * plan on the interpreting kernels.  No counterpart in the reference (a
 * device code generator). */
pimi_status pimi_plan_banded_programs(const pimi_plan *plan, int32_t compile, char *source, int64_t cap,
                                      int64_t *len, int32_t *nprograms, int32_t *pid);
/* fused plan export of a compiled model (host arrays); lengths via pimi_plan_export_sizes */
pimi_status pimi_plan_export_sizes(const pimi_plan *plan, int64_t *sizes /* 12 entries */);
pimi_status pimi_plan_export(const pimi_plan *plan, int64_t *cp_base, int64_t *colptr, int32_t *rows,
                             double *vals, double *cg_scale, double *ob_scale, int64_t *ob_br_base,
                             int32_t *br_out, int32_t *br_in, double *br_r, int32_t *br_use_support,
                             int64_t *br_sup_base, int32_t *sup_rows, int32_t *sup_cols, double *sup_vals,
                             int64_t *br_dense_base, double *dense);
/* Host copies (n complex each, PI block layout) of the stationary-solve
 * preconditioner factors: U (eigenvectors of each fused Hamiltonian block)
 * and inv_den (see pimi_preconditioner).  No counterpart in the reference. */
pimi_status pimi_sector_preconditioner_factors(const pimi_plan *plan, double *U, double *inv_den);
/* Build a device plan's stationary-solver tables ahead of use (the sector
 * preconditioner factors and the sector-weighting tables), so that a later
 * solve or batch submission does no host setup.  Tables built on first use
 * are built once per plan under a per-plan lock, so workspaces of one plan
 * may run on several threads.  No counterpart in the reference. */
pimi_status pimi_plan_prepare(pimi_plan *plan);
void pimi_plan_destroy(pimi_plan *plan);

/* One workspace per concurrent caller: owns a HIP stream and the Krylov basis. */
pimi_status pimi_workspace_create(pimi_plan *plan, int32_t precision, int32_t krylovdim, pimi_workspace **out);
void pimi_workspace_destroy(pimi_workspace *ws);

/* y = L x for ncols column vectors of length n (column-major).
 * on_device = 1: x, y are device pointers (complex of the workspace
 * precision); 0: host pointers (complex double). */
pimi_status pimi_apply(pimi_workspace *ws, const void *x, void *y, int64_t ncols, int32_t on_device);

/* y = L^dagger x, the adjoint under the PI (Hilbert-Schmidt) inner product
 * (apply_adjoint!, src/liouvillian.jl:3275-3293; fused kernel 3107-3143):
 * i[H, X] - {Q, X}/2 + sum conj(c) K^H X K + sum_branches conj(s) C^H X C.
 * Same arguments as pimi_apply; the adjoint stencil is built on first use. */
pimi_status pimi_apply_adjoint(pimi_workspace *ws, const void *x, void *y, int64_t ncols, int32_t on_device);

/* Stationary state: solve (L/s + v t^H) rho = v from rho0 = x0, or from
 * rho0 = v when x0 is NULL (krylov_steady_state's initial_state,
 * src/krylov.jl:797-826).  x0 (may be NULL) and rho_out are host buffers of
 * n complex doubles. */
pimi_status pimi_steady_gmres(pimi_workspace *ws, const pimi_gmres_opts *opts, const double *x0, double *rho_out,
                              pimi_solve_info *info);

/* Dynamics: rho(t_k) by expv over each interval, recording
 * obs_out[k*nobs + o] = <obs_o, rho(t_k)>.  rho0/obs/rho_final host buffers. */
pimi_status pimi_expv_series(pimi_workspace *ws, const double *rho0, const double *ts, int64_t nts,
                             const pimi_expv_opts *opts, int32_t nobs, const double *obs, double *obs_out,
                             double *rho_final, pimi_solve_info *info);

/* Throughput probe: `reps` back-to-back applies to ncols synthetic columns
 * resident on the device, timed with HIP events on the workspace stream;
 * returns the average milliseconds per apply launch.  No counterpart in the
 * reference (measurement only). */
pimi_status pimi_apply_bench(pimi_workspace *ws, int64_t ncols, int32_t reps, double *ms_avg);

/* Autotuner of the banded apply's tile schedule: times `reps` applies of this
 * workspace's plan at each tile width (4, 8 and 16 columns per lane) for one
 * vector and for `ncols` vectors, in the workspace precision, on synthetic
 * device vectors, and keeps the fastest width of each for every later apply
 * of the plan in that precision.  widths_out[0], widths_out[1]: the chosen
 * single-vector and multi-column widths (0 when the plan uses ELL tables,
 * which have no width to tune).  No counterpart in the reference (the
 * schedule of a device kernel). */
pimi_status pimi_workspace_autotune_apply(pimi_workspace *ws, int64_t ncols, int32_t reps, int32_t *widths_out);

/* Device-to-device copy of `bytes` bytes, `reps` times, event-timed: the
 * calibration kernel for HBM byte counters.  Measurement only. */
pimi_status pimi_copy_bench(pimi_workspace *ws, int64_t bytes, int32_t reps, double *ms_avg);

/* y = scale * M^{-1} x with the sector preconditioner on the device (host
 * buffers, complex double; computed in the workspace precision).  For tests;
 * no counterpart in the reference. */
pimi_status pimi_precond_apply(pimi_workspace *ws, const double *x, double *y, double scale);

/* Throughput probe of the sector preconditioner (four grouped MFMA GEMMs per
 * application, in the workspace precision) on a synthetic device vector,
 * `reps` times, event-timed on the workspace stream: average milliseconds
 * per application and its algorithmic flops: 4 x 8 sum_nu m_nu^3 for complex
 * factors U, 4 x 4 sum_nu m_nu^3 when every U entry is real (a real
 * Hamiltonian block; each product is then complex by real).
 * Measurement only. */
pimi_status pimi_precond_bench(pimi_workspace *ws, int32_t reps, double *ms_avg, double *flops);

/* Time of the last apply kernel launch sequence, in milliseconds (HIP events
 * recorded on the workspace stream around pimi_apply). */
pimi_status pimi_workspace_last_apply_ms(const pimi_workspace *ws, double *ms);

/* ---------------------------------------------------------------------------
 * Conditional (no-jump) trajectory cohorts (src/batched_trajectories.jl:266-598):
 * ncols PI states in one n x ncols column-major matrix (host, or device with
 * on_device = 1), complex FP64 workspace.  Jump channels are the model's jump
 * terms in term order, numbered 1..count.  Intensities follow the reference's
 * sign check (nonfinite or below -1e-11: PIMI_EINVAL) and are clamped at zero.
 * ------------------------------------------------------------------------- */
pimi_status pimi_trajectory_channels(const pimi_plan *plan, int32_t *count);
/* batched_conditional_action!: Y = -i[H, X] - {E, X}/2 + lambda X per column,
 * E = sum_k rate_k L_k^dagger L_k, intensities[c] = lambda of column c. */
pimi_status pimi_conditional_action(pimi_workspace *ws, const void *X, void *Y, int64_t ncols, double *intensities,
                                    int32_t on_device);
/* Allocation-free variant of pimi_conditional_action, for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops
 * that forbid hipMalloc).  No counterpart in the reference.  X, Y: device
 * pointers (n x ncols complex double, column-major); intensities: device
 * pointer to ncols doubles, receiving max(0, lambda_c) as the host array of
 * pimi_conditional_action does.  The call allocates nothing: the intensity
 * reduction uses a fixed device array of the library (one per GPU, held
 * until the process ends, about 1 MB), so ncols is limited to 1..256
 * (PIMI_EINVAL beyond).  The plan's trajectory tables must already exist
 * (they are built, with allocation, by the first pimi_conditional_action,
 * pimi_conditional_rk4, pimi_channel_intensities or pimi_apply_jumps call
 * on the plan); otherwise PIMI_EINVAL.  The same sign check as
 * pimi_conditional_action gives PIMI_EINVAL.  Calls that overlap on one GPU
 * share the array and race; the call returns after its work completes. */
pimi_status pimi_conditional_action_noalloc(pimi_workspace *ws, const void *X, void *Y, int64_t ncols,
                                            double *intensities);
/* batched_conditional_rk4!: one fixed RK4 step of size h for every column,
 * then each column trace-normalized (X in place). */
pimi_status pimi_conditional_rk4(pimi_workspace *ws, void *X, int64_t ncols, double h, int32_t on_device);
/* batched_channel_intensities!: rates[k + c*count] = intensity of channel k+1 in column c. */
pimi_status pimi_channel_intensities(pimi_workspace *ws, const void *X, int64_t ncols, double *rates,
                                     int32_t on_device);
/* batched_apply_jumps!: column c with channels[c] = k > 0 becomes gain_k(X_c) / (t . gain_k(X_c));
 * channels[c] = 0 leaves it unchanged.  Columns of one channel go through one multi-column apply. */
pimi_status pimi_apply_jumps(pimi_workspace *ws, void *X, int64_t ncols, const int32_t *channels, int32_t on_device);

/* ---- Arnoldi spectrum (src/krylov.jl:914-961, krylov_liouvillian_spectrum) ----
 * Arnoldi factorization L V_k = V_{k+1} H_k started from q0 (host, n complex;
 * normalized here) with two classical Gram-Schmidt passes per step; at most m
 * steps (m <= the workspace krylovdim), stopping early once
 * H[j+1, j] <= breakfactor * max(|L v_j|, DBL_MIN) (the reference's breakdown
 * test, krylov.jl:941-945).  H: (m+1) x m column-major complex host array,
 * zeroed first; *k_out = steps taken.  The basis stays in the workspace for
 * pimi_arnoldi_vectors.  FP64 workspaces only.  PIMI_EINVAL for a zero or
 * non-finite q0. */
pimi_status pimi_arnoldi(pimi_workspace *ws, const double *q0, int32_t m, double breakfactor, double *H,
                         int32_t *k_out);
/* Ritz vectors of the last pimi_arnoldi call: out[:, c] = V[:, 0:k] Y[:, c]
 * for c < ncols (Y: k x ncols column-major complex host; out: n x ncols). */
pimi_status pimi_arnoldi_vectors(pimi_workspace *ws, int32_t k, const double *Y, int32_t ncols, double *out);

/* ---- fixed-step RK4 (src/evolution.jl:75-118, evolve!) ----
 * x_out = RK4 propagation of x0 over [t0, t1] in `steps` equal steps, with the
 * reference's stage arithmetic (k2 accumulates k1 + 2 k2' + 2 k3' + k4',
 * x += (h/6) k2).  Host vectors (n complex); all stages on the device.  FP64
 * workspaces only; PIMI_EINVAL for steps <= 0 or a non-finite tspan. */
pimi_status pimi_evolve_rk4(pimi_workspace *ws, const double *x0, double t0, double t1, int64_t steps,
                            double *x_out);

/* ---------------------------------------------------------------------------
 * Batch engine: many independent stationary solves advanced together on one
 * device, one kernel launch per operation for all of them.  No counterpart in
 * the reference, which solves one model at a time; each job runs the
 * stationary solve of src/krylov.jl:797-891 (bordered system, right sector
 * preconditioner, scale-invariant criterion) as GMRES cycles inside FP64
 * iterative refinement.  Jobs wait in a FIFO queue and are admitted while
 * the engine's device memory reservation holds their state (see
 * pimi_batch_job_bytes); a finished job frees its memory at once.  A job's
 * result does not depend on which other jobs share the device.
 * ------------------------------------------------------------------------- */
typedef struct pimi_batch pimi_batch;

typedef struct {
    int32_t precision;     /* PIMI_C64, or PIMI_C32: FP32 Arnoldi (basis, apply, preconditioner) in FP64 refinement */
    int32_t max_jobs;      /* jobs resident at once (0: 1024) */
    int64_t memory_bytes;  /* device memory reserved for job state (0: three quarters of the free memory) */
    int32_t flags;         /* PIMI_BATCH_TIMING: time every launch with HIP events (see pimi_batch_stats) */
    int32_t reserved;
} pimi_batch_opts;

enum { PIMI_BATCH_TIMING = 1 };
/* Operation classes of pimi_batch_stats. */
enum {
    PIMI_OP_PRECOND_GEMM = 0, /* sector preconditioner (MFMA GEMMs) */
    PIMI_OP_APPLY = 1,        /* stencil apply of the Arnoldi step */
    PIMI_OP_DOTS = 2,         /* Gram-Schmidt projections V^H w */
    PIMI_OP_UPDATE = 3,       /* Gram-Schmidt updates w -= V h */
    PIMI_OP_CYCLE_END = 4,    /* x += M^{-1} V y */
    PIMI_OP_FP64 = 5,         /* FP64 residuals, operator-scale probes, final checks */
    PIMI_OP_OTHER = 6         /* norms, scalings, trace terms, mailbox */
};

pimi_status pimi_batch_create(pimi_context *ctx, const pimi_batch_opts *opts, pimi_batch **out);
void pimi_batch_destroy(pimi_batch *batch);
/* device bytes one job of this plan holds while admitted */
pimi_status pimi_batch_job_bytes(const pimi_plan *plan, int32_t precision, int32_t krylovdim, int64_t *bytes);
/* Queue a stationary solve of `plan` (which must outlive the job).  opts:
 * preconditioner PIMI_PRECOND_SECTOR, criterion PIMI_CRITERION_BACKWARD,
 * krylovdim = Krylov dimension of one GMRES cycle, inner_rtol in (0, 1) =
 * residual reduction per cycle, maxiter = Arnoldi steps in total,
 * max_refinements = cycles, operator_scale 0 = estimate it. */
pimi_status pimi_batch_submit(pimi_batch *batch, pimi_plan *plan, const pimi_gmres_opts *opts, int64_t *job_id);
/* Advance up to max_ticks ticks (admitting queued jobs as memory frees); with
 * stop_when_finished, return after the first tick in which a job finished.
 * *finished = jobs that finished during the call. */
pimi_status pimi_batch_run(pimi_batch *batch, int32_t max_ticks, int32_t stop_when_finished, int64_t *finished);
/* counts[6]: queued, active, finished (not yet fetched), ticks so far,
 * reserved bytes in use, reservation size */
pimi_status pimi_batch_counts(pimi_batch *batch, int64_t *counts);
/* counts[3]: ticks launched as HIP graphs, graphs instantiated, graphs
 * updated in place (a tick of a shape seen before).  The engine captures
 * each tick's launches as a HIP graph unless PIMI_NO_GRAPH=1 or
 * PIMI_BATCH_TIMING.  No counterpart in the reference. */
pimi_status pimi_batch_graph_counts(pimi_batch *batch, int64_t *counts);
/* Accumulated statistics per operation class c (PIMI_OP_*): out[c] = device
 * milliseconds (only with PIMI_BATCH_TIMING), out[8+c] = algorithmic bytes
 * (each array counted once per launch), out[16+c] = flops; 24 entries.
 * reset != 0 clears them afterwards.  Measurement only. */
pimi_status pimi_batch_stats(pimi_batch *batch, double *out, int32_t reset);
/* next finished job id, or -1 */
pimi_status pimi_batch_next_finished(pimi_batch *batch, int64_t *job_id);
/* Copy a finished job's state (trace one, n complex doubles) and info, and
 * forget the job; PIMI_ENOTCONVERGED (state and info filled) for a job that
 * did not converge. */
pimi_status pimi_batch_fetch(pimi_batch *batch, int64_t job_id, double *rho_out, pimi_solve_info *info);
/* A dynamics job in the batch engine: the adaptive Krylov exponential of
 * pimi_expv_series (solve_dynamics with ExpvAlgorithm, src/highlevel.jl:559-733;
 * krylov_expv!, src/krylov_extensions.jl:1452-1568) from rho0 (n complex)
 * over the output times ts[0..nts-1] (nondecreasing), observables
 * <obs_q, rho(t)> of nobs vectors (nobs x n complex, may be 0).  Its Arnoldi
 * steps, projections and observations run in the engine's ticks beside the
 * other jobs (one launch per step kind over all jobs); the step control and
 * the small exponential are the single-vector path's.  FP64 engines only.
 * Results agree with pimi_expv_series to the exponential's tolerance (the
 * Arnoldi basis is orthogonalized by delayed reorthogonalization, not CGS2). */
pimi_status pimi_batch_submit_expv(pimi_batch *batch, pimi_plan *plan, const double *rho0, const double *ts,
                                   int64_t nts, const pimi_expv_opts *opts, int32_t nobs, const double *obs,
                                   int64_t *job_id);
/* A finished dynamics job's observables (nts x nobs complex, row per output
 * time; may be NULL when nobs = 0), its final state (n complex; may be NULL)
 * and info (accepted, rejected, applications, factorizations, iterations =
 * attempts); forgets the job.  PIMI_ENOTCONVERGED (outputs filled up to the
 * failure) when the exponential did not reach the last output time. */
pimi_status pimi_batch_fetch_expv(pimi_batch *batch, int64_t job_id, double *obs_out, double *rho_final,
                                  pimi_solve_info *info);
/* Multi-model apply: y_k = L_k x_k for every plan k in one launch (banded
 * and ELL kernels over all plans' tiles), in the engine precision.  x, y:
 * host buffers holding the plans' vectors one after another (complex
 * double, sum of the plans' n entries); they must not overlap.  The batch
 * apply! of src/liouvillian.jl:3232-3251 runs many columns of one plan; this
 * runs one column of many plans. */
pimi_status pimi_batch_apply(pimi_batch *batch, pimi_plan *const *plans, int32_t nplans, const double *x, double *y);
/* Throughput probe: one launch applying L to a synthetic vector of every plan
 * (engine precision), `reps` times, event-timed; algorithmic bytes per launch
 * = sum over plans of 2 n s + stencil tables.  Measurement only. */
pimi_status pimi_batch_apply_bench(pimi_batch *batch, pimi_plan *const *plans, int32_t nplans, int32_t reps,
                                   double *ms_avg, int64_t *algorithmic_bytes);
/* Autotuner of the engine's banded-apply tile width, per plan size: times one
 * launch over a synthetic vector of every plan (as pimi_batch_apply_bench) at
 * each width (4, 8, 16 columns), `reps` times each, and keeps the fastest for
 * plans of sizes it has not tuned (*width_out = that width); then the same
 * over the plans of each distinct size (PI dimension and sector count, i.e.
 * each N of one local dimension) on their own, and keeps that size's
 * fastest for the engine's later launches, which group the jobs of one tick
 * by width (one launch per width).  widths_out (nplans entries, or null) =
 * the width now used for each plan (0: not banded).  Results do not depend
 * on the widths.  No counterpart in the reference. */
pimi_status pimi_batch_autotune_apply(pimi_batch *batch, pimi_plan *const *plans, int32_t nplans, int32_t reps,
                                      int32_t *width_out, int32_t *widths_out);

#ifdef __cplusplus
}
#endif
#endif
