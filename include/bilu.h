/*
 * bilu.h -- C ABI of the MI355X batched shared-pattern ILU(0)/ParILU(0)/ISAI
 * preconditioned BiCGSTAB solver (libbilu.so).
 *
 * Plain C: pointers, sizes and an opaque pattern handle.  Device pointers are
 * HIP device addresses owned by the caller; `stream` is a hipStream_t passed
 * as void* (NULL = the default stream).  Every entry point returns a status
 * code (BILU_OK = 0) and never aborts the process.
 *
 * Reference interfaces these replace (aggarwal2000/Merging-ILU0,
 * Batched-ILU0-merged/ILU0_ISAI_bicgstab/):
 *   bilu_solve            <- Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_merged_Gpu
 *                            (PrecondBiCGSTAB.h:26-27, merged.cu:610-650), _segregated_
 *                            (PrecondBiCGSTAB.h:29-30), Batched_ILU_Preconditioned_BiCGSTAB_Gpu
 *                            (PrecondBiCGSTAB.h:15-17), Batched_ILU_ISAI_... (:19),
 *                            Batched_General_ISAI_... (:21-23)
 *   bilu_ilu0             <- ComputeILU0Approach1 (ILU_0.h:11) / ILU_0_Factorization_Gpu (ILU_0.h:9)
 *   bilu_parilu0          <- ParILU_0_Factorization_Gpu (parILU_0.h:7)
 *   bilu_split_lu         <- Update_row_pointers_L_and_U_and_Allocate_Memory +
 *                            Fill_L_and_U_col_idxs_and_vals (factorization.h:23-26)
 *   bilu_isai             <- GenerateISAI_gpu (isai.h:9), power 1
 *   bilu_pattern_create   <- Count_Missing_Diagonal_Elements + Find_locations_of_diagonal_elements
 *                            (factorization.h:10,17), done once per pattern on the host
 *   bilu_true_residual    <- KernelFillTrueResNorms (Bicgstab_common.hpp:294-342)
 *   bilu_pattern_add_diagonal, bilu_expand_values
 *                         <- Count_Missing_Diagonal_Elements + Add_Missing_Diagonal_Elements
 *                            (factorization.h:10-15, factorization.cu:399-467)
 */
#ifndef BILU_H
#define BILU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define BILU_OK 0
#define BILU_ERR_INVALID_ARG 1
#define BILU_ERR_INVALID_PATTERN 2   /* unsorted / out-of-range columns, bad row_ptrs */
#define BILU_ERR_MISSING_DIAGONAL 3  /* the ILU paths need a full diagonal (merged.cu:557-560) */
#define BILU_ERR_TOO_LARGE 4         /* n, nnz or ISAI row size beyond what the kernels support */
#define BILU_ERR_HIP 5               /* a HIP runtime call failed */
#define BILU_ERR_NO_DEVICE 6

/* ---- per-system convergence flags (conv output); 1/-1 keep the reference meaning ---- */
#define BILU_CONVERGED 1
#define BILU_NOT_CONVERGED (-1)   /* max_iter reached or rho == 0 breakdown */
#define BILU_ZERO_PIVOT (-2)      /* zero pivot in ILU(0) / zero diagonal in ISAI */
#define BILU_NONFINITE (-3)       /* NaN/Inf reached the residual */

/* ---- preconditioners (PrecondBiCGSTAB.h:15-30) ---- */
#define BILU_PRECOND_ILU0 0           /* exact ILU(0), in place */
#define BILU_PRECOND_PARILU0 1        /* ParILU(0) Jacobi sweeps from L=tril(A)+I, U=triu(A) */
#define BILU_PRECOND_ILU0_ISAI 2      /* exact ILU(0), then ISAI(L), ISAI(U) on the factor pattern */
#define BILU_PRECOND_PARILU0_ISAI 3   /* ParILU(0), then ISAI(L), ISAI(U) */
#define BILU_PRECOND_GENERAL_ISAI 4   /* ISAI(A), Gauss-Jordan per row */
#define BILU_PRECOND_NONE 5           /* identity (testing) */
#define BILU_PRECOND_FACTORS 6        /* in-place ILU factors supplied by the caller in bilu_solve_ex's
                                         d_F_init (the factor-then-solve flow of ilu_solve.cu:443-509) */

#define BILU_TOL_ABSOLUTE 0   /* ||r|| < tol (reference, ATOL = 1e-11) */
#define BILU_TOL_RELATIVE 1   /* ||r|| < tol * ||b|| */

#define BILU_MERGED 0         /* factor and solve in one launch (merged.cu) */
#define BILU_SEGREGATED 1     /* factor launch, factors through HBM, solve launch (segregated.cu) */
#define BILU_MERGE_AUTO 2     /* FP64: merged (ParILU(0), and ILU(0) solved on the one-wave kernel:
                                 segregated); FP32_IR: segregated (the faster at
                                 n = 144 when measured); bilu_uneven_tune picks per pattern.  Results do
                                 not depend on it */

#define BILU_KERNEL_AUTO 0        /* fastest kernel that supports the pattern and options */
#define BILU_KERNEL_WAVE 1        /* one wave64 workgroup per system, schedules read from global */
#define BILU_KERNEL_PERSISTENT 2  /* persistent multi-wave workgroups, schedules in LDS */

#define BILU_PRECISION_FP64 0     /* everything in FP64 (the reference arithmetic) */
#define BILU_PRECISION_FP32_IR 1  /* FP32 preconditioner and inner BiCGSTAB inside FP64
                                     iterative refinement (outer residual and its dot in FP64);
                                     ILU(0) / ParILU(0) factors, their ISAI(L, U), or the general
                                     ISAI of A (generated in FP64, stored FP32, segregated only) */

/* exact ILU(0) schedules (the same arithmetic and bits, different parallel orders) */
#define BILU_ILU_SCHED_AUTO 0     /* the pattern's tuned schedule (bilu_ilu0_tune), else ROWS */
#define BILU_ILU_SCHED_ROWS 1     /* rows level by level, a lane per row (approach 1 order) */
#define BILU_ILU_SCHED_ENTRIES 2  /* entries level by level from dependency lists (approach 3) */
#define BILU_ILU_SCHED_ERAS 3     /* right-looking pivot eras (approach 4) */
#define BILU_ILU_SCHED_STREAMS 4  /* persistent kernel, list-scheduled lane streams */

/* ParILU(0) sweep order (bilu_opts.parilu_order) */
#define BILU_PARILU_JACOBI 0   /* every entry reads the iterate the sweep started from */
#define BILU_PARILU_INPLACE 1  /* rounds of 64 entries (fewest products first), each reading the
                                  iterate as the earlier rounds of the sweep left it; the
                                  persistent factor kernel only (FP32_IR, or FP64 ILU-type
                                  preconditioners through the split factor + solve launches) */

#define BILU_ISAI_GENERAL 0
#define BILU_ISAI_LOWER 1     /* strictly-lower values + unit diagonal */
#define BILU_ISAI_UPPER 2     /* diagonal + strictly-upper values */

typedef struct bilu_pattern_s *bilu_pattern_t;

typedef struct {
    int precond;          /* BILU_PRECOND_* */
    int parilu_sweeps;    /* sweeps for the ParILU variants (>= 0) */
    double tol;           /* default 1e-11 */
    int tol_mode;         /* BILU_TOL_* */
    int max_iter;         /* full BiCGSTAB iterations, default 500 */
    int merge;            /* BILU_MERGED / BILU_SEGREGATED / BILU_MERGE_AUTO (default).  FP32_IR:
                             merged = one persistent kernel factors (and generates the ISAI
                             of) each system in LDS and solves it; segregated = the factor
                             launch writes the FP32 factors / ISAI values to the workspace
                             (the general ISAI: segregated only, merged gives INVALID_ARG) */
    int true_res_restart; /* 1: restart from b - A x when the true residual misses tol */
    int max_restarts;     /* bound on such restarts (default 3) */
    int kernel;           /* BILU_KERNEL_*: which kernel family runs the solve */
    int precision;        /* BILU_PRECISION_* (default FP64) */
    double inner_tol;     /* FP32_IR: relative tolerance of each inner solve (default 1e-5) */
    int max_refine;       /* FP32_IR: refinement steps (default 10) */
    double parilu_stop_tol; /* ParILU: > 0 stops the sweeps once ||(LU - A) on the pattern||_F
                               <= parilu_stop_tol * ||A||_F (parilu_sweeps is then the cap)
                               and falls back to exact ILU(0) when the cap is reached
                               without it; 0 = fixed sweep count (default) */
    int isai_power;       /* ISAI preconditioners: sparsity of |L|^k, |U|^k (|A|^k for the
                             general ISAI); 1 = the factors' / matrix's own pattern (default).
                             The reference's ILU-ISAI solver uses 2 (PrecondBiCGSTAB_ilu_isai_gpu.cu:
                             297-299).  k > 1: FP64, one wave per system, no warm start; rows of
                             the general ISAI at most 32 entries (the reference's row_size_limit),
                             of ISAI(L) / ISAI(U) at most 128 (beyond the reference's assert) */
    double inner_adapt;   /* FP32_IR: an inner solve stops at the relative residual
                             max(inner_tol, inner_adapt * tol / |r|), |r| the FP64 residual it
                             starts from, so the last refinement step solves no further than the
                             outer tolerance needs (default 0.5; 0 = inner_tol always) */
    int parilu_order;     /* BILU_PARILU_JACOBI (default) / BILU_PARILU_INPLACE.  With
                             parilu_stop_tol > 0, an in-place sweep stops on the residuals its
                             rounds saw and keeps the iterate after the sweep */
} bilu_opts;

typedef struct {
    int n, nnz;
    int l_nnz, u_nnz;          /* reference L/U layout sizes (L stores its unit diagonal) */
    int levels_lower, levels_upper;
    int ilu_updates;           /* number of (i,k,j) ILU(0) updates */
    int max_row_nnz;
    int levels_entries;        /* levels of the entry (dependency-list) ILU(0) schedule */
    int ilu_eras;              /* eras of the right-looking pivot ILU(0) schedule */
} bilu_pattern_info;

/* defaults: ILU(0), 5 sweeps, tol 1e-11 absolute, 500 iterations, merge auto, no restart,
 * FP64, inner_tol 1e-5, 10 refinement steps, fixed ParILU sweeps */
void bilu_opts_default(bilu_opts *o);
const char *bilu_version(void);
const char *bilu_status_string(int status);

/* Host arrays; columns sorted within each row.  Analyses the pattern once
 * (diagonal pointers, level schedules, ILU update lists, ParILU dependency
 * lists, ISAI gather maps) and uploads it to the current device.  A handle's
 * launches share its work counters and its workspace: calls on one handle
 * must be ordered (one stream, or the caller's events); two solves on one
 * handle running concurrently on different streams would take systems from
 * each other's counter.  Different handles are independent (an uneven plan
 * takes each of its patterns' handles for its own branches). */
int bilu_pattern_create(int n, int nnz, const int *row_ptrs, const int *col_idx,
                        bilu_pattern_t *out);
int bilu_pattern_destroy(bilu_pattern_t p);
int bilu_pattern_get_info(bilu_pattern_t p, bilu_pattern_info *info);
/* Sizes the handle's solve workspace for batches of up to `batch` systems on
 * every solve path (the split factor + solve launches keep the factors or
 * ISAI values there), so that no later bilu_solve / bilu_solve_ex /
 * bilu_mlp_solve of at most that many systems allocates or waits for the
 * device: the calls a caller times or captures into a hipGraph.  Without it
 * the workspace grows on first use (a device synchronize and hipMalloc),
 * and a solve under graph capture that would grow it returns
 * BILU_ERR_INVALID_ARG instead.  No counterpart in the reference, which
 * allocates and frees its scratch inside every solve (merged.cu:553,565). */
int bilu_pattern_reserve(bilu_pattern_t p, int batch);
/* Missing-diagonal insertion (factorization.cu:399-467): the pattern with an
 * explicit entry (i,i) inserted at its sorted place in every row that lacks
 * one.  Host arrays, columns sorted within each row (checked as in
 * bilu_pattern_create).  *n_missing receives the number of inserted entries;
 * with row_ptrs_out == NULL only counts.  Otherwise row_ptrs_out[n+1],
 * col_idx_out[nnz + n_missing] and src_out[nnz + n_missing] (position of each
 * new entry in the input pattern, -1 for an inserted diagonal) are written. */
int bilu_pattern_add_diagonal(int n, int nnz, const int *row_ptrs, const int *col_idx,
                              int *n_missing, int *row_ptrs_out, int *col_idx_out, int *src_out);
/* Values of a batch on the augmented pattern (update_col_idxs_and_values_diagonal_add,
 * factorization.cu:169-189): d_out[s*nnz_out + k] = d_src[k] < 0 ? 0 :
 * d_vals[s*nnz + d_src[k]].  d_src[nnz_out] is the device copy of src_out
 * above (entries < nnz). */
int bilu_expand_values(int batch, int nnz, int nnz_out, const int *d_src, const double *d_vals,
                       double *d_out, void *stream);
/* host copies of the reference L/U split patterns (sizes from get_info) */
int bilu_pattern_lu_split(bilu_pattern_t p, int *l_row_ptrs, int *l_col_idx,
                          int *u_row_ptrs, int *u_col_idx);

/* Batched solve A_s x_s = b_s, s < batch.  d_A[batch*nnz], d_b/d_x[batch*n],
 * per-system outputs d_iter (half steps, float), d_conv, d_iter_res (last
 * recursive residual norm), d_true_res (||b - A x||_2); d_iter_res and
 * d_true_res may be NULL. */
int bilu_solve(bilu_pattern_t p, int batch, const double *d_A, const double *d_b,
               double *d_x, const bilu_opts *opts, float *d_iter, int *d_conv,
               double *d_iter_res, double *d_true_res, void *stream);

/* bilu_solve with the extensions of the mixed-precision and warm-started
 * paths: d_F_init[batch*nnz] (may be NULL) is the initial guess of the
 * ParILU(0) sweeps (in-place factor layout, FP64); d_refine[batch] (may be
 * NULL) receives the FP64 refinement steps of BILU_PRECISION_FP32_IR (0 for
 * FP64); d_factor_status[batch] (may be NULL) receives the ParILU sweeps
 * used, -1 when the residual stop fell back to exact ILU(0), 0 for exact
 * ILU(0).  With BILU_PRECISION_FP32_IR, d_iter counts the inner FP32 half
 * steps summed over the refinement steps and d_iter_res = d_true_res = the
 * last FP64 residual norm.  No counterpart in the reference (its solvers are
 * FP64 only and start ParILU from A, parILU_0_gpu.cu:293-358). */
int bilu_solve_ex(bilu_pattern_t p, int batch, const double *d_A, const double *d_b,
                  double *d_x, const bilu_opts *opts, const double *d_F_init, float *d_iter,
                  int *d_conv, double *d_iter_res, double *d_true_res, int *d_refine,
                  int *d_factor_status, void *stream);

/* bilu_solve_ex for callers that cannot allocate (code under hipGraph
 * capture, serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  Same arguments, options and results, bit for bit; no
 * hipMalloc, hipFree, device synchronize or allocating library call inside.
 * The factors / ISAI values between the factor and solve launches live in a
 * fixed 256 MiB device array of the library (one per GPU, held until the
 * process ends) instead of the handle's workspace; batches larger than it
 * holds (about 16,000 systems at nnz = 2082) run in consecutive chunks on
 * `stream`.  Every call on a GPU shares that array: calls that overlap on one
 * GPU (other streams, other threads, other handles) race and must not be
 * issued.  Nothing in the array needs to start at zero (each system's values
 * are written before they are read).  BILU_ERR_TOO_LARGE when one system
 * does not fit; BILU_ERR_INVALID_ARG for isai_power > 1 (those pattern maps
 * are allocated at their first use). */
int bilu_solve_ex_noalloc(bilu_pattern_t p, int batch, const double *d_A, const double *d_b,
                          double *d_x, const bilu_opts *opts, const double *d_F_init, float *d_iter,
                          int *d_conv, double *d_iter_res, double *d_true_res, int *d_refine,
                          int *d_factor_status, void *stream);

/* Kernel timing for benchmarks (no counterpart in the reference): between
 * bilu_timing_begin and bilu_timing_end, the first `max_launches` launches
 * of one kernel kind by any solve entry point are bracketed by HIP events
 * recorded on the stream the kernel is launched on (events created by
 * _begin).  _end waits for those events and returns the average kernel
 * duration (ms) and the number of launches timed.  One window at a time. */
#define BILU_TIMING_FACTOR 1   /* the persistent factorization kernel of the split paths */
#define BILU_TIMING_SOLVE 2    /* the solve kernel (persistent FP64, mixed precision, one wave per system) */
int bilu_timing_begin(int kernel, int max_launches);
int bilu_timing_end(float *avg_ms, int *launches);

/* Exact ILU(0) in place on the pattern: d_F[batch*nnz] <- factors of d_A
 * (strict lower = L, unit diagonal implicit; upper = U).  d_status[batch]
 * (may be NULL) receives 0 or BILU_ZERO_PIVOT. */
int bilu_ilu0(bilu_pattern_t p, int batch, const double *d_A, double *d_F,
              int *d_status, void *stream);
/* bilu_ilu0 with a chosen schedule (BILU_ILU_SCHED_*): ILU_0_gpu.cu's
 * approach switch (:1241-1343; approaches 1/3/4 at :105-186, :693-935,
 * :947-1229), all of them deterministic here.  BILU_ERR_TOO_LARGE when the
 * pattern does not fit the schedule's kernel. */
int bilu_ilu0_ex(bilu_pattern_t p, int batch, const double *d_A, double *d_F, int schedule, int *d_status,
                 void *stream);
/* times every schedule the pattern supports on this batch (one warm-up and
 * two timed runs each, the stream is synchronized), keeps the fastest for
 * BILU_ILU_SCHED_AUTO and returns it in *best (may be NULL); ms[4] (may be
 * NULL) gets the time per run of ROWS, ENTRIES, ERAS, STREAMS (-1: not
 * supported).  d_F holds the factors of d_A on return. */
int bilu_ilu0_tune(bilu_pattern_t p, int batch, const double *d_A, double *d_F, float *ms, int *best,
                   void *stream);
/* ParILU(0): `sweeps` Jacobi sweeps from d_F_init (NULL: from A) */
int bilu_parilu0(bilu_pattern_t p, int batch, const double *d_A, const double *d_F_init,
                 double *d_F, int sweeps, void *stream);
/* ||(L U - A) on the pattern||_F per system */
int bilu_factor_residual(bilu_pattern_t p, int batch, const double *d_A, const double *d_F,
                         double *d_res, void *stream);
/* in-place factor -> reference L (unit diagonal stored) and U value arrays */
int bilu_split_lu(bilu_pattern_t p, int batch, const double *d_F, double *d_L, double *d_U,
                  void *stream);
/* ISAI (power 1) of the typed view of d_vals; output on the pattern masked
 * to the type's triangle (other positions are left untouched). */
int bilu_isai(bilu_pattern_t p, int batch, const double *d_vals, int type, double *d_out,
              int *d_status, void *stream);
/* ISAI with the sparsity of |P|^power, P = the typed view of the pattern
 * (LOWER: columns < i plus the unit diagonal, UPPER: columns >= i, GENERAL:
 * all) -- the reference's chosen-pattern ISAI (GenerateISAI_gpu(aiA, A, type,
 * power), isai.h:9, pattern isai_gpu.cu:40-178, taken structurally).
 * bilu_isai_pattern: host copy of that pattern (row_ptrs == NULL: *nnz only).
 * bilu_isai_power: values d_out[batch * nnz] on it; BILU_ERR_TOO_LARGE for a
 * GENERAL row of more than 32 entries (the reference's row_size_limit) or a
 * LOWER / UPPER row of more than 128 (the reference asserts beyond 32). */
int bilu_isai_pattern(bilu_pattern_t p, int type, int power, int *nnz, int *row_ptrs, int *col_idx);
int bilu_isai_power(bilu_pattern_t p, int batch, const double *d_vals, int type, int power,
                    double *d_out, int *d_status, void *stream);
/* ||b - A x||_2 per system */
int bilu_true_residual(bilu_pattern_t p, int batch, const double *d_A, const double *d_b,
                       const double *d_x, double *d_res, void *stream);

/* ---- uneven batches (no counterpart in the reference: its batches share one
 * pattern).  A request list over several patterns, solved by the
 * mixed-precision path (opts: precision BILU_PRECISION_FP32_IR, precond
 * ILU0 / PARILU0 / ILU0_ISAI / PARILU0_ISAI).  Request r solves system
 * req_index[r] of pattern req_pattern[r] (arrays d_A[p] [*, nnz_p], d_b[p],
 * d_x[p] per pattern); per-request results d_iter[r], d_conv[r],
 * d_iter_res[r], d_true_res[r], d_refine[r] (last three may be NULL).  One
 * persistent factor + solve kernel chain per pattern, as concurrent branches
 * of one hipGraph, captured by the first solve and replayed while the
 * arguments are the same.  Patterns must be distinct handles; a plan is used
 * from one host thread at a time. */
typedef struct bilu_uneven_s *bilu_uneven_t;

int bilu_uneven_create(int npat, const bilu_pattern_t *patterns, int nreq, const int *req_pattern,
                       const int *req_index, const bilu_opts *opts, bilu_uneven_t *out);
int bilu_uneven_destroy(bilu_uneven_t u);
/* autotuner: times every pattern's candidates -- the merge point (merged:
 * one kernel factors each system in LDS and solves it; segregated: factor
 * kernel, FP32 factors / ISAI values through HBM, solve kernel; opts.merge
 * BILU_MERGED / BILU_SEGREGATED fixes it, BILU_MERGE_AUTO lets the tuner
 * choose) and the workgroup sizes (waves per factor workgroup {2, 4, 8}, per
 * solve workgroup {4, 8, 12}, for either kernel; for the ISAI templates of 48 and 72 slots per
 * lane also the placement of the solve kernel's slot records: the workgroup's LDS or global
 * memory) -- first alone on the pattern's requests, then
 * pattern by pattern (slowest first) inside the whole launch with the other
 * patterns running, and keeps the fastest (results do not depend on the
 * choice) */
int bilu_uneven_tune(bilu_uneven_t u, const double *const *d_A, const double *const *d_b, double *const *d_x,
                     float *d_iter, int *d_conv, double *d_iter_res, double *d_true_res, int *d_refine,
                     void *stream);
/* chosen configuration of pattern p and its number of requests (factor_waves
 * 0: the merged kernel) */
int bilu_uneven_config(bilu_uneven_t u, int p, int *factor_waves, int *solve_waves, int *jobs);
/* 1 when pattern p runs the merged kernel */
int bilu_uneven_merge_point(bilu_uneven_t u, int p, int *merged);
/* pins pattern p's configuration instead of the tuner's choice (merged 0/1,
 * a merge point the plan's opts.merge allows; waves per workgroup, capped by
 * what the kernel and the LDS hold; records_global 1 takes effect for the ISAI
 * templates of 48 and 72 slots).  Results do not depend on it. */
int bilu_uneven_configure(bilu_uneven_t u, int p, int merged, int factor_waves, int solve_waves, int records_global);
/* 1 when pattern p's solve kernel reads its slot records from global memory */
int bilu_uneven_placement(bilu_uneven_t u, int p, int *records_global);
/* the candidates the last bilu_uneven_tune timed for pattern p (up to cap of
 * them; *count = how many): merge point, waves per factor / solve workgroup,
 * record placement, ms alone and inside the whole launch (-1: not timed there) */
int bilu_uneven_tune_table(bilu_uneven_t u, int p, int cap, int *merged, int *factor_waves, int *solve_waves,
                           int *records_global, float *alone_ms, float *corun_ms, int *count);
int bilu_uneven_solve(bilu_uneven_t u, const double *const *d_A, const double *const *d_b, double *const *d_x,
                      float *d_iter, int *d_conv, double *d_iter_res, double *d_true_res, int *d_refine,
                      void *stream);

/* ---- learned warm start of ParILU(0) (no counterpart in the reference) ----
 * An MLP (features asinh(a/1e-8)/24 of the nnz values of A, two hidden
 * layers of `hidden` GELU units, nnz outputs) predicts the in-place ILU(0)
 * factor; its output feeds bilu_solve_ex's d_F_init.  BF16 MFMA GEMMs with
 * FP32 accumulation; FP32 master weights trained with Adam on exact ILU(0)
 * factors (e.g. from bilu_ilu0).  Weights are exchanged in the packed FP32
 * layout W1[hidden][nnz], b1[hidden], W2[hidden][hidden], b2[hidden],
 * W3[nnz][hidden], b3[nnz]. */
typedef struct bilu_mlp_s *bilu_mlp_t;

int bilu_mlp_create(bilu_pattern_t p, int hidden, unsigned long long seed, bilu_mlp_t *out);
int bilu_mlp_destroy(bilu_mlp_t m);
int bilu_mlp_num_params(bilu_mlp_t m, long long *count);
int bilu_mlp_get_weights(bilu_mlp_t m, float *d_w, void *stream);
int bilu_mlp_set_weights(bilu_mlp_t m, const float *d_w, void *stream);
/* mean squared error of the scaled factors over the minibatch (d_loss, one
 * device float, may be NULL) and its gradient (d_grad[num_params], packed
 * layout, may be NULL), without updating the weights */
int bilu_mlp_gradient(bilu_mlp_t m, int batch, const double *d_A, const double *d_F, float *d_loss,
                      float *d_grad, void *stream);
/* one Adam step (lr) on the minibatch d_A[batch*nnz] with exact factors d_F */
int bilu_mlp_train_step(bilu_mlp_t m, int batch, const double *d_A, const double *d_F, float lr,
                        float *d_loss, void *stream);
/* predicted factors d_F_init[batch*nnz] (FP64, in-place factor layout); the
 * network's output y is rounded to FP16 before it enters the guess (as in
 * bilu_mlp_solve's fused warm start) */
int bilu_mlp_infer(bilu_mlp_t m, int batch, const double *d_A, double *d_F_init, void *stream);
/* learned + mixed-precision solve in one call: the same results, bit for bit,
 * as bilu_mlp_infer followed by bilu_solve_ex with that d_F_init and the
 * FP32-IR ParILU(0) or ParILU(0)+ISAI options, without the FP64 d_F_init
 * round trip (the factor kernel turns the network's raw outputs into the warm
 * start itself).  m must have been created on the handle p; opts->precond is
 * BILU_PRECOND_PARILU0 or _PARILU0_ISAI, opts->precision BILU_PRECISION_FP32_IR,
 * isai_power 1.  Outputs as bilu_solve_ex (d_factor_status per system). */
int bilu_mlp_solve(bilu_mlp_t m, bilu_pattern_t p, int batch, const double *d_A, const double *d_b, double *d_x,
                   const bilu_opts *opts, float *d_iter, int *d_conv, double *d_iter_res, double *d_true_res,
                   int *d_refine, int *d_factor_status, void *stream);

#ifdef __cplusplus
}
#endif
#endif
