/*
 * oracle.h -- CPU restatement of the reference's batched shared-pattern
 * ILU(0) / ParILU(0) / ISAI-preconditioned BiCGSTAB arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or the timed CPU baseline).  The product (libbilu) never links,
 * loads or calls it.
 *
 * Everything works on one CSR pattern (row_ptrs rp[n+1], sorted column
 * indices ci[nnz]) shared by all systems of a batch; values are page-major
 * (system s owns vals[s*nnz .. s*nnz+nnz)).  Factors are kept "in place" on
 * that pattern: strictly-lower positions hold L (unit diagonal implicit),
 * diagonal and strictly-upper positions hold U -- the layout the reference
 * merged kernel factors into (PrecondBiCGSTAB_exact_ilu0_app1_merged_gpu.cu
 * :145-246).  ISAI factors are stored on the same pattern, masked to their
 * triangle (lower: rp[i]..diag[i], upper: diag[i]..rp[i+1]-1).
 *
 * Floating-point contract: this file is compiled with -ffp-contract=off and
 * every fused multiply-add is an explicit fma(), written where the CUDA
 * reference's `a += b*c` / `a -= b*c` is contracted by nvcc.  Reduction
 * orders are parameterised (width w, see orc_dot) so one routine covers the
 * reference order (w = 32: warp-strided partial sums + shfl_down tree,
 * Bicgstab_common.hpp:97-128) and the wave64 order of the HIP kernels
 * (w = 64).
 */
#ifndef BILU_ORACLE_H
#define BILU_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* per-system status codes (conv flag); 1 / -1 keep the reference meaning
 * (PrecondBiCGSTAB_exact_ilu0_app1_merged_gpu.cu:337,445,500) */
#define ORC_CONVERGED 1
#define ORC_NOT_CONVERGED (-1)
#define ORC_ZERO_PIVOT (-2)
#define ORC_NONFINITE (-3)

/* preconditioner kinds for orc_bicgstab */
#define ORC_PC_NONE 0       /* identity */
#define ORC_PC_LU 1         /* in-place L\U factor: lower then upper trisolve */
#define ORC_PC_ISAI_LU 2    /* aiU * (aiL * v), two SpMVs (ilu_isai_solve.cu:25-45) */
#define ORC_PC_ISAI_GEN 3   /* aiA * v (gisai_solve.cu:23-30) */
#define ORC_PC_ISAI_LU_PAT 4   /* aiU * (aiL * v) with aiL, aiU on their own patterns */
#define ORC_PC_ISAI_GEN_PAT 5  /* aiA * v with aiA on its own pattern */

/* ISAI matrix types (isai.h:7 enum class mat_type) */
#define ORC_ISAI_GENERAL 0
#define ORC_ISAI_LOWER 1
#define ORC_ISAI_UPPER 2

#define ORC_TOL_ABSOLUTE 0  /* reference: ||r|| < ATOL (merged.cu:438,496) */
#define ORC_TOL_RELATIVE 1  /* ||r|| < tol * ||b|| */

typedef struct {
    double tol;
    int tol_mode;
    int max_iter;            /* full iterations (reference MAX_ITER = 500) */
    int dot_width;           /* 1 = serial, 32 = reference warp, 64 = wave64 */
    int spmv_width;          /* 1 = sequential row sum, 32 = reference warp tree */
    int true_res_restart;    /* restart from b - A x when the true residual misses tol */
    int max_restarts;
    int trsv_order;          /* 0: CSR order (reference), 1: availability order (see orc_lu_solve_avail) */
    double inner_adapt;      /* FP32 inner solves: stop at max(inner_tol, inner_adapt * tol / |r|) */
} orc_opts;

int orc_find_diag(int n, const int *rp, const int *ci, int *diag);
int orc_sweep_order(int n, const int *rp, const int *ci, const int *diag, int *order);
int orc_isai_pattern(int n, const int *rp, const int *ci, int type, int power, int *mrp, int *mci);
int orc_isai_on(int n, const int *rp, const int *ci, const double *vals, int type,
                const int *mrp, const int *mci, double *out);
int orc_solve_isai_pow(int n, const int *rp, const int *ci, const double *A, const double *b, double *x,
                       int precond, int sweeps, int power, const orc_opts *o, float *iter, int *conv,
                       double *iter_res, double *true_res);
int orc_add_diagonal(int n, const int *rp, const int *ci, int *rp_out, int *ci_out, int *holder);
void orc_expand_values(int batch, int nnz, int nnz_out, const int *holder, const double *vals, double *out);

double orc_dot(int n, const double *a, const double *b, int width);
void orc_spmv(int n, const int *rp, const int *ci, const double *vals,
              const double *x, double *y, int width);

int orc_ilu0(int n, const int *rp, const int *ci, const int *diag, double *F);
int orc_parilu0_sweeps(int n, const int *rp, const int *ci, const int *diag,
                       const double *A, double *F, int sweeps);
double orc_factor_residual(int n, const int *rp, const int *ci, const int *diag,
                           const double *A, const double *F);
void orc_lu_solve(int n, const int *rp, const int *ci, const int *diag,
                  const double *F, const double *v, double *tmp, double *out);
void orc_trsv_levels(int n, const int *rp, const int *ci, const int *diag, int *lev_l, int *lev_u);
void orc_lu_solve_avail(int n, const int *rp, const int *ci, const int *diag,
                        const int *lev_l, const int *lev_u, const double *F,
                        const double *v, double *tmp, double *out);
int orc_isai(int n, const int *rp, const int *ci, const int *diag,
             const double *vals, int type, double *out);

int orc_bicgstab(int n, const int *rp, const int *ci, const int *diag,
                 const double *A, const double *b, double *x,
                 int pc_kind, const double *P1, const double *P2,
                 const orc_opts *o, float *iter, int *conv,
                 double *iter_res, double *true_res);

/* whole-batch driver: precond setup + solve per system, optionally on
 * `threads` POSIX threads (systems statically partitioned).
 * precond: 0 = exact ILU(0), 1 = ParILU(0) with `sweeps` Jacobi sweeps,
 * 2 = ILU(0)+ISAI(L,U), 3 = ParILU(0)+ISAI(L,U), 4 = general ISAI(A),
 * 5 = none. */
int orc_batch_solve(int n, int nnz, const int *rp, const int *ci, int batch,
                    const double *A, const double *b, double *x,
                    int precond, int sweeps, const orc_opts *o,
                    float *iter, int *conv, double *iter_res, double *true_res,
                    int threads);

/* mixed precision: FP32 inner BiCGSTAB over FP32 factors F32 inside FP64
 * iterative refinement (tol / tol_mode / max_iter from o; the solves in
 * availability order with lev_l / lev_u from orc_trsv_levels) */
int orc_bicgstab_mixed(int n, const int *rp, const int *ci, const int *diag,
                       const int *lev_l, const int *lev_u, const double *A, const double *b,
                       const float *F32, const orc_opts *o, double inner_tol, int max_refine,
                       double *x, float *iter, int *conv, double *iter_res,
                       double *true_res, int *refine);
int orc_bicgstab_mixed_pc(int n, const int *rp, const int *ci, const int *diag,
                          const int *lev_l, const int *lev_u, const double *A, const double *b,
                          const float *F32, int isai, int zp, const orc_opts *o, double inner_tol,
                          int max_refine, double *x, float *iter, int *conv, double *iter_res,
                          double *true_res, int *refine);
int orc_isai_lu_f32(int n, const int *rp, const int *ci, const int *diag, const double *F, float *M);
int orc_entry_order(int n, const int *rp, const int *ci, const int *diag, int *order);
/* ParILU(0) sweeps in the factor kernel's round order (orc_sweep_order):
 * inplace = 0 Jacobi (every entry reads the sweep's starting iterate),
 * inplace = 1 rounds of 64 entries one after the other, each reading the
 * iterate the earlier rounds left (entries of one round read its start).
 * stop_tol > 0 stops on the summed squared residual (Jacobi: of the
 * starting iterate, which is kept; in place: as each round saw it, the
 * iterate after the sweep is kept) and falls back to exact ILU(0). */
int orc_parilu0_stop(int n, const int *rp, const int *ci, const int *diag, const double *A,
                     const double *init, double *F, int sweeps, double stop_tol, int *status,
                     int inplace);
int orc_batch_solve_split(int n, int nnz, const int *rp, const int *ci, int batch,
                          const double *A, const double *b, double *x, int precond, int sweeps,
                          const double *init, double stop_tol, int precision, double inner_tol,
                          int max_refine, const orc_opts *o, float *iter, int *conv,
                          double *iter_res, double *true_res, int *refine, int *fstat, int threads,
                          int parilu_inplace);

#ifdef __cplusplus
}
#endif
#endif
