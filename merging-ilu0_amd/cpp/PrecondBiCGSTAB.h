// Drop-in batched preconditioned BiCGSTAB entry points: the reference's
// solver-call surface (PrecondBiCGSTAB.h:15-30, same names, parameter lists
// and constants) over the HIP library libbilu, plus one overload carrying
// the extensions (tolerance mode, precision, stream, warm start).
//
// Reference meaning kept: tolerance = ATOL on the absolute recursive
// residual, MAX_ITER full iterations, x_0 = 0, per-page iteration count in
// half steps and flags 1 / -1 (library flags -2 zero pivot, -3 non-finite
// extend them).  After the solve, x and the results are copied back to the
// host arrays (allocated if absent), as the reference's helpers do, and
// when subdir is not empty the first subdir.size() pages of x are written
// to subdir[p]/x_gpu_<solver>.mtx (x_scaled_gpu_... when is_scaled), the
// reference's answer files.  Patterns without a full diagonal: the ILU and
// ILU-ISAI solvers insert explicit zero diagonals first, as the reference's
// factorizations do, and then solve on that pattern; the merged and
// segregated solvers reject them, as the reference asserts, and so does the
// general ISAI solver.  Errors raise std::runtime_error or
// std::invalid_argument (never exit / abort).
#pragma once

#include <string>
#include <vector>

class PagedCSRMatrices;
class PagedVectors;
class SolverResults;

#define MAX_ITER 500

#define ATOL 0.00000000001

// ILU(0) (is_parilu = false; every reference approach 1-4 computes the same
// exact ILU(0) factors, so `ilu_approach` picks only the parallel schedule:
// 1 rows by levels, 2 and 3 dependency lists by entry levels, 4 pivot eras)
// or ParILU(0) with `parilu_sweeps` sweeps, then BiCGSTAB on those factors.
// PST = factorization time, IST = solve time, OET = both (ms, HIP events).
void Batched_ILU_Preconditioned_BiCGSTAB_Gpu(const std::vector<std::string> &subdir, const PagedCSRMatrices &A_pages,
                                             const PagedVectors &b_pages, PagedVectors &x_pages, const bool is_scaled,
                                             SolverResults &solver_results, const bool is_parilu,
                                             const int ilu_approach, const int parilu_sweeps, double &PST,
                                             double &IST, double &OET);

// ILU(0) / ParILU(0) factors, ISAI(L), ISAI(U) (power 1), BiCGSTAB with the
// two-SpMV apply.  Setup and solve run in one call: PST = 0, IST = OET.
void Batched_ILU_ISAI_Preconditioned_BiCGSTAB_Gpu(const std::vector<std::string> &subdir, const PagedCSRMatrices &A_pages,
                                                  const PagedVectors &b_pages, PagedVectors &x_pages,
                                                  const bool is_scaled, SolverResults &solver_results,
                                                  const bool is_parilu, const int parilu_sweeps, double &PST,
                                                  double &IST, double &OET);

// general ISAI(A) (power 1), BiCGSTAB.  PST = 0, IST = OET.
void Batched_General_ISAI_Preconditioned_BiCGSTAB_Gpu(const std::vector<std::string> &subdir,
                                                      const PagedCSRMatrices &A_pages, const PagedVectors &b_pages,
                                                      PagedVectors &x_pages, const bool is_scaled,
                                                      SolverResults &solver_results, double &PST, double &IST,
                                                      double &OET);

// merged exact ILU(0) + BiCGSTAB (one launch); time in solver_results
void Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_merged_Gpu(const std::vector<std::string> &subdir,
                                                              const PagedCSRMatrices &A_pages,
                                                              const PagedVectors &b_pages, PagedVectors &x_pages,
                                                              const bool is_scaled, SolverResults &solver_results);

// factorization launch, factors through device memory, solve launch
void Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_segregated_Gpu(const std::vector<std::string> &subdir,
                                                                  const PagedCSRMatrices &A_pages,
                                                                  const PagedVectors &b_pages, PagedVectors &x_pages,
                                                                  const bool is_scaled, SolverResults &solver_results);

// ---- extension (no counterpart in the reference) ----
struct BiluSolveOptions {
    int precond = 0;              // BILU_PRECOND_* (include/bilu.h)
    int parilu_sweeps = 5;
    double tol = ATOL;
    int tol_mode = 0;             // BILU_TOL_ABSOLUTE (reference) or BILU_TOL_RELATIVE
    int max_iter = MAX_ITER;
    int merge = 0;                // BILU_MERGED / BILU_SEGREGATED
    int true_res_restart = 0;
    int precision = 0;            // BILU_PRECISION_FP64 / BILU_PRECISION_FP32_IR
    double inner_tol = 1e-5;
    int max_refine = 10;
    double parilu_stop_tol = 0.0;
    int isai_power = 1;           // ISAI sparsity |L|^k, |U|^k (|A|^k); the reference's ILU-ISAI uses 2
    int parilu_order = 0;         // BILU_PARILU_JACOBI / BILU_PARILU_INPLACE
    double inner_adapt = 0.5;     // FP32_IR: inner solves stop at max(inner_tol, inner_adapt tol / |r|)
    const double *d_F_init = nullptr;   // ParILU warm start (device, in-place factor layout)
    void *stream = nullptr;             // hipStream_t; nullptr = default stream
    bool copy_back = true;              // copy x and the results to the host arrays
};

// any preconditioner / precision through one call; returns the solve time (ms)
double Batched_Preconditioned_BiCGSTAB_Gpu(const PagedCSRMatrices &A_pages, const PagedVectors &b_pages,
                                           PagedVectors &x_pages, SolverResults &solver_results,
                                           const BiluSolveOptions &opts);
