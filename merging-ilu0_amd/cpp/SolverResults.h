// Per-page solver outputs of the drop-in C++ interface (reference
// SolverResults.h:7-128): iteration count (half steps, float), convergence
// flag, true and recursive residual norms, the tolerance and the time taken.
#pragma once

#include "location_enums.h"

class SolverResults {
    int npages = 0;
    double time_taken = 0, abs_tol = 0;
    CPU_EXISTENCE cpu_exists = CPU_EXISTENCE::NON_EXISTENT;
    GPU_EXISTENCE gpu_exists = GPU_EXISTENCE::NON_EXISTENT;
    float *iter_count_cpu = nullptr, *iter_count_gpu = nullptr;
    int *conv_flag_cpu = nullptr, *conv_flag_gpu = nullptr;
    double *true_residual_norm_cpu = nullptr, *true_residual_norm_gpu = nullptr;
    double *iter_residual_norm_cpu = nullptr, *iter_residual_norm_gpu = nullptr;

public:
    SolverResults() = default;
    SolverResults(int num_pages, double atol, CPU_EXISTENCE to_allocate_on_cpu, GPU_EXISTENCE to_allocate_on_gpu);
    ~SolverResults();
    SolverResults(const SolverResults &) = delete;
    SolverResults &operator=(const SolverResults &) = delete;

    bool ExistsCPU() const { return cpu_exists == CPU_EXISTENCE::EXISTENT; }
    bool ExistsGPU() const { return gpu_exists == GPU_EXISTENCE::EXISTENT; }
    void SetNumPages(int pages) { npages = pages; }
    int GetNumPages() const { return npages; }
    void SetATOL(double tol) { abs_tol = tol; }
    double GetATOL() const { return abs_tol; }
    void SetTimeTaken(double t) { time_taken = t; }
    double GetTimeTaken() const { return time_taken; }
    void AllocateMemory(LOCATION loc);
    void DeallocateMemory(LOCATION loc);
    float *GetPtrToCpuIterCount() const { return iter_count_cpu; }
    double *GetPtrToCpuTrueResNorm() const { return true_residual_norm_cpu; }
    double *GetPtrToCpuIterResNorm() const { return iter_residual_norm_cpu; }
    int *GetPtrToCpuConvFlag() const { return conv_flag_cpu; }
    float *GetPtrToGpuIterCount() const { return iter_count_gpu; }
    int *GetPtrToGpuConvFlag() const { return conv_flag_gpu; }
    double *GetPtrToGpuTrueResNorm() const { return true_residual_norm_gpu; }
    double *GetPtrToGpuIterResNorm() const { return iter_residual_norm_gpu; }
    void CopyFromCpuToGpu();
    void CopyFromGpuToCpu();
};
