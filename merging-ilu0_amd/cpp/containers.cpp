// HIP-backed batch containers of the drop-in C++ interface.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "SolverResults.h"
#include "matrix.h"

namespace {

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
void host_alloc(T *&p, size_t n)
{
    std::free(p);
    p = static_cast<T *>(std::calloc(n ? n : 1, sizeof(T)));
    if (!p) throw std::runtime_error("host allocation failed");
}

template <typename T>
void dev_alloc(T *&p, size_t n)
{
    if (p) (void)hipFree(p);
    p = nullptr;
    hip_check(hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)), "hipMalloc");
}

template <typename T>
void host_free(T *&p)
{
    std::free(p);
    p = nullptr;
}

template <typename T>
void dev_free(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename T>
void h2d(T *d, const T *h, size_t n) { hip_check(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D"); }
template <typename T>
void d2h(T *h, const T *d, size_t n) { hip_check(hipMemcpy(h, d, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H"); }

}  // namespace

// ---------------------------------------------------------------- PagedCSRMatrices
void PagedCSRMatrices::AllocateMemory(LOCATION loc)
{
    const size_t nv = (size_t)npages * nnz;
    if (loc == LOCATION::CPU) {
        host_alloc(col_ind_cpu, (size_t)nnz);
        host_alloc(row_ptrs_cpu, (size_t)nrows + 1);
        host_alloc(values_cpu, nv);
        cpu_exists = CPU_EXISTENCE::EXISTENT;
    } else {
        dev_alloc(col_ind_gpu, (size_t)nnz);
        dev_alloc(row_ptrs_gpu, (size_t)nrows + 1);
        dev_alloc(values_gpu, nv);
        gpu_exists = GPU_EXISTENCE::EXISTENT;
    }
}

void PagedCSRMatrices::DeallocateMemory(LOCATION loc)
{
    if (loc == LOCATION::CPU) {
        host_free(col_ind_cpu); host_free(row_ptrs_cpu); host_free(values_cpu);
        cpu_exists = CPU_EXISTENCE::NON_EXISTENT;
    } else {
        dev_free(col_ind_gpu); dev_free(row_ptrs_gpu); dev_free(values_gpu);
        gpu_exists = GPU_EXISTENCE::NON_EXISTENT;
    }
}

PagedCSRMatrices::~PagedCSRMatrices()
{
    DeallocateMemory(LOCATION::CPU);
    DeallocateMemory(LOCATION::GPU);
}

void PagedCSRMatrices::CopyFromCpuToGpu()
{
    if (!ExistsCPU()) throw std::runtime_error("PagedCSRMatrices: no host arrays to copy");
    if (!ExistsGPU()) AllocateMemory(LOCATION::GPU);
    h2d(col_ind_gpu, col_ind_cpu, (size_t)nnz);
    h2d(row_ptrs_gpu, row_ptrs_cpu, (size_t)nrows + 1);
    h2d(values_gpu, values_cpu, (size_t)npages * nnz);
}

void PagedCSRMatrices::CopyFromGpuToCpu()
{
    if (!ExistsGPU()) throw std::runtime_error("PagedCSRMatrices: no device arrays to copy");
    if (!ExistsCPU()) AllocateMemory(LOCATION::CPU);
    d2h(col_ind_cpu, col_ind_gpu, (size_t)nnz);
    d2h(row_ptrs_cpu, row_ptrs_gpu, (size_t)nrows + 1);
    d2h(values_cpu, values_gpu, (size_t)npages * nnz);
}

// ---------------------------------------------------------------- PagedVectors
void PagedVectors::AllocateMemory(LOCATION loc)
{
    const size_t n = (size_t)npages * nelements;
    if (loc == LOCATION::CPU) { host_alloc(values_cpu, n); cpu_exists = CPU_EXISTENCE::EXISTENT; }
    else { dev_alloc(values_gpu, n); gpu_exists = GPU_EXISTENCE::EXISTENT; }
}

void PagedVectors::DeallocateMemory(LOCATION loc)
{
    if (loc == LOCATION::CPU) { host_free(values_cpu); cpu_exists = CPU_EXISTENCE::NON_EXISTENT; }
    else { dev_free(values_gpu); gpu_exists = GPU_EXISTENCE::NON_EXISTENT; }
}

PagedVectors::~PagedVectors()
{
    DeallocateMemory(LOCATION::CPU);
    DeallocateMemory(LOCATION::GPU);
}

void PagedVectors::CopyFromCpuToGpu()
{
    if (!ExistsCPU()) throw std::runtime_error("PagedVectors: no host array to copy");
    if (!ExistsGPU()) AllocateMemory(LOCATION::GPU);
    h2d(values_gpu, values_cpu, (size_t)npages * nelements);
}

void PagedVectors::CopyFromGpuToCpu()
{
    if (!ExistsGPU()) throw std::runtime_error("PagedVectors: no device array to copy");
    if (!ExistsCPU()) AllocateMemory(LOCATION::CPU);
    d2h(values_cpu, values_gpu, (size_t)npages * nelements);
}

// ---------------------------------------------------------------- SolverResults
SolverResults::SolverResults(int num_pages, double atol, CPU_EXISTENCE cpu, GPU_EXISTENCE gpu)
    : npages(num_pages), abs_tol(atol)
{
    if (cpu == CPU_EXISTENCE::EXISTENT) AllocateMemory(LOCATION::CPU);
    if (gpu == GPU_EXISTENCE::EXISTENT) AllocateMemory(LOCATION::GPU);
}

void SolverResults::AllocateMemory(LOCATION loc)
{
    const size_t n = (size_t)npages;
    if (loc == LOCATION::CPU) {
        host_alloc(iter_count_cpu, n); host_alloc(conv_flag_cpu, n);
        host_alloc(true_residual_norm_cpu, n); host_alloc(iter_residual_norm_cpu, n);
        cpu_exists = CPU_EXISTENCE::EXISTENT;
    } else {
        dev_alloc(iter_count_gpu, n); dev_alloc(conv_flag_gpu, n);
        dev_alloc(true_residual_norm_gpu, n); dev_alloc(iter_residual_norm_gpu, n);
        gpu_exists = GPU_EXISTENCE::EXISTENT;
    }
}

void SolverResults::DeallocateMemory(LOCATION loc)
{
    if (loc == LOCATION::CPU) {
        host_free(iter_count_cpu); host_free(conv_flag_cpu);
        host_free(true_residual_norm_cpu); host_free(iter_residual_norm_cpu);
        cpu_exists = CPU_EXISTENCE::NON_EXISTENT;
    } else {
        dev_free(iter_count_gpu); dev_free(conv_flag_gpu);
        dev_free(true_residual_norm_gpu); dev_free(iter_residual_norm_gpu);
        gpu_exists = GPU_EXISTENCE::NON_EXISTENT;
    }
}

SolverResults::~SolverResults()
{
    DeallocateMemory(LOCATION::CPU);
    DeallocateMemory(LOCATION::GPU);
}

void SolverResults::CopyFromCpuToGpu()
{
    if (!ExistsCPU()) throw std::runtime_error("SolverResults: no host arrays to copy");
    if (!ExistsGPU()) AllocateMemory(LOCATION::GPU);
    const size_t n = (size_t)npages;
    h2d(iter_count_gpu, iter_count_cpu, n); h2d(conv_flag_gpu, conv_flag_cpu, n);
    h2d(true_residual_norm_gpu, true_residual_norm_cpu, n); h2d(iter_residual_norm_gpu, iter_residual_norm_cpu, n);
}

void SolverResults::CopyFromGpuToCpu()
{
    if (!ExistsGPU()) throw std::runtime_error("SolverResults: no device arrays to copy");
    if (!ExistsCPU()) AllocateMemory(LOCATION::CPU);
    const size_t n = (size_t)npages;
    d2h(iter_count_cpu, iter_count_gpu, n); d2h(conv_flag_cpu, conv_flag_gpu, n);
    d2h(true_residual_norm_cpu, true_residual_norm_gpu, n); d2h(iter_residual_norm_cpu, iter_residual_norm_gpu, n);
}
