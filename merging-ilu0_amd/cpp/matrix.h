// Batch containers of the drop-in C++ interface (reference matrix.h:6-120,
// 249-322): a batch ("pages") of CSR matrices sharing one pattern, and a
// batch of vectors, each with paired host / HIP device arrays.  Values are
// page-major (page p owns values[p * nnz .. p * nnz + nnz)).  Same accessor
// names and meaning as the reference; device memory is HIP memory.
#pragma once

#include "location_enums.h"

class PagedCSRMatrices {
    int nrows = 0, ncols = 0, nnz = 0, npages = 0;
    CPU_EXISTENCE cpu_exists = CPU_EXISTENCE::NON_EXISTENT;
    GPU_EXISTENCE gpu_exists = GPU_EXISTENCE::NON_EXISTENT;
    int *col_ind_cpu = nullptr, *row_ptrs_cpu = nullptr;
    double *values_cpu = nullptr;
    int *col_ind_gpu = nullptr, *row_ptrs_gpu = nullptr;
    double *values_gpu = nullptr;

public:
    PagedCSRMatrices() = default;
    ~PagedCSRMatrices();
    PagedCSRMatrices(const PagedCSRMatrices &) = delete;
    PagedCSRMatrices &operator=(const PagedCSRMatrices &) = delete;

    bool ExistsCPU() const { return cpu_exists == CPU_EXISTENCE::EXISTENT; }
    bool ExistsGPU() const { return gpu_exists == GPU_EXISTENCE::EXISTENT; }
    void SetNumPages(int pages) { npages = pages; }
    int GetNumPages() const { return npages; }
    void SetNumRows(int rows) { nrows = rows; }
    int GetNumRows() const { return nrows; }
    void SetNumCols(int cols) { ncols = cols; }
    int GetNumCols() const { return ncols; }
    void SetNumNz(int nz) { nnz = nz; }
    int GetNumNz() const { return nnz; }

    // allocate / free the arrays at `loc` for the current sizes (throws
    // std::runtime_error when a HIP allocation fails)
    void AllocateMemory(LOCATION loc);
    void DeallocateMemory(LOCATION loc);

    double *GetPtrToCpuValues() const { return values_cpu; }
    int *GetPtrToCpuColInd() const { return col_ind_cpu; }
    int *GetPtrToCpuRowPtrs() const { return row_ptrs_cpu; }
    double *GetPtrToGpuValues() const { return values_gpu; }
    int *GetPtrToGpuColInd() const { return col_ind_gpu; }
    int *GetPtrToGpuRowPtrs() const { return row_ptrs_gpu; }

    // copies between existing allocations (no reallocation, unlike the
    // reference's CopyFromGpuToCpu, matrix.cu:128-131)
    void CopyFromCpuToGpu();
    void CopyFromGpuToCpu();
};

class PagedVectors {
    int npages = 0, nelements = 0;
    CPU_EXISTENCE cpu_exists = CPU_EXISTENCE::NON_EXISTENT;
    GPU_EXISTENCE gpu_exists = GPU_EXISTENCE::NON_EXISTENT;
    double *values_cpu = nullptr, *values_gpu = nullptr;

public:
    PagedVectors() = default;
    ~PagedVectors();
    PagedVectors(const PagedVectors &) = delete;
    PagedVectors &operator=(const PagedVectors &) = delete;

    bool ExistsCPU() const { return cpu_exists == CPU_EXISTENCE::EXISTENT; }
    bool ExistsGPU() const { return gpu_exists == GPU_EXISTENCE::EXISTENT; }
    void SetNumPages(int pages) { npages = pages; }
    int GetNumPages() const { return npages; }
    void SetNumElements(int eles) { nelements = eles; }
    int GetNumElements() const { return nelements; }
    void AllocateMemory(LOCATION loc);
    void DeallocateMemory(LOCATION loc);
    double *GetPtrToCpuValues() const { return values_cpu; }
    double *GetPtrToGpuValues() const { return values_gpu; }
    void CopyFromCpuToGpu();
    void CopyFromGpuToCpu();
};
