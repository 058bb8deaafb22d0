// Drop-in ParILU(0) entry point (reference parILU_0.h:7): Jacobi sweeps
// from L = strict-lower(A) + I, U = upper(A); L / U in the reference split
// layout.  (The reference sweeps asynchronously; Jacobi has the same fixed
// point and is deterministic.)  Missing diagonal entries are inserted as
// explicit zeros first (parILU_0_gpu.cu:309-319).
#pragma once

class PagedCSRMatrices;

void ParILU_0_Factorization_Gpu(const PagedCSRMatrices &A_pages, PagedCSRMatrices &L_pages, PagedCSRMatrices &U_pages,
                                const int num_iterations);
