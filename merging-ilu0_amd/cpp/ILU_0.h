// Drop-in factor-level entry points (reference ILU_0.h:7-11).
#pragma once

class PagedCSRMatrices;

// exact ILU(0) of every page; L (explicit unit diagonal) and U in the
// reference split layout (factorization.cu:216-344), device arrays
// allocated here (host pattern arrays too).  `approach_num` (1-4) selects
// the schedule of the same arithmetic (identical factors): 1 rows level by
// level, 2 and 3 entries by dependency lists, 4 right-looking pivot eras;
// other values throw std::invalid_argument.  A
// pattern without a full diagonal gets explicit zero diagonal entries first
// (ILU_0_gpu.cu:1262-1272); L and U are then split from that pattern.
void ILU_0_Factorization_Gpu(const PagedCSRMatrices &A_pages, PagedCSRMatrices &L_pages, PagedCSRMatrices &U_pages,
                             const int approach_num);

// in place on the device values: strictly lower = L (unit diagonal
// implicit), diagonal + upper = U; diagonal positions are found from the
// pattern (`diag_ptrs` may be null)
void ComputeILU0Approach1(PagedCSRMatrices &Factored_Pages, const int *const diag_ptrs);
