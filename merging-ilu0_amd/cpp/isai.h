// Drop-in ISAI entry point (reference isai.h:7-9).  Power 1: aiA on the
// pattern of the input (values on the typed triangle, zeros elsewhere).
// Power k > 1: aiA on the pattern of |A|^k of the typed view (the reference
// takes the nonzeros of page 0's A^k, isai_gpu.cu:40-178; this is the
// structural pattern), host and device pattern arrays allocated here.  Rows
// of more than 32 entries raise std::runtime_error (the reference asserts,
// row_size_limit 32).
#pragma once

class PagedCSRMatrices;

enum class mat_type { general, lower_tri, upper_tri };

void GenerateISAI_gpu(PagedCSRMatrices &aiA_pages, const PagedCSRMatrices &A_pages, const mat_type A_type,
                      const int power);
