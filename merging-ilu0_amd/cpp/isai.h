// Drop-in ISAI entry point (reference isai.h:7-9).  Power 1 (the pattern of
// the input) is supported; other powers raise std::invalid_argument.
#pragma once

class PagedCSRMatrices;

enum class mat_type { general, lower_tri, upper_tri };

void GenerateISAI_gpu(PagedCSRMatrices &aiA_pages, const PagedCSRMatrices &A_pages, const mat_type A_type,
                      const int power);
