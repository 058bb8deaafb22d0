// Batch input / output of the drop-in C++ interface (reference
// ReadWriteData.h): a category directory holds one subdirectory per
// problem with A.mtx / b.mtx (or A_scaled.mtx / b_scaled.mtx) in Matrix
// Market format; all problems share one sparsity pattern.  Same names and
// meaning as the reference; errors raise std::runtime_error instead of
// exiting, and the subdirectories are listed in sorted order.
#pragma once

#include <string>
#include <vector>

class PagedVectors;
class PagedCSRMatrices;
class SolverResults;

// write page p < subdir.size() of Ans_pages to subdir[p] + "/" + ans_name
// (Matrix Market array, %0.17lg)
void Print_ans(const std::vector<std::string> &subdir, const PagedVectors &Ans_pages, const std::string &ans_name);

// every entry of main_dir except . and .. (main_dir ends with '/')
void FillSubdir(const std::string &main_dir, std::vector<std::string> &subdir);

// Problem_Size * subdir.size() pages, page p read from subdir[p % subdir.size()];
// host arrays of A_pages / b_pages allocated here
void ReadData(const int Problem_Size, const std::vector<std::string> &subdir, PagedCSRMatrices &A_pages,
              PagedVectors &b_pages, const bool is_scaled);

// one line "conv_flag   iteration_count" per subdirectory
void PrintSolverResults(const std::vector<std::string> &subdir, const SolverResults &solver_results,
                        const std::string &f_name);
