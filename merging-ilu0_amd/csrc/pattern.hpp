// Host-side analysis of one shared CSR pattern: everything the batched
// kernels need that depends only on the pattern, computed once and uploaded.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bilu {

// one triangular / general view of the pattern for ISAI (power 1)
struct IsaiMap {
    // rows: J = columns of row i inside the view, stored as positions
    // beg[i]..beg[i]+m[i]-1 of the pattern; gather[goff[i] + t*m + k] is the
    // pattern position of (J[k], J[t]) (the transposed dense block A(J,J)^T),
    // -1 for a structural zero, -2 for the implicit unit diagonal of L.
    std::vector<int> beg, m, goff, gather, rhs_pos;
    int max_m = 0;
};

struct PatternAnalysis {
    int n = 0, nnz = 0;
    std::vector<int> rp, ci, diag;
    // level schedules: rows grouped by dependency level
    std::vector<int> lev_l_ptr, lev_l_rows;  // lower solve and ILU(0) rows
    std::vector<int> lev_u_ptr, lev_u_rows;  // upper solve rows (bottom-up)
    // ILU(0) update lists, per strictly-lower position p=(i,k):
    // pairs (dst=(i,j), src=(k,j)) for j>k, (i,j) and (k,j) in the pattern
    std::vector<int> upd_ptr;   // nnz+1
    std::vector<int> upd;       // 2 ints per pair
    // ParILU(0) dependency lists, per position e=(i,j):
    // pairs (pos(i,k), pos(k,j)), k < min(i,j) increasing
    std::vector<int> dep_ptr;   // nnz+1
    std::vector<int> dep;       // 2 ints per pair
    // reference L/U split (L keeps an explicit unit diagonal)
    std::vector<int> l_rp, l_ci, l_src;   // l_src: pattern position or -1 (unit diagonal)
    std::vector<int> u_rp, u_ci, u_src;
    IsaiMap isai[3];  // indexed by BILU_ISAI_* type
    int max_row = 0;
};

// returns BILU_OK or an error code; `why` gets a message
int analyse_pattern(int n, int nnz, const int *rp, const int *ci, PatternAnalysis &pa,
                    std::string &why);

}  // namespace bilu
