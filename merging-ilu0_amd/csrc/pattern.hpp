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
    // beg[i]..beg[i]+m[i]-1 of the pattern; gather[goff[i] + k*m + t] is the
    // pattern position of (J[k], J[t]) (the transposed dense block A(J,J)^T:
    // lane t holds its row t, and the lanes of one k read consecutive ints),
    // -1 for a structural zero, -2 for the implicit unit diagonal of L.
    std::vector<int> beg, m, goff, gather, rhs_pos;
    int max_m = 0;
    // rows by decreasing m (ties by row index); the first n_pair rows (those
    // with m > 16, rounded up to an even count) are solved two per wave, the
    // rest four per wave
    std::vector<int> order;
    int n_pair = 0;
};

// fills M.order / M.n_pair from M.m
void isai_row_order(IsaiMap &M);

// Schedule tables of the persistent multi-wave kernels (global memory; the
// waves read them through the caches), plus the per-lane SpMV slot layout
// that is held in VGPRs.
//   Triangular solves, "pull by availability": lane l owns rows l + 64 k
//   (k < NR) and keeps their partial sums in registers.  After level ℓ
//   finalises its rows, every lane folds in the entries of its rows whose
//   column lies in level ℓ.  tr{L,U}[s][64]: one u32 record per slot and
//   lane = 8*pos | (8*col + k) << 16 (LDS byte offsets into the factor and
//   the vector, k = the lane's row slot); padding records point at zero
//   cells (factor index nnz, vector index n).  Each level owns cnt[ℓ]
//   chunks of CH consecutive slots; a lane's entries of one row run in
//   increasing column order within a level, so each row sum is an fma chain
//   in (level of column, column) order.
//   rlev{L,U}[k][64]: level of row l + 64 k (0xFFFF: no such row).
//   A level's rows finalise only in the row slots k named by its mask.
//   Factorization: etab records (pos, dep_begin, dep_end, diag position of
//   the column or 0xFFFF for upper entries) grouped by dependency level of
//   the entry DAG (eptr); dep holds (pos(i,k), pos(k,j)) pairs, k increasing.
// solve records per lane and chunk (one chunk = CH instruction slots)
#ifndef BILU_SOLVE_CH
#define BILU_SOLVE_CH 4
#endif

struct FastPlan {
    bool ok = false;          // pattern fits the packed tables
    std::string why;          // reason when !ok
    int S = 0;                // SpMV slots per lane (balanced row -> lane assignment)
    // slot tables [S][64]: meta = 8*col | (8*row << 16) as LDS byte offsets,
    // 0xFFFF in the high half if the slot does not end a row; pos = pattern
    // position (-1 = padding)
    std::vector<uint32_t> slot_meta;
    std::vector<int32_t> slot_pos;
    static constexpr int CH = BILU_SOLVE_CH;   // solve slots per chunk
    int nr = 0;
    std::vector<uint32_t> trL, trU;
    // per level (index < 64): chunk count | (mask of the row slots k holding a
    // row of the level) << 8
    std::vector<uint16_t> cntL, cntU;
    std::vector<uint16_t> rlevL, rlevU;
    std::vector<uint16_t> etab, eptr, dep;   // etab: 4 u16 per record, dep: 2 u16 per pair
    int nlev_e = 0, ne = 0;
    std::vector<uint16_t> wcode;      // per position: column of a strictly lower entry, 0xFFFE an upper
                                      // entry with work, 0xFFFF no work
    // ISAI(L) / ISAI(U) generation streams [step][64] (two u32 records per step)
    std::vector<uint32_t> isL0, isL1, isU0, isU1;
    int isL_steps = 0, isU_steps = 0;
    // slot layouts of ISAI(L) (strict lower + unit diagonal) and ISAI(U)
    // (diagonal + strict upper) for the FP32 applies: packed records [S][64]
    // (x offset | output offset << 16 | not-a-row-end << 31, FP32 LDS byte
    // offsets) and coefficient positions (-2: unit diagonal, -1: padding);
    // SL = SU = 0 when the pattern lacks a diagonal entry
    int SL = 0, SU = 0;
    std::vector<uint32_t> isl_meta, isu_meta;
    std::vector<int32_t> isl_pos, isu_pos;
    // exact ILU(0) as per-lane step streams [step][64] (two u32 records per
    // step; list-scheduled over the entry DAG, padded to a multiple of
    // ILU_BATCH steps)
    std::vector<uint32_t> ilu0, ilu1;
    int ilu_steps = 0;
    // one ParILU(0) sweep (Jacobi or in place) in lane-synchronous rounds:
    // the entries with work sorted by product count (fewest first, stable:
    // an in-place sweep then meets most entries after their operands),
    // entry q of that order on lane q % 64 in round q / 64.  Products [step][64] (round r
    // takes max(1, its largest count) steps; bit 0 marks a round's last
    // step), padded to a multiple of PAR_BATCH; finish records [round][64].
    std::vector<uint32_t> parp, parf;
    int par_rounds = 0, par_psteps = 0;
};

// steps per record batch of the ParILU sweep streams
constexpr int PAR_BATCH = 8;
// steps per record batch of the ISAI generation streams (padded with no-op steps)
constexpr int ISAI_BATCH = 8;

// steps per record batch of the ILU(0) streams (records of a batch are
// loaded together, one batch ahead)
constexpr int ILU_BATCH = 8;
// steps between the write of a factor value and its first read
constexpr int ILU_GAP = 2;

struct PatternAnalysis {
    int n = 0, nnz = 0;
    std::vector<int> rp, ci, diag;
    // level schedules: rows grouped by dependency level
    std::vector<int> lev_l_ptr, lev_l_rows;  // lower solve and ILU(0) rows
    std::vector<int> lev_u_ptr, lev_u_rows;  // upper solve rows (bottom-up)
    std::vector<int> lev_l, lev_u;           // level of each row
    // ILU(0) update lists, per strictly-lower position p=(i,k):
    // pairs (dst=(i,j), src=(k,j)) for j>k, (i,j) and (k,j) in the pattern
    std::vector<int> upd_ptr;   // nnz+1
    std::vector<int> upd;       // 2 ints per pair
    // ParILU(0) dependency lists, per position e=(i,j):
    // pairs (pos(i,k), pos(k,j)), k < min(i,j) increasing
    std::vector<int> dep_ptr;   // nnz+1
    std::vector<int> dep;       // 2 ints per pair
    // exact ILU(0) by entries (the dependency-list schedule): the positions
    // with work (a product or a division), grouped by dependency level
    std::vector<int> lev_e_ptr, lev_e_ent;
    // exact ILU(0) by pivot eras (right-looking): era r divides the lower
    // entries (i,k) of its pivots k by u_kk (pairs pos, pos(k,k)), then
    // applies its pivots' updates; each updated entry (target) carries its
    // pairs (pos(i,k), pos(k,j)) of the era, k increasing
    std::vector<int> era_div_ptr, era_div;    // eras+1; 2 ints per division
    std::vector<int> era_tgt_ptr, era_tgt;    // eras+1; target positions
    std::vector<int> tgt_pair_ptr, tgt_pair;  // targets+1; 2 ints per pair
    // reference L/U split (L keeps an explicit unit diagonal)
    std::vector<int> l_rp, l_ci, l_src;   // l_src: pattern position or -1 (unit diagonal)
    std::vector<int> u_rp, u_ci, u_src;
    IsaiMap isai[3];  // indexed by BILU_ISAI_* type
    int max_row = 0;
    FastPlan fast;
};

void build_fast_plan(PatternAnalysis &pa);
void build_ilu_schedules(PatternAnalysis &pa);

// ISAI with a chosen sparsity: the pattern of |P|^power for the typed view P
// (LOWER: columns < i plus the unit diagonal, UPPER: columns >= i, GENERAL:
// all), and the gather map of its rows (positions beg[i] = mrp[i] of that
// pattern) from the pattern positions of the typed view.
void isai_power_pattern(const PatternAnalysis &pa, int type, int power, std::vector<int> &mrp,
                        std::vector<int> &mci);
void build_isai_on(const PatternAnalysis &pa, int type, const std::vector<int> &mrp, const std::vector<int> &mci,
                   IsaiMap &M);

// returns BILU_OK or an error code; `why` gets a message
int analyse_pattern(int n, int nnz, const int *rp, const int *ci, PatternAnalysis &pa,
                    std::string &why);

}  // namespace bilu
