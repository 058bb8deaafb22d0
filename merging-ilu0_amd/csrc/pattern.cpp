// Host-side pattern analysis (once per pattern, outside any timed region).
//
// The reference recomputes parts of this inside every solve call: the
// diagonal locator (factorization.cu:392-480, with a one-thread prefix sum
// and a blocking D2H), the L/U split (factorization.cu:487-575), the ParILU
// dependency graph (parILU_0_gpu.cu:118-220, host) and the ISAI dense index
// maps (isai_gpu.cu:209-287).  Here they are derived once from the host CSR
// arrays and cached with the pattern handle.
#include "pattern.hpp"

#include <algorithm>
#include <array>

#include "../../include/bilu.h"

namespace bilu {

static int find_pos(const std::vector<int> &rp, const std::vector<int> &ci, int r, int c)
{
    auto b = ci.begin() + rp[r], e = ci.begin() + rp[r + 1];
    auto it = std::lower_bound(b, e, c);
    return (it != e && *it == c) ? int(it - ci.begin()) : -1;
}

static void build_levels(int n, const std::vector<int> &lev, std::vector<int> &ptr,
                         std::vector<int> &rows)
{
    int nlev = 0;
    for (int i = 0; i < n; i++) nlev = std::max(nlev, lev[i] + 1);
    ptr.assign(nlev + 1, 0);
    for (int i = 0; i < n; i++) ptr[lev[i] + 1]++;
    for (int l = 0; l < nlev; l++) ptr[l + 1] += ptr[l];
    rows.assign(n, 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; i++) rows[fill[lev[i]]++] = i;
}

void isai_row_order(IsaiMap &M)
{
    const int n = int(M.m.size());
    M.order.resize(n);
    for (int i = 0; i < n; i++) M.order[i] = i;
    std::stable_sort(M.order.begin(), M.order.end(), [&](int a, int b) { return M.m[a] > M.m[b]; });
    int big = 0;
    while (big < n && M.m[M.order[big]] > 16) big++;
    M.n_pair = std::min(n, (big + 1) / 2 * 2);
}

static void build_isai(const PatternAnalysis &pa, int type, IsaiMap &M)
{
    const int n = pa.n;
    M.beg.assign(n, 0);
    M.m.assign(n, 0);
    M.goff.assign(n + 1, 0);
    M.rhs_pos.assign(n, 0);
    M.gather.clear();
    M.max_m = 0;
    for (int i = 0; i < n; i++) {
        int b, e;
        if (type == BILU_ISAI_LOWER) { b = pa.rp[i]; e = pa.diag[i] + 1; }
        else if (type == BILU_ISAI_UPPER) { b = pa.diag[i]; e = pa.rp[i + 1]; }
        else { b = pa.rp[i]; e = pa.rp[i + 1]; }
        const int m = e - b;
        M.beg[i] = b;
        M.m[i] = m;
        M.max_m = std::max(M.max_m, m);
        for (int q = 0; q < m; q++)
            if (pa.ci[b + q] == i) M.rhs_pos[i] = q;
        for (int k = 0; k < m; k++) {      // k-major: lane t reads k * m + t
            for (int t = 0; t < m; t++) {
                const int r = pa.ci[b + k], c = pa.ci[b + t];  // entry (J[k], J[t])
                int g;
                if (type == BILU_ISAI_LOWER && r == c) g = -2;
                else if (type == BILU_ISAI_LOWER && r < c) g = -1;
                else if (type == BILU_ISAI_UPPER && r > c) g = -1;
                else g = find_pos(pa.rp, pa.ci, r, c);
                M.gather.push_back(g);
            }
        }
        M.goff[i + 1] = int(M.gather.size());
    }
    isai_row_order(M);
}

// Rows of P^k by unions: row i of P^k = the union over j in row i of P^(k-1)
// of row j of P (a marker array, then a scan for the sorted order).  The
// reference takes the nonzeros of the first page's A^power computed densely
// (isai_gpu.cu:40-178); the structural pattern is that set without exact
// numeric cancellations.
void isai_power_pattern(const PatternAnalysis &pa, int type, int power, std::vector<int> &mrp,
                        std::vector<int> &mci)
{
    const int n = pa.n;
    std::vector<int> brp(n + 1, 0), bci;
    for (int i = 0; i < n; i++) {
        for (int p = pa.rp[i]; p < pa.rp[i + 1]; p++) {
            const int c = pa.ci[p];
            if ((type == BILU_ISAI_LOWER && c < i) || (type == BILU_ISAI_UPPER && c >= i) || type == BILU_ISAI_GENERAL)
                bci.push_back(c);
        }
        if (type == BILU_ISAI_LOWER) bci.push_back(i);
        brp[i + 1] = int(bci.size());
    }
    mrp = brp;
    mci = bci;
    std::vector<char> mark(n, 0);
    for (int k = 2; k <= power; k++) {
        std::vector<int> nrp(n + 1, 0), nci;
        for (int i = 0; i < n; i++) {
            int len = 0;
            for (int p = mrp[i]; p < mrp[i + 1]; p++) {
                const int j = mci[p];
                for (int q = brp[j]; q < brp[j + 1]; q++)
                    if (!mark[bci[q]]) { mark[bci[q]] = 1; len++; }
            }
            for (int c = 0, w = 0; c < n && w < len; c++)
                if (mark[c]) { mark[c] = 0; nci.push_back(c); w++; }
            nrp[i + 1] = int(nci.size());
        }
        mrp.swap(nrp);
        mci.swap(nci);
    }
}

void build_isai_on(const PatternAnalysis &pa, int type, const std::vector<int> &mrp, const std::vector<int> &mci,
                   IsaiMap &M)
{
    const int n = pa.n;
    M.beg.assign(n, 0);
    M.m.assign(n, 0);
    M.goff.assign(n + 1, 0);
    M.rhs_pos.assign(n, 0);
    M.gather.clear();
    M.max_m = 0;
    for (int i = 0; i < n; i++) {
        const int b = mrp[i], m = mrp[i + 1] - b;
        M.beg[i] = b;
        M.m[i] = m;
        M.max_m = std::max(M.max_m, m);
        for (int q = 0; q < m; q++)
            if (mci[b + q] == i) M.rhs_pos[i] = q;
        for (int k = 0; k < m; k++) {      // k-major: lane t reads k * m + t
            for (int t = 0; t < m; t++) {
                const int r = mci[b + k], c = mci[b + t];  // entry (J[k], J[t]) of the typed view
                int g;
                if (type == BILU_ISAI_LOWER && r == c) g = -2;
                else if (type == BILU_ISAI_LOWER && r < c) g = -1;
                else if (type == BILU_ISAI_UPPER && r > c) g = -1;
                else g = find_pos(pa.rp, pa.ci, r, c);
                M.gather.push_back(g);
            }
        }
        M.goff[i + 1] = int(M.gather.size());
    }
    isai_row_order(M);
}

int analyse_pattern(int n, int nnz, const int *rp, const int *ci, PatternAnalysis &pa,
                    std::string &why)
{
    if (n <= 0 || nnz <= 0 || !rp || !ci) { why = "empty pattern"; return BILU_ERR_INVALID_ARG; }
    if (rp[0] != 0 || rp[n] != nnz) { why = "row_ptrs[0] must be 0 and row_ptrs[n] == nnz"; return BILU_ERR_INVALID_PATTERN; }
    for (int i = 0; i < n; i++) {
        if (rp[i + 1] < rp[i]) { why = "row_ptrs not monotone"; return BILU_ERR_INVALID_PATTERN; }
        for (int p = rp[i]; p < rp[i + 1]; p++) {
            if (ci[p] < 0 || ci[p] >= n) { why = "column index out of range"; return BILU_ERR_INVALID_PATTERN; }
            if (p > rp[i] && ci[p] <= ci[p - 1]) { why = "columns not strictly increasing within a row"; return BILU_ERR_INVALID_PATTERN; }
        }
    }
    pa.n = n;
    pa.nnz = nnz;
    pa.rp.assign(rp, rp + n + 1);
    pa.ci.assign(ci, ci + nnz);
    pa.diag.assign(n, -1);
    pa.max_row = 0;
    for (int i = 0; i < n; i++) {
        pa.diag[i] = find_pos(pa.rp, pa.ci, i, i);
        pa.max_row = std::max(pa.max_row, rp[i + 1] - rp[i]);
        if (pa.diag[i] < 0) { why = "missing diagonal entry in row " + std::to_string(i); return BILU_ERR_MISSING_DIAGONAL; }
    }
    // level schedules
    std::vector<int> lev(n, 0);
    for (int i = 0; i < n; i++) {
        int l = 0;
        for (int p = rp[i]; p < pa.diag[i]; p++) l = std::max(l, lev[ci[p]] + 1);
        lev[i] = l;
    }
    build_levels(n, lev, pa.lev_l_ptr, pa.lev_l_rows);
    pa.lev_l = lev;
    std::fill(lev.begin(), lev.end(), 0);
    for (int i = n - 1; i >= 0; i--) {
        int l = 0;
        for (int p = pa.diag[i] + 1; p < rp[i + 1]; p++) l = std::max(l, lev[ci[p]] + 1);
        lev[i] = l;
    }
    build_levels(n, lev, pa.lev_u_ptr, pa.lev_u_rows);
    pa.lev_u = lev;
    // ILU(0) update lists
    pa.upd_ptr.assign(nnz + 1, 0);
    pa.upd.clear();
    for (int i = 0; i < n; i++) {
        for (int p = rp[i]; p < rp[i + 1]; p++) {
            const int k = ci[p];
            if (k < i) {
                for (int q = pa.diag[k] + 1; q < rp[k + 1]; q++) {
                    const int dst = find_pos(pa.rp, pa.ci, i, ci[q]);
                    if (dst >= 0) { pa.upd.push_back(dst); pa.upd.push_back(q); }
                }
            }
            pa.upd_ptr[p + 1] = int(pa.upd.size() / 2);
        }
    }
    // ParILU(0) dependency lists
    pa.dep_ptr.assign(nnz + 1, 0);
    pa.dep.clear();
    for (int i = 0; i < n; i++) {
        for (int e = rp[i]; e < rp[i + 1]; e++) {
            const int j = ci[e];
            const int kmax = std::min(i, j) - 1;
            for (int p = rp[i]; p < pa.diag[i] && ci[p] <= kmax; p++) {
                const int q = find_pos(pa.rp, pa.ci, ci[p], j);
                if (q >= 0) { pa.dep.push_back(p); pa.dep.push_back(q); }
            }
            pa.dep_ptr[e + 1] = int(pa.dep.size() / 2);
        }
    }
    build_ilu_schedules(pa);
    // reference L/U split (factorization.cu:216-344)
    pa.l_rp.assign(n + 1, 0);
    pa.u_rp.assign(n + 1, 0);
    pa.l_ci.clear(); pa.l_src.clear(); pa.u_ci.clear(); pa.u_src.clear();
    for (int i = 0; i < n; i++) {
        for (int p = rp[i]; p < pa.diag[i]; p++) { pa.l_ci.push_back(ci[p]); pa.l_src.push_back(p); }
        pa.l_ci.push_back(i); pa.l_src.push_back(-1);
        for (int p = pa.diag[i]; p < rp[i + 1]; p++) { pa.u_ci.push_back(ci[p]); pa.u_src.push_back(p); }
        pa.l_rp[i + 1] = int(pa.l_ci.size());
        pa.u_rp[i + 1] = int(pa.u_ci.size());
    }
    for (int t = 0; t < 3; t++) build_isai(pa, t, pa.isai[t]);
    build_fast_plan(pa);
    return BILU_OK;
}


// Alternative schedules of the same exact ILU(0) arithmetic (each entry:
// acc = a_ij, acc = fma(-l_ik, u_kj, acc) for k increasing, then acc / u_jj
// below the diagonal), so every schedule gives the same bits:
//  - entries: an entry's level is one above the levels of its operands (its
//    pairs and, below the diagonal, u_jj); the dependency lists of
//    ILU_0_gpu.cu:693-935 processed level by level instead of by ready flags;
//  - eras (right-looking, ILU_0_gpu.cu:947-1229): pivot k joins the era after
//    every pivot k' < k that writes row k or column k ((k,k') or (k',k) in the
//    pattern), and no earlier era than the previous pivot of any entry it
//    updates, so each entry sees its updates in increasing k.
void build_ilu_schedules(PatternAnalysis &pa)
{
    const int n = pa.n, nnz = pa.nnz;
    const auto &rp = pa.rp, &ci = pa.ci, &diag = pa.diag;
    // entry levels
    std::vector<int> elev(nnz, 0);
    int maxlev = 0;
    for (int i = 0; i < n; i++)
        for (int e = rp[i]; e < rp[i + 1]; e++) {
            int l = 0;
            for (int d = pa.dep_ptr[e]; d < pa.dep_ptr[e + 1]; d++)
                l = std::max(l, std::max(elev[pa.dep[2 * d]], elev[pa.dep[2 * d + 1]]) + 1);
            if (ci[e] < i) l = std::max(l, elev[diag[ci[e]]] + 1);
            elev[e] = l;
            maxlev = std::max(maxlev, l);
        }
    pa.lev_e_ptr.assign(1, 0);
    pa.lev_e_ent.clear();
    for (int l = 1; l <= maxlev; l++) {
        for (int e = 0; e < nnz; e++)
            if (elev[e] == l) pa.lev_e_ent.push_back(e);
        pa.lev_e_ptr.push_back(int(pa.lev_e_ent.size()));
    }
    // pivot eras
    std::vector<std::vector<int>> upper_rows(n), prev_piv(n);
    for (int i = 0; i < n; i++)
        for (int p = diag[i] + 1; p < rp[i + 1]; p++) upper_rows[ci[p]].push_back(i);
    for (int e = 0; e < nnz; e++) {
        int prev = -1;
        for (int d = pa.dep_ptr[e]; d < pa.dep_ptr[e + 1]; d++) {
            const int k = ci[pa.dep[2 * d]];
            if (prev >= 0) prev_piv[k].push_back(prev);
            prev = k;
        }
    }
    std::vector<int> era(n, 0);
    int neras = 0;
    for (int k = 0; k < n; k++) {
        int r = 0;
        for (int p = rp[k]; p < diag[k]; p++) r = std::max(r, era[ci[p]] + 1);
        for (int kk : upper_rows[k]) r = std::max(r, era[kk] + 1);
        for (int kk : prev_piv[k]) r = std::max(r, era[kk]);
        era[k] = r;
        neras = std::max(neras, r + 1);
    }
    if (n == 0) neras = 0;
    // per era: divisions (column k below the diagonal, pivots increasing)
    std::vector<std::vector<int>> col_lower(n);
    for (int i = 0; i < n; i++)
        for (int p = rp[i]; p < diag[i]; p++) col_lower[ci[p]].push_back(p);
    std::vector<std::vector<int>> era_piv(neras);
    for (int k = 0; k < n; k++) era_piv[era[k]].push_back(k);
    pa.era_div_ptr.assign(1, 0);
    pa.era_div.clear();
    for (int r = 0; r < neras; r++) {
        for (int k : era_piv[r])
            for (int p : col_lower[k]) { pa.era_div.push_back(p); pa.era_div.push_back(diag[k]); }
        pa.era_div_ptr.push_back(int(pa.era_div.size() / 2));
    }
    // per era: targets with their pairs (a target's pivots have non-decreasing
    // eras, so each era's pairs are one run of its dependency list)
    std::vector<std::vector<std::pair<int, int>>> runs(neras);   // (target, first dep) per era
    for (int e = 0; e < nnz; e++) {
        int d = pa.dep_ptr[e];
        while (d < pa.dep_ptr[e + 1]) {
            const int r = era[ci[pa.dep[2 * d]]];
            runs[r].push_back({e, d});
            while (d < pa.dep_ptr[e + 1] && era[ci[pa.dep[2 * d]]] == r) d++;
        }
    }
    pa.era_tgt_ptr.assign(1, 0);
    pa.era_tgt.clear();
    pa.tgt_pair_ptr.assign(1, 0);
    pa.tgt_pair.clear();
    for (int r = 0; r < neras; r++) {
        for (const auto &td : runs[r]) {
            const int e = td.first;
            pa.era_tgt.push_back(e);
            for (int d = td.second; d < pa.dep_ptr[e + 1] && era[ci[pa.dep[2 * d]]] == r; d++) {
                pa.tgt_pair.push_back(pa.dep[2 * d]);
                pa.tgt_pair.push_back(pa.dep[2 * d + 1]);
            }
            pa.tgt_pair_ptr.push_back(int(pa.tgt_pair.size() / 2));
        }
        pa.era_tgt_ptr.push_back(int(pa.era_tgt.size()));
    }
}

// Longest-processing-time assignment of whole rows (len(i) entries each) to
// the 64 lanes of a wave, improved by moves and swaps between the fullest
// lane and the others; every lane keeps its rows in increasing order.
// Returns the slots per lane (the fullest lane's load).
template <typename Len>
static int balance_rows(int n, Len len, std::vector<std::vector<int>> &rows)
{
    std::vector<int> order;
    for (int i = 0; i < n; i++)
        if (len(i) > 0) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return len(a) > len(b); });
    std::vector<int> load(64, 0);
    rows.assign(64, {});
    for (int i : order) {
        int best = 0;
        for (int l = 1; l < 64; l++) if (load[l] < load[best]) best = l;
        load[best] += len(i);
        rows[best].push_back(i);
    }
    for (int iter = 0; iter < 4096; iter++) {
        int hi = 0;
        for (int l = 1; l < 64; l++) if (load[l] > load[hi]) hi = l;
        bool improved = false;
        for (int l = 0; l < 64 && !improved; l++) {
            if (l == hi) continue;
            for (size_t a = 0; a < rows[hi].size() && !improved; a++) {
                const int ra = rows[hi][a];
                if (load[l] + len(ra) < load[hi]) {   // move
                    load[l] += len(ra); load[hi] -= len(ra);
                    rows[l].push_back(ra); rows[hi].erase(rows[hi].begin() + a);
                    improved = true;
                    break;
                }
                for (size_t b = 0; b < rows[l].size(); b++) {   // swap
                    const int rb = rows[l][b];
                    const int d = len(ra) - len(rb);
                    if (d > 0 && load[l] + d < load[hi]) {
                        load[l] += d; load[hi] -= d;
                        std::swap(rows[hi][a], rows[l][b]);
                        improved = true;
                        break;
                    }
                }
            }
        }
        if (!improved) break;
    }
    int S = 0;
    for (int l = 0; l < 64; l++) { std::sort(rows[l].begin(), rows[l].end()); S = std::max(S, load[l]); }
    return S;
}

void build_fast_plan(PatternAnalysis &pa)
{
    FastPlan &fp = pa.fast;
    fp = FastPlan();
    const int n = pa.n, nnz = pa.nnz;
    if (n >= 0xFFFF || nnz >= 0xFFFF) { fp.why = "n or nnz beyond the u16 schedule tables"; return; }
    // balanced SpMV slot layout of A: whole rows per lane, each row sum an
    // fma chain left to right over the CSR entries
    std::vector<std::vector<int>> rows;
    const int S = balance_rows(n, [&](int i) { return pa.rp[i + 1] - pa.rp[i]; }, rows);
    fp.S = S;
    fp.slot_meta.assign(size_t(S) * 64, 0xFFFF0000u);
    fp.slot_pos.assign(size_t(S) * 64, -1);
    for (int l = 0; l < 64; l++) {
        int s = 0;
        for (int i : rows[l]) {
            for (int p = pa.rp[i]; p < pa.rp[i + 1]; p++, s++) {
                const uint32_t row = (p == pa.rp[i + 1] - 1) ? uint32_t(i) : 0xFFFFu;
                fp.slot_meta[size_t(s) * 64 + l] = uint32_t(pa.ci[p]) | (row << 16);
                fp.slot_pos[size_t(s) * 64 + l] = p;
            }
        }
    }
    // slot layouts of the two triangles for the FP32 ISAI applies: ISAI(L)
    // over the strictly lower entries and the (unit) diagonal, ISAI(U) over
    // the diagonal and the strictly upper entries; whole rows per lane, CSR
    // order.  Records hold FP32 LDS byte offsets: x offset (bits 0-15) | output
    // offset (bits 16-30: the row's cell at a row end, else the scratch cell
    // n + 1) | bit 31 when the slot does not end its row; padding slots read
    // the zero cell n.  Positions: the pattern position of the coefficient,
    // -2 for ISAI(L)'s unit diagonal, -1 for padding.
    if (n < 8190) {
        for (int t = 0; t < 2; t++) {
            const bool lower = t == 0;
            auto beg = [&](int i) { return lower ? pa.rp[i] : pa.diag[i]; };
            auto end = [&](int i) { return lower ? pa.diag[i] + 1 : pa.rp[i + 1]; };
            std::vector<std::vector<int>> trows;
            bool diag_ok = true;
            for (int i = 0; i < n; i++) diag_ok &= pa.diag[i] >= 0;
            if (!diag_ok) break;
            const int ST = balance_rows(n, [&](int i) { return end(i) - beg(i); }, trows);
            const uint32_t pad = uint32_t(n) * 4u | ((uint32_t(n + 1) * 4u) << 16) | (1u << 31);
            std::vector<uint32_t> &meta = lower ? fp.isl_meta : fp.isu_meta;
            std::vector<int32_t> &pos = lower ? fp.isl_pos : fp.isu_pos;
            (lower ? fp.SL : fp.SU) = ST;
            meta.assign(size_t(ST) * 64, pad);
            pos.assign(size_t(ST) * 64, -1);
            for (int l = 0; l < 64; l++) {
                int s = 0;
                for (int i : trows[l]) {
                    for (int p = beg(i); p < end(i); p++, s++) {
                        const bool last = p == end(i) - 1;
                        const uint32_t y = last ? uint32_t(i) * 4u : uint32_t(n + 1) * 4u;
                        meta[size_t(s) * 64 + l] = uint32_t(pa.ci[p]) * 4u | (y << 16) | (last ? 0u : 1u << 31);
                        pos[size_t(s) * 64 + l] = (lower && p == pa.diag[i]) ? -2 : p;
                    }
                }
            }
        }
    }
    // pull-by-availability triangular-solve schedules
    const int NR = (n + 63) / 64;
    fp.nr = NR;
    if (nnz > 8188 || n > 4094 || NR > 7) { fp.why = "pattern beyond the packed solve records"; return; }
    // padding records point at the zero cells (factor index nnz, vector index n), row slot 0
    const uint32_t pad_rec = uint32_t(nnz * 8) | (uint32_t(n * 8) << 16);
    auto trsv_sched = [&](bool lower, std::vector<uint32_t> &rec, std::vector<uint16_t> &cnt,
                          std::vector<uint16_t> &rlev) -> bool {
        const std::vector<int> &lv = lower ? pa.lev_l : pa.lev_u;
        int nlev = 0;
        for (int i = 0; i < n; i++) nlev = std::max(nlev, lv[i] + 1);
        if (nlev > 64) return false;
        // E[level][lane]: (col, pos, k) of the entries of the lane's rows whose column sits in that level
        std::vector<std::vector<std::vector<std::array<int, 3>>>> E(nlev, std::vector<std::vector<std::array<int, 3>>>(64));
        for (int i = 0; i < n; i++) {
            const int b = lower ? pa.rp[i] : pa.diag[i] + 1, e = lower ? pa.diag[i] : pa.rp[i + 1];
            for (int p = b; p < e; p++) E[lv[pa.ci[p]]][i % 64].push_back({pa.ci[p], p, i / 64});
        }
        rec.clear();
        cnt.assign(64, 0);
        for (int l = 0; l < nlev; l++) {
            size_t S = 0;
            for (int t = 0; t < 64; t++) {
                std::sort(E[l][t].begin(), E[l][t].end());
                S = std::max(S, E[l][t].size());
            }
            const size_t chunks = (S + FastPlan::CH - 1) / FastPlan::CH;
            if (chunks > 255) return false;
            cnt[l] = uint16_t(chunks);
            for (size_t sl = 0; sl < chunks * FastPlan::CH; sl++)
                for (int t = 0; t < 64; t++) {
                    if (sl < E[l][t].size()) {
                        const auto &x = E[l][t][sl];
                        rec.push_back(uint32_t(x[1] * 8) | (uint32_t(x[0] * 8 + x[2]) << 16));
                    } else {
                        rec.push_back(pad_rec);
                    }
                }
        }
        rec.insert(rec.end(), size_t(64) * FastPlan::CH, pad_rec);   // one chunk of tail padding (prefetch)
        rlev.assign(size_t(NR) * 64, 0xFFFF);
        for (int i = 0; i < n; i++) {
            rlev[size_t(i / 64) * 64 + i % 64] = uint16_t(lv[i]);
            cnt[lv[i]] |= uint16_t(1u << (8 + i / 64));
        }
        return true;
    };
    if (!trsv_sched(true, fp.trL, fp.cntL, fp.rlevL) || !trsv_sched(false, fp.trU, fp.cntU, fp.rlevU)) {
        fp.why = "more than 64 solve levels";
        return;
    }
    // SpMV slot metadata as LDS byte offsets: x offset | (row offset or 0xFFFF) << 16
    for (auto &m : fp.slot_meta) {
        const uint32_t col = m & 0xFFFFu, row = m >> 16;
        m = (col * 8) | ((row == 0xFFFFu ? 0xFFFFu : row * 8) << 16);
    }
    // entry-DAG levels of the factorization (pull form of the KIJ updates)
    std::vector<int> elev(nnz, 0), pdiv(nnz, 0xFFFF);
    int maxlev = 0;
    for (int i = 0; i < n; i++) {
        for (int e = pa.rp[i]; e < pa.rp[i + 1]; e++) {
            int l = 0;
            for (int d = pa.dep_ptr[e]; d < pa.dep_ptr[e + 1]; d++)
                l = std::max(l, std::max(elev[pa.dep[2 * d]], elev[pa.dep[2 * d + 1]]) + 1);
            if (pa.ci[e] < i) { pdiv[e] = pa.diag[pa.ci[e]]; l = std::max(l, elev[pdiv[e]] + 1); }
            elev[e] = l;
            maxlev = std::max(maxlev, l);
        }
    }
    fp.etab.clear(); fp.dep.clear(); fp.eptr.assign(1, 0);
    for (int l = 1; l <= maxlev; l++) {
        for (int e = 0; e < nnz; e++) {
            if (elev[e] != l) continue;
            const int db = int(fp.dep.size() / 2);
            for (int d = pa.dep_ptr[e]; d < pa.dep_ptr[e + 1]; d++) {
                fp.dep.push_back(uint16_t(pa.dep[2 * d])); fp.dep.push_back(uint16_t(pa.dep[2 * d + 1]));
            }
            const int de = int(fp.dep.size() / 2);
            if (de >= 0xFFFF) { fp.why = "factorization dependency list beyond u16"; return; }
            fp.etab.push_back(uint16_t(e)); fp.etab.push_back(uint16_t(db));
            fp.etab.push_back(uint16_t(de)); fp.etab.push_back(uint16_t(pdiv[e]));
        }
        fp.eptr.push_back(uint16_t(fp.etab.size() / 4));
    }
    fp.nlev_e = maxlev;
    fp.ne = int(fp.etab.size() / 4);
    // per position: the column j of a strictly lower entry (every one has
    // work: a division), 0xFFFE for an upper entry with work, 0xFFFF without
    fp.wcode.assign(size_t(nnz) + 1, 0xFFFF);
    for (int q = 0; q < fp.ne; q++) fp.wcode[fp.etab[4 * q]] = 0xFFFE;
    for (int i = 0; i < n; i++)
        for (int p = pa.rp[i]; p < pa.rp[i + 1]; p++)
            if (pa.ci[p] < i) fp.wcode[p] = uint16_t(pa.ci[p]);

    // ParILU sweep rounds: a product record holds the byte offsets of the pair
    // (l, u) in the source iterate (the zero cell nnz for padding) | 1 on the
    // round's last step; a finish record the byte offset of the entry | 1 for
    // a lane with an entry | byte offset of its pivot << 16 (0xFFFF: none).
    {
        const uint32_t ZC = uint32_t(nnz) * 8u;
        std::vector<int> ord(fp.ne);
        for (int q = 0; q < fp.ne; q++) ord[q] = q;
        auto cnt = [&](int q) { return int(fp.etab[4 * q + 2]) - int(fp.etab[4 * q + 1]); };
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cnt(x) < cnt(y); });
        const int R = (fp.ne + 63) / 64;
        std::vector<int> base(R + 1, 0);
        for (int r = 0; r < R; r++) {
            const int last = std::min(fp.ne, 64 * (r + 1)) - 1;   // the round's largest count
            base[r + 1] = base[r] + std::max(1, cnt(ord[last]));
        }
        const int PS = std::max(PAR_BATCH, (base[R] + PAR_BATCH - 1) / PAR_BATCH * PAR_BATCH);
        fp.par_rounds = R;
        fp.par_psteps = PS;
        fp.parp.assign(size_t(PS) * 64, ZC | (ZC << 16));
        fp.parf.assign(size_t(std::max(R, 1)) * 64, 0xFFFFu << 16);
        for (int r = 0; r < R; r++) {
            for (int l = 0; l < 64; l++) {
                const int w = 64 * r + l;
                if (w < fp.ne) {
                    const int q = ord[w];
                    const int e = fp.etab[4 * q], db = fp.etab[4 * q + 1], de = fp.etab[4 * q + 2];
                    const int pv = fp.etab[4 * q + 3];
                    for (int t = 0; t < de - db; t++)
                        fp.parp[size_t(base[r] + t) * 64 + l] = (uint32_t(fp.dep[2 * (db + t)]) * 8u) |
                                                                ((uint32_t(fp.dep[2 * (db + t) + 1]) * 8u) << 16);
                    fp.parf[size_t(r) * 64 + l] = uint32_t(e) * 8u | 1u |
                                                  ((pv != 0xFFFF ? uint32_t(pv) * 8u : 0xFFFFu) << 16);
                }
                fp.parp[size_t(base[r + 1] - 1) * 64 + l] |= 1u;
            }
        }
    }
    if (fp.dep.empty()) { fp.dep.assign(2, 0); }
    if (fp.etab.empty()) { fp.etab.assign(4, 0); }

    // ISAI(L), ISAI(U) generation streams (power-1 pattern).  Row i of the
    // ISAI solves A(J,J)^T m = e by the column sweep of the reference
    // (isai_gpu.cu:325-389); in pull form every target m_t is an fma chain
    // acc = fma(-m_c, A(J_c, J_t), acc) over the already final m_c in the
    // sweep's order (L: c = |J|-1 down to t+1, U: c = 0 up to t-1), then
    // (U only) divided by A(J_t, J_t).  Lane l runs the targets of its rows
    // l + 64 k one after the other; records are LDS byte offsets for 8-byte
    // elements: rec0 = 8 q | 8 mp << 16 (q: factor position, mp: ISAI
    // position of m_c), rec1 = 8 p | 8 d << 16 (p: target written after this
    // step, d: divisor position; 0xFFFF = none).  Cells past nnz: factor
    // zero (nnz) and one (nnz + 1); ISAI zero (nnz), one (nnz + 1), minus one
    // (nnz + 2).
    const uint32_t FZ = uint32_t(nnz), FONE = uint32_t(nnz + 1);
    const uint32_t MZ = uint32_t(nnz), MONE = uint32_t(nnz + 1), MNEG = uint32_t(nnz + 2);
    auto posof = [&](int r, int c) -> int {
        const auto b = pa.ci.begin() + pa.rp[r], e = pa.ci.begin() + pa.rp[r + 1];
        const auto it = std::lower_bound(b, e, c);
        return (it != e && *it == c) ? int(it - pa.ci.begin()) : -1;
    };
    auto rec0 = [](uint32_t q, uint32_t mp) { return (q * 8) | ((mp * 8) << 16); };
    auto rec1 = [](uint32_t p, uint32_t d) { return (p == 0xFFFFu ? 0xFFFFu : p * 8) | ((d == 0xFFFFu ? 0xFFFFu : d * 8) << 16); };
    auto isai_stream = [&](bool lower, std::vector<uint32_t> &r0, std::vector<uint32_t> &r1, int &steps) {
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> L(64);
        for (int i = 0; i < n; i++) {
            auto &v = L[i % 64];
            if (lower) {
                const int b = pa.rp[i], m = pa.diag[i] - b + 1;
                for (int t = m - 2; t >= 0; t--) {
                    const int jt = pa.ci[b + t];
                    size_t first = v.size();
                    for (int c = m - 1; c > t; c--) {
                        const int q = posof(pa.ci[b + c], jt);
                        if (q < 0) continue;
                        v.push_back({rec0(uint32_t(q), c == m - 1 ? MONE : uint32_t(b + c)), rec1(0xFFFFu, 0xFFFFu)});
                    }
                    if (v.size() == first) v.push_back({rec0(FZ, MZ), 0});   // (no pair: cannot happen, c = m-1 exists)
                    v.back().second = rec1(uint32_t(b + t), 0xFFFFu);
                }
            } else {
                const int b = pa.diag[i], m = pa.rp[i + 1] - b;
                v.push_back({rec0(FONE, MNEG), rec1(uint32_t(b), uint32_t(pa.diag[i]))});
                for (int t = 1; t < m; t++) {
                    const int jt = pa.ci[b + t];
                    size_t first = v.size();
                    for (int c = 0; c < t; c++) {
                        const int q = posof(pa.ci[b + c], jt);
                        if (q < 0) continue;
                        v.push_back({rec0(uint32_t(q), uint32_t(b + c)), rec1(0xFFFFu, 0xFFFFu)});
                    }
                    if (v.size() == first) v.push_back({rec0(FZ, MZ), 0});
                    v.back().second = rec1(uint32_t(b + t), uint32_t(pa.diag[jt]));
                }
            }
        }
        size_t S = 1;
        for (auto &v : L) S = std::max(S, v.size());
        S = (S + ISAI_BATCH - 1) / ISAI_BATCH * ISAI_BATCH;   // whole record batches
        steps = int(S);
        r0.assign(S * 64, rec0(FZ, MZ));
        r1.assign(S * 64, rec1(0xFFFFu, 0xFFFFu));
        for (int l = 0; l < 64; l++)
            for (size_t k = 0; k < L[l].size(); k++) {
                r0[k * 64 + l] = L[l][k].first;
                r1[k * 64 + l] = L[l][k].second;
            }
    };
    isai_stream(true, fp.isL0, fp.isL1, fp.isL_steps);
    isai_stream(false, fp.isU0, fp.isU1, fp.isU_steps);

    // Exact ILU(0) as per-lane step streams.  Entry e = (i, j) with work
    // folds its dependency pairs acc = fma(-l_ik, u_kj, acc) from acc = a_ij
    // in k order (the reference KIJ order, merged.cu:168-219), one pair per
    // step, then (lower entries) divides by u_jj and writes the entry.  A
    // lane runs its entries one after another; entries are list-scheduled in
    // entry-level order onto the lane where they can start first, starting no
    // earlier than ILU_GAP steps after every value they read was written (a
    // wave's LDS accesses execute in program order, and the kernel loads a
    // step's operands one step ahead).  Records: r0 = 8 pL | 8 pU << 16 (the
    // pair; zero cell for entries without pairs), r1 = 8 pos | begin (bit 0)
    // | end (bit 1) | (8 pdiv or 0xFFFF) << 16.  Idle steps: zero cells, no
    // flags.
    {
        const uint32_t ZC = uint32_t(nnz) * 8u;
        std::vector<int> fin(nnz, -1000000);   // step writing each entry (entries without work: initial)
        std::vector<int> freel(64, 0);
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> L(64);
        const int ne = int(fp.etab.size() / 4);
        for (int q = 0; q < ne; q++) {
            const int e = fp.etab[4 * q], db = fp.etab[4 * q + 1], de = fp.etab[4 * q + 2], pv = fp.etab[4 * q + 3];
            int ready = 0;
            for (int d = db; d < de; d++)
                ready = std::max(ready, std::max(fin[fp.dep[2 * d]], fin[fp.dep[2 * d + 1]]) + ILU_GAP);
            if (pv != 0xFFFF) ready = std::max(ready, fin[pv] + ILU_GAP);
            const int k = std::max(1, de - db);
            int best = 0, bstart = 1 << 30, bidle = 1 << 30;
            for (int l = 0; l < 64; l++) {
                const int st = std::max(freel[l], ready), idle = st - freel[l];
                if (st < bstart || (st == bstart && idle < bidle)) { best = l; bstart = st; bidle = idle; }
            }
            auto &v = L[best];
            while ((int)v.size() < bstart) v.push_back({ZC | (ZC << 16), ZC | (0xFFFFu << 16)});
            for (int t = 0; t < k; t++) {
                const uint32_t r0 = (de > db) ? (uint32_t(fp.dep[2 * (db + t)]) * 8u) | ((uint32_t(fp.dep[2 * (db + t) + 1]) * 8u) << 16)
                                              : (ZC | (ZC << 16));
                uint32_t r1 = uint32_t(e) * 8u | (t == 0 ? 1u : 0u) | (t == k - 1 ? 2u : 0u);
                r1 |= (pv != 0xFFFF ? uint32_t(pv) * 8u : 0xFFFFu) << 16;
                v.push_back({r0, r1});
            }
            fin[e] = bstart + k - 1;
            freel[best] = bstart + k;
        }
        int S = 1;
        for (auto &v : L) S = std::max(S, (int)v.size());
        S = (S + ILU_BATCH - 1) / ILU_BATCH * ILU_BATCH;
        fp.ilu_steps = S;
        fp.ilu0.assign(size_t(S) * 64, ZC | (ZC << 16));
        fp.ilu1.assign(size_t(S) * 64, ZC | (0xFFFFu << 16));
        for (int l = 0; l < 64; l++)
            for (size_t t = 0; t < L[l].size(); t++) {
                fp.ilu0[t * 64 + l] = L[l][t].first;
                fp.ilu1[t * 64 + l] = L[l][t].second;
            }
    }
    fp.ok = true;
}

}  // namespace bilu

// Missing-diagonal insertion on the host: one pass over the rows, the new
// entry (i,i) going before the first column greater than i.
extern "C" int bilu_pattern_add_diagonal(int n, int nnz, const int *rp, const int *ci, int *n_missing, int *rp_out,
                                         int *ci_out, int *src_out)
{
    if (n <= 0 || nnz < 0 || !rp || (nnz > 0 && !ci) || !n_missing) return BILU_ERR_INVALID_ARG;
    if (rp_out && (!ci_out || !src_out)) return BILU_ERR_INVALID_ARG;
    if (rp[0] != 0 || rp[n] != nnz) return BILU_ERR_INVALID_PATTERN;
    int missing = 0;
    for (int i = 0; i < n; i++) {
        if (rp[i + 1] < rp[i]) return BILU_ERR_INVALID_PATTERN;
        bool has = false;
        for (int p = rp[i]; p < rp[i + 1]; p++) {
            if (ci[p] < 0 || ci[p] >= n || (p > rp[i] && ci[p] <= ci[p - 1])) return BILU_ERR_INVALID_PATTERN;
            has |= ci[p] == i;
        }
        missing += has ? 0 : 1;
    }
    *n_missing = missing;
    if (!rp_out) return BILU_OK;
    int q = 0;
    rp_out[0] = 0;
    for (int i = 0; i < n; i++) {
        bool placed = false;
        for (int p = rp[i]; p < rp[i + 1]; p++) {
            if (!placed && ci[p] >= i) {
                if (ci[p] > i) { ci_out[q] = i; src_out[q++] = -1; }
                placed = true;
            }
            ci_out[q] = ci[p];
            src_out[q++] = p;
        }
        if (!placed) { ci_out[q] = i; src_out[q++] = -1; }
        rp_out[i + 1] = q;
    }
    return BILU_OK;
}
