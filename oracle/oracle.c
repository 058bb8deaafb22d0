/*
 * oracle.c -- CPU restatement of the reference arithmetic (test
 * infrastructure only; see oracle.h for who may load it).
 *
 * Reference paths cited relative to
 * Batched-ILU0-merged/ILU0_ISAI_bicgstab/ of aggarwal2000/Merging-ILU0.
 * Short names: merged.cu = PrecondBiCGSTAB_exact_ilu0_app1_merged_gpu.cu,
 * common.hpp = Bicgstab_common.hpp.
 *
 * Build: oracle/build.py (gcc -O3 -march=x86-64-v3 -ffp-contract=off).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* position of column c in row r of a sorted CSR pattern, -1 if absent */
static int find_pos(const int *rp, const int *ci, int r, int c)
{
    int lo = rp[r], hi = rp[r + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (ci[mid] == c) return mid;
        if (ci[mid] < c) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* factorization.cu:192-212 (find_locn_of_diag_elements_kernel) and
 * :399-421 (Count_Missing_Diagonal_Elements). */
int orc_find_diag(int n, const int *rp, const int *ci, int *diag)
{
    int missing = 0;
    for (int i = 0; i < n; i++) {
        diag[i] = find_pos(rp, ci, i, i);
        if (diag[i] < 0) missing++;
    }
    return missing;
}

/* Missing-diagonal insertion, in the steps of factorization.cu:31-65
 * (diag_info[i] = 1 for a row without (i,i), then an inclusive prefix sum),
 * :70-78 (new row pointers = old + diag_info), :82-165 (per new position the
 * holder = old position, or -(i+1) for the inserted (i,i): entries left of
 * the diagonal keep their offset, entries right of it move one further, the
 * diagonal takes the first position whose column exceeds i, or the row's
 * end) and :169-189 (column = old column or i, value = old value or 0).
 * Returns the number of inserted entries; rp_out == NULL only counts.
 * The reference's warp code also marks lane 0 of every later 32-entry chunk
 * that lies right of the diagonal (factorization.cu:134-136), which in a row
 * of more than 32 entries overwrites a moved entry; this restatement keeps
 * the one diagonal per row that the rest of the routine assumes, so such
 * rows follow the brute-force union instead (pinned against it). */
int orc_add_diagonal(int n, const int *rp, const int *ci, int *rp_out, int *ci_out, int *holder)
{
    int *info = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) info[i] = find_pos(rp, ci, i, i) < 0 ? 1 : 0;
    for (int i = 1; i < n; i++) info[i] += info[i - 1];
    const int missing = info[n - 1];
    if (rp_out) {
        rp_out[0] = 0;
        for (int i = 0; i < n; i++) rp_out[i + 1] = rp[i + 1] + info[i];
        for (int i = 0; i < n; i++) {
            const int ob = rp[i], oe = rp[i + 1], nb = rp_out[i], ne = rp_out[i + 1], disp = nb - ob;
            if (ne - nb == oe - ob) {
                for (int p = ob; p < oe; p++) holder[p + disp] = p;
                continue;
            }
            int prev_flag = 0;
            for (int p = ob; p < oe; p++) {
                const int flag = ci[p] > i;
                if (ci[p] < i) holder[p + disp] = p;
                else holder[p + disp + 1] = p;
                if (flag && (p == ob || !prev_flag)) holder[p + disp] = -i - 1;
                if (p == oe - 1 && !flag) holder[p + disp + 1] = -i - 1;
                prev_flag = flag;
            }
            if (ob == oe) holder[nb] = -i - 1;
        }
        for (int k = 0; k < rp_out[n]; k++) ci_out[k] = holder[k] < 0 ? -(holder[k] + 1) : ci[holder[k]];
    }
    free(info);
    return missing;
}

void orc_expand_values(int batch, int nnz, int nnz_out, const int *holder, const double *vals, double *out)
{
    for (long long i = 0; i < (long long)batch * nnz_out; i++) {
        const long long s = i / nnz_out;
        const int h = holder[i % nnz_out];
        out[i] = h < 0 ? 0.0 : vals[h + s * nnz];
    }
}

/* Strided partial sums followed by a shfl_down tree, the order of the
 * reference inner_product (common.hpp:97-128: lane l of warp 0 sums
 * idx = l, l+32, ...; then offsets 16,8,4,2,1).  width 1 = serial sum
 * (common.hpp:82-95), width 64 = the wave64 kernels' order. */
double orc_dot(int n, const double *a, const double *b, int width)
{
    double t[64];
    if (width < 1) width = 1;
    for (int l = 0; l < width; l++) {
        double s = 0.0;
        for (int idx = l; idx < n; idx += width) s = fma(a[idx], b[idx], s);
        t[l] = s;
    }
    for (int off = width / 2; off > 0; off /= 2)
        for (int l = 0; l < off; l++) t[l] = t[l] + t[l + off];
    return t[0];
}

static double row_sum(const double *vals, const int *ci, int beg, int end,
                      const double *x, int width)
{
    double t[64];
    if (width <= 1) {
        double s = 0.0;
        for (int k = beg; k < end; k++) s = fma(vals[k], x[ci[k]], s);
        return s;
    }
    for (int l = 0; l < width; l++) {
        double s = 0.0;
        for (int k = beg + l; k < end; k += width) s = fma(vals[k], x[ci[k]], s);
        t[l] = s;
    }
    for (int off = width / 2; off > 0; off /= 2)
        for (int l = 0; l < off; l++) t[l] = t[l] + t[l + off];
    return t[0];
}

/* SpMV (common.hpp:254-290): width 32 = one warp per row with lane-strided
 * partial sums and a shfl_down tree; width 1 = left-to-right row sum. */
void orc_spmv(int n, const int *rp, const int *ci, const double *vals,
              const double *x, double *y, int width)
{
    for (int i = 0; i < n; i++) y[i] = row_sum(vals, ci, rp[i], rp[i + 1], x, width);
}

/* Exact ILU(0) in place on the pattern.
 * Reference: the right-looking KIJ loop of merged.cu:145-246
 * (compute_exact_ilu_0_approach1; twin ILU_0_gpu.cu:105-171): for k, for rows
 * i > k holding (i,k): l_ik = a_ik / a_kk, then a_ij -= l_ik * a_kj for
 * j > k.  Written here row by row (IKJ); every entry receives the same
 * sequence of fma updates in the same k order, so the result is bitwise the
 * KIJ result.  The reference also "updates" entries whose (k,j) is not in
 * the pattern with a zero; that is skipped here (it changes nothing but the
 * sign of zero).  Returns ORC_ZERO_PIVOT if a pivot or a U diagonal is 0
 * (the reference asserts, merged.cu:198). */
int orc_ilu0(int n, const int *rp, const int *ci, const int *diag, double *F)
{
    int *map = (int *)malloc(sizeof(int) * (size_t)n);
    int status = 0;
    for (int j = 0; j < n; j++) map[j] = -1;
    for (int i = 0; i < n && status == 0; i++) {
        for (int p = rp[i]; p < rp[i + 1]; p++) map[ci[p]] = p;
        for (int p = rp[i]; p < diag[i]; p++) {
            const int k = ci[p];
            const double ukk = F[diag[k]];
            if (ukk == 0.0) { status = ORC_ZERO_PIVOT; break; }
            const double l = F[p] / ukk;
            F[p] = l;
            for (int q = diag[k] + 1; q < rp[k + 1]; q++) {
                const int dst = map[ci[q]];
                if (dst >= 0) F[dst] = fma(-l, F[q], F[dst]);
            }
        }
        for (int p = rp[i]; p < rp[i + 1]; p++) map[ci[p]] = -1;
    }
    if (status == 0)
        for (int i = 0; i < n; i++)
            if (F[diag[i]] == 0.0) { status = ORC_ZERO_PIVOT; break; }
    free(map);
    return status;
}

/* ParILU(0) fixed-point sweeps, Jacobi form (every sweep reads the previous
 * sweep's factors).  Per entry (parILU_0_gpu.cu:18-57 with the dependency
 * lists of :118-220): sum over k < min(i,j) with (i,k),(k,j) in the pattern,
 * k increasing, of l_ik*u_kj; lower entries (i>j) become (a_ij - sum)/u_jj,
 * upper entries a_ij - sum.  Initial guess = A itself, which on the in-place
 * layout is the reference's L = strict-lower(A)+I, U = upper(A)
 * (factorization.cu:288-344).  The reference sweeps asynchronously in place
 * (results after finitely many sweeps depend on scheduling); Jacobi is the
 * deterministic variant with the same fixed point. */
int orc_parilu0_sweeps(int n, const int *rp, const int *ci, const int *diag,
                       const double *A, double *F, int sweeps)
{
    const int nnz = rp[n];
    double *old = (double *)malloc(sizeof(double) * (size_t)nnz);
    for (int s = 0; s < sweeps; s++) {
        memcpy(old, F, sizeof(double) * (size_t)nnz);
        for (int i = 0; i < n; i++) {
            for (int e = rp[i]; e < rp[i + 1]; e++) {
                const int j = ci[e];
                const int kmax = (i < j ? i : j) - 1;
                double sum = 0.0;
                for (int p = rp[i]; p < diag[i] && ci[p] <= kmax; p++) {
                    const int k = ci[p];
                    const int q = find_pos(rp, ci, k, j);
                    if (q >= 0) sum = fma(old[p], old[q], sum);
                }
                if (i > j) F[e] = (A[e] - sum) / old[diag[j]];
                else F[e] = A[e] - sum;
            }
        }
    }
    free(old);
    return 0;
}

/* ||(L*U - A) restricted to the pattern||_F with unit-diagonal L. */
double orc_factor_residual(int n, const int *rp, const int *ci, const int *diag,
                           const double *A, const double *F)
{
    double acc = 0.0;
    for (int i = 0; i < n; i++) {
        for (int e = rp[i]; e < rp[i + 1]; e++) {
            const int j = ci[e];
            const int kmax = (i < j ? i : j) - 1;
            double sum = 0.0;
            for (int p = rp[i]; p < diag[i] && ci[p] <= kmax; p++) {
                const int q = find_pos(rp, ci, ci[p], j);
                if (q >= 0) sum = fma(F[p], F[q], sum);
            }
            /* (LU)_ij = sum + (i > j ? l_ij * u_jj : u_ij) */
            const double lu = (i > j) ? fma(F[e], F[diag[j]], sum) : sum + F[e];
            const double d = lu - A[e];
            acc = fma(d, d, acc);
        }
    }
    return sqrt(acc);
}

/* v_hat = U^{-1} L^{-1} v on the in-place factor.
 * merged.cu:22-64: y_i = (v_i - sum_{j<i} l_ij y_j) / 1, the sum running
 * left to right; merged.cu:67-108: x_i = (y_i - sum_{j>i} u_ij x_j) / u_ii,
 * the sum running right to left. */
void orc_lu_solve(int n, const int *rp, const int *ci, const int *diag,
                  const double *F, const double *v, double *tmp, double *out)
{
    for (int i = 0; i < n; i++) {
        double sum = 0.0;
        for (int p = rp[i]; p < diag[i]; p++) sum = fma(F[p], tmp[ci[p]], sum);
        tmp[i] = v[i] - sum;
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = 0.0;
        for (int p = rp[i + 1] - 1; p > diag[i]; p--) sum = fma(F[p], out[ci[p]], sum);
        out[i] = (tmp[i] - sum) / F[diag[i]];
    }
}

/* Dependency levels of the triangular solves: lower level(i) = 1 + max
 * level(j) over strictly-lower columns j (0 if none); upper from the bottom
 * over strictly-upper columns. */
void orc_trsv_levels(int n, const int *rp, const int *ci, const int *diag, int *lev_l, int *lev_u)
{
    for (int i = 0; i < n; i++) {
        int l = 0;
        for (int p = rp[i]; p < diag[i]; p++) if (lev_l[ci[p]] + 1 > l) l = lev_l[ci[p]] + 1;
        lev_l[i] = l;
    }
    for (int i = n - 1; i >= 0; i--) {
        int l = 0;
        for (int p = diag[i] + 1; p < rp[i + 1]; p++) if (lev_u[ci[p]] + 1 > l) l = lev_u[ci[p]] + 1;
        lev_u[i] = l;
    }
}

/* Same solves as orc_lu_solve in the persistent wave64 kernel's order: each
 * row sum adds rounded products (no fma) in "availability" order, the row's
 * entries sorted by (level of the column, column) -- the order in which the
 * kernel adds entries into the row's accumulator as each level of rows
 * becomes final -- and the upper solve multiplies by the reciprocal of u_ii
 * (computed once per factor) instead of dividing. */
static void sort_avail(int *pairs, int m, const int *lev)
{
    /* insertion sort of (col, pos) pairs keyed on (lev[col], col) */
    for (int a = 1; a < m; a++) {
        const int c = pairs[2 * a], p = pairs[2 * a + 1];
        int b = a - 1;
        while (b >= 0 && (lev[pairs[2 * b]] > lev[c] || (lev[pairs[2 * b]] == lev[c] && pairs[2 * b] > c))) {
            pairs[2 * b + 2] = pairs[2 * b];
            pairs[2 * b + 3] = pairs[2 * b + 1];
            b--;
        }
        pairs[2 * b + 2] = c;
        pairs[2 * b + 3] = p;
    }
}

void orc_lu_solve_avail(int n, const int *rp, const int *ci, const int *diag,
                        const int *lev_l, const int *lev_u, const double *F,
                        const double *v, double *tmp, double *out)
{
    int pairs[2 * 1024];
    for (int i = 0; i < n; i++) {
        int m = 0;
        for (int p = rp[i]; p < diag[i] && m < 1024; p++, m++) { pairs[2 * m] = ci[p]; pairs[2 * m + 1] = p; }
        sort_avail(pairs, m, lev_l);
        double sum = 0.0;
        for (int q = 0; q < m; q++) sum = sum + F[pairs[2 * q + 1]] * tmp[pairs[2 * q]];
        tmp[i] = v[i] - sum;
    }
    for (int i = n - 1; i >= 0; i--) {
        int m = 0;
        for (int p = diag[i] + 1; p < rp[i + 1] && m < 1024; p++, m++) { pairs[2 * m] = ci[p]; pairs[2 * m + 1] = p; }
        sort_avail(pairs, m, lev_u);
        double sum = 0.0;
        for (int q = 0; q < m; q++) sum = sum + F[pairs[2 * q + 1]] * out[pairs[2 * q]];
        out[i] = (tmp[i] - sum) * (1.0 / F[diag[i]]);
    }
}

/* value of the typed view of `vals` at (r,c): lower = strict-lower part with
 * an explicit unit diagonal (the L of factorization.cu:313-314), upper =
 * diagonal + strict upper, general = everything. */
static double view_at(const int *rp, const int *ci, const double *vals,
                      int type, int r, int c)
{
    if (type == ORC_ISAI_LOWER) {
        if (r == c) return 1.0;
        if (r < c) return 0.0;
    } else if (type == ORC_ISAI_UPPER) {
        if (r > c) return 0.0;
    }
    const int p = find_pos(rp, ci, r, c);
    return p >= 0 ? vals[p] : 0.0;
}

/* Largest ISAI row: 32 for the general (Gauss-Jordan) solve, as the
 * reference's row_size_limit (isai_gpu.cu:17,223); the triangular sweeps take
 * rows up to 128 entries (power-k patterns, where the reference stops on its
 * assert). */
#define ORC_ISAI_GEN_MAX 32
#define ORC_ISAI_TRI_MAX 128

/* Reference pivot choice (isai_gpu.cu:417-451, choose_pivot_row_1): 32
 * lanes; lane t outside [dp, m) stands in with the diagonal candidate; a
 * shfl_down tree keeps the other lane's candidate only if strictly larger in
 * magnitude. */
static int ref_pivot(double Dt[][ORC_ISAI_TRI_MAX], int m, int dp)
{
    double val[32];
    int piv[32];
    for (int t = 0; t < 32; t++) {
        if (t >= m || t < dp) { val[t] = Dt[dp][dp]; piv[t] = dp; }
        else { val[t] = Dt[t][dp]; piv[t] = t; }
    }
    for (int off = 16; off > 0; off /= 2)
        for (int l = 0; l + off < 32; l++)
            if (fabs(val[l + off]) > fabs(val[l])) { val[l] = val[l + off]; piv[l] = piv[l + off]; }
    return piv[0];
}

/* ISAI with the sparsity pattern of the input (power 1).
 * isai_gpu.cu:559-626: for row i with J = cols(i), the thread holding column
 * t of the dense block A(J,J) ("implicit transpose") solves A(J,J)^T m = e,
 * e at the position of i in J; row i of the ISAI is m^T.
 * lower: upper-triangular column sweep, rhs at the last position
 * (isai_gpu.cu:325-355); upper: lower-triangular sweep, rhs at 0
 * (:359-389); general: Gauss-Jordan with partial pivoting (:417-554).
 * Output on the pattern, masked to the triangle of `type`. */
/* one ISAI row: solve Dt^T-form system for `sol` (Dt[t][k] = A(J[k], J[t])),
 * rhs e_ipos; lower: upper-triangular column sweep from the last position
 * (isai_gpu.cu:325-355), upper: lower-triangular sweep from 0 (:359-389),
 * general: Gauss-Jordan with the reference pivot choice (:417-554). */
static int isai_row_solve(double Dt[][ORC_ISAI_TRI_MAX], int m, int type, int ipos, double *sol)
{
    if (type == ORC_ISAI_LOWER) {
        for (int t = 0; t < m; t++) sol[t] = (t == m - 1) ? 1.0 : 0.0;
        for (int c = m - 1; c >= 0; c--) {
            if (Dt[c][c] == 0.0) return ORC_ZERO_PIVOT;
            sol[c] = sol[c] / Dt[c][c];
            for (int t = 0; t < c; t++) sol[t] = fma(-sol[c], Dt[t][c], sol[t]);
        }
    } else if (type == ORC_ISAI_UPPER) {
        for (int t = 0; t < m; t++) sol[t] = (t == 0) ? 1.0 : 0.0;
        for (int c = 0; c < m; c++) {
            if (Dt[c][c] == 0.0) return ORC_ZERO_PIVOT;
            sol[c] = sol[c] / Dt[c][c];
            for (int t = c + 1; t < m; t++) sol[t] = fma(-sol[c], Dt[t][c], sol[t]);
        }
    } else {
        for (int t = 0; t < m; t++) sol[t] = (t == ipos) ? 1.0 : 0.0;
        for (int dp = 0; dp < m; dp++) {
            const int pv = ref_pivot(Dt, m, dp);
            if (pv != dp) {
                for (int c = 0; c < m; c++) { double tmp = Dt[dp][c]; Dt[dp][c] = Dt[pv][c]; Dt[pv][c] = tmp; }
                double tmp = sol[dp]; sol[dp] = sol[pv]; sol[pv] = tmp;
            }
            const double de = Dt[dp][dp];
            if (de == 0.0) return ORC_ZERO_PIVOT;
            for (int t = 0; t < m; t++) {
                if (t == dp) continue;
                const double mult = Dt[t][dp] / de;
                for (int c = 0; c < m; c++) Dt[t][c] = fma(-mult, Dt[dp][c], Dt[t][c]);
                sol[t] = fma(-mult, sol[dp], sol[t]);
            }
        }
        for (int t = 0; t < m; t++) sol[t] = sol[t] / Dt[t][t];
    }
    return 0;
}

/* ISAI on the rows J_i = mci[mrp[i] .. mrp[i+1]) of a chosen pattern:
 * Dt[t][k] = A(J[k], J[t]) from the typed view of (rp, ci, vals), row i of
 * the ISAI = the solution, written to out[mrp[i] + q]. */
static int isai_on_rows(int n, const int *rp, const int *ci, const double *vals, int type,
                        const int *mrp, const int *mci, double *out)
{
    static __thread double Dt[ORC_ISAI_TRI_MAX][ORC_ISAI_TRI_MAX];
    double sol[ORC_ISAI_TRI_MAX];
    for (int i = 0; i < n; i++) {
        const int beg = mrp[i], m = mrp[i + 1] - beg;
        if (m > ORC_ISAI_TRI_MAX || (type == ORC_ISAI_GENERAL && m > ORC_ISAI_GEN_MAX)) return -4;
        int ipos = 0;
        for (int q = 0; q < m; q++) if (mci[beg + q] == i) ipos = q;
        for (int k = 0; k < m; k++)
            for (int t = 0; t < m; t++) Dt[t][k] = view_at(rp, ci, vals, type, mci[beg + k], mci[beg + t]);
        const int st = isai_row_solve(Dt, m, type, ipos, sol);
        if (st != 0) return st;
        for (int q = 0; q < m; q++) out[beg + q] = sol[q];
    }
    return 0;
}

/* ISAI with the sparsity pattern of the input (power 1).
 * isai_gpu.cu:559-626: for row i with J = cols(i), the thread holding column
 * t of the dense block A(J,J) ("implicit transpose") solves A(J,J)^T m = e,
 * e at the position of i in J; row i of the ISAI is m^T.
 * Output on the pattern, masked to the triangle of `type` (rows start at
 * rp[i] or diag[i], so the typed rows are windows of the pattern). */
int orc_isai(int n, const int *rp, const int *ci, const int *diag,
             const double *vals, int type, double *out)
{
    static __thread double Dt[ORC_ISAI_TRI_MAX][ORC_ISAI_TRI_MAX];
    double sol[ORC_ISAI_TRI_MAX];
    for (int i = 0; i < n; i++) {
        int beg, end;
        if (type == ORC_ISAI_LOWER) { beg = rp[i]; end = diag[i] + 1; }
        else if (type == ORC_ISAI_UPPER) { beg = diag[i]; end = rp[i + 1]; }
        else { beg = rp[i]; end = rp[i + 1]; }
        const int m = end - beg;
        if (m > ORC_ISAI_TRI_MAX || (type == ORC_ISAI_GENERAL && m > ORC_ISAI_GEN_MAX)) return -4;
        int ipos = 0;
        for (int q = 0; q < m; q++) if (ci[beg + q] == i) ipos = q;
        /* Dt[t][k] = D[k][t] = A(J[k], J[t]) */
        for (int k = 0; k < m; k++)
            for (int t = 0; t < m; t++) Dt[t][k] = view_at(rp, ci, vals, type, ci[beg + k], ci[beg + t]);
        const int st = isai_row_solve(Dt, m, type, ipos, sol);
        if (st != 0) return st;
        for (int q = 0; q < m; q++) out[beg + q] = sol[q];
    }
    return 0;
}

/* Pattern of |P|^power for the typed view P of (rp, ci): LOWER = columns
 * <= i plus the (unit) diagonal, UPPER = columns >= i, GENERAL = all.  The
 * reference chooses the ISAI pattern as the nonzeros of A_0^power computed
 * from the first page's values (isai_gpu.cu:40-178); this is the structural
 * version (no entry lost to an exact numeric cancellation).  Row i of P^k =
 * union over j in row i of P^(k-1) of row j of P, sorted.  Returns nnz;
 * mrp == NULL only counts. */
int orc_isai_pattern(int n, const int *rp, const int *ci, int type, int power, int *mrp, int *mci)
{
    int *brp = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *bci = (int *)malloc(sizeof(int) * (size_t)(rp[n] + n));
    brp[0] = 0;
    for (int i = 0; i < n; i++) {
        int q = brp[i];
        for (int p = rp[i]; p < rp[i + 1]; p++) {
            const int c = ci[p];
            if ((type == ORC_ISAI_LOWER && c < i) || (type == ORC_ISAI_UPPER && c >= i) ||
                type == ORC_ISAI_GENERAL)
                bci[q++] = c;
        }
        if (type == ORC_ISAI_LOWER) bci[q++] = i;
        brp[i + 1] = q;
    }
    int *cur_rp = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int cap = brp[n] > 0 ? brp[n] : 1;
    int *cur_ci = (int *)malloc(sizeof(int) * (size_t)cap);
    memcpy(cur_rp, brp, sizeof(int) * (size_t)(n + 1));
    memcpy(cur_ci, bci, sizeof(int) * (size_t)brp[n]);
    char *mark = (char *)calloc((size_t)n, 1);
    int *row = (int *)malloc(sizeof(int) * (size_t)n);
    for (int k = 2; k <= power; k++) {
        int *nrp = (int *)malloc(sizeof(int) * (size_t)(n + 1));
        int ncap = cur_rp[n] + n, nq = 0;
        int *nci = (int *)malloc(sizeof(int) * (size_t)ncap);
        nrp[0] = 0;
        for (int i = 0; i < n; i++) {
            int len = 0;
            for (int p = cur_rp[i]; p < cur_rp[i + 1]; p++) {
                const int j = cur_ci[p];
                for (int q = brp[j]; q < brp[j + 1]; q++)
                    if (!mark[bci[q]]) { mark[bci[q]] = 1; row[len++] = bci[q]; }
            }
            for (int c = 0, w = 0; c < n && w < len; c++)   /* sorted by scanning the marks */
                if (mark[c]) { mark[c] = 0; row[w++] = c; }
            if (nq + len > ncap) { ncap = 2 * (nq + len); nci = (int *)realloc(nci, sizeof(int) * (size_t)ncap); }
            memcpy(nci + nq, row, sizeof(int) * (size_t)len);
            nq += len;
            nrp[i + 1] = nq;
        }
        free(cur_rp); free(cur_ci);
        cur_rp = nrp; cur_ci = nci;
    }
    const int nnz = cur_rp[n];
    if (mrp) {
        memcpy(mrp, cur_rp, sizeof(int) * (size_t)(n + 1));
        memcpy(mci, cur_ci, sizeof(int) * (size_t)nnz);
    }
    free(brp); free(bci); free(cur_rp); free(cur_ci); free(mark); free(row);
    return nnz;
}

/* ISAI of the typed view of (rp, ci, vals) on the pattern (mrp, mci), as
 * orc_isai row by row; out[mrp[n]]. */
int orc_isai_on(int n, const int *rp, const int *ci, const double *vals, int type,
                const int *mrp, const int *mci, double *out)
{
    return isai_on_rows(n, rp, ci, vals, type, mrp, mci, out);
}

/* patterns of the ISAI arrays when they are not the matrix's own (powers > 1) */
typedef struct {
    const int *lrp, *lci, *urp, *uci;
} isai_pats;

/* y = M^{-1} v for the chosen preconditioner */
static void apply_pc(int n, const int *rp, const int *ci, const int *diag,
                     int kind, const double *P1, const double *P2, int spmv_w,
                     const int *lev, const isai_pats *pats, const double *v, double *tmp, double *out)
{
    if (kind == ORC_PC_ISAI_LU_PAT) {
        /* aiL then aiU on their own patterns, rows left to right */
        for (int i = 0; i < n; i++) tmp[i] = row_sum(P1, pats->lci, pats->lrp[i], pats->lrp[i + 1], v, spmv_w);
        for (int i = 0; i < n; i++) out[i] = row_sum(P2, pats->uci, pats->urp[i], pats->urp[i + 1], tmp, spmv_w);
    } else if (kind == ORC_PC_ISAI_GEN_PAT) {
        for (int i = 0; i < n; i++) out[i] = row_sum(P1, pats->lci, pats->lrp[i], pats->lrp[i + 1], v, spmv_w);
    } else if (kind == ORC_PC_LU && lev) {
        orc_lu_solve_avail(n, rp, ci, diag, lev, lev + n, P1, v, tmp, out);
    } else if (kind == ORC_PC_LU) {
        orc_lu_solve(n, rp, ci, diag, P1, v, tmp, out);
    } else if (kind == ORC_PC_ISAI_LU) {
        /* ilu_isai_solve.cu:25-45: temp = aiL * v; out = aiU * temp */
        for (int i = 0; i < n; i++) tmp[i] = row_sum(P1, ci, rp[i], diag[i] + 1, v, spmv_w);
        for (int i = 0; i < n; i++) out[i] = row_sum(P2, ci, diag[i], rp[i + 1], tmp, spmv_w);
    } else if (kind == ORC_PC_ISAI_GEN) {
        orc_spmv(n, rp, ci, P1, v, out, spmv_w);  /* gisai_solve.cu:23-30 */
    } else {
        memcpy(out, v, sizeof(double) * (size_t)n);
    }
}

/* Right-preconditioned BiCGSTAB with the reference control flow
 * (merged.cu:296-524; common.hpp:151-248): x0 = 0, r = b - A x0, r_hat = r,
 * p = v = 0, rho_old = omega_old = alpha = 1; iteration counted in half
 * steps (float); convergence on the recursive residual norm < tol_abs
 * after each half step; rho == 0 => break with flag -1; b == 0 => x = 0,
 * iter 0, flag 1; ||r0|| < tol_abs => iter 0, flag 1.
 * Additions (documented options): relative tolerance mode, a restart from
 * the true residual when it misses the tolerance, and non-finite detection
 * (flag -3) in place of the reference's hang/garbage.
 * true_res = ||b - A x||_2 (absolute, as common.hpp:294-342). */
static int bicgstab_core(int n, const int *rp, const int *ci, const int *diag,
                         const double *A, const double *b, double *x,
                         int pc_kind, const double *P1, const double *P2, const isai_pats *pats,
                         const orc_opts *o, float *iter_out, int *conv_out,
                         double *iter_res_out, double *true_res_out);

int orc_bicgstab(int n, const int *rp, const int *ci, const int *diag,
                 const double *A, const double *b, double *x,
                 int pc_kind, const double *P1, const double *P2,
                 const orc_opts *o, float *iter_out, int *conv_out,
                 double *iter_res_out, double *true_res_out)
{
    return bicgstab_core(n, rp, ci, diag, A, b, x, pc_kind, P1, P2, NULL, o, iter_out, conv_out,
                         iter_res_out, true_res_out);
}

static int bicgstab_core(int n, const int *rp, const int *ci, const int *diag,
                         const double *A, const double *b, double *x,
                         int pc_kind, const double *P1, const double *P2, const isai_pats *pats,
                         const orc_opts *o, float *iter_out, int *conv_out,
                         double *iter_res_out, double *true_res_out)
{
    const int dw = o->dot_width, sw = o->spmv_width;
    double *buf = (double *)malloc(sizeof(double) * (size_t)n * 10);
    int *lev = NULL;
    if (o->trsv_order == 1 && pc_kind == ORC_PC_LU) {
        lev = (int *)malloc(sizeof(int) * 2 * (size_t)n);
        orc_trsv_levels(n, rp, ci, diag, lev, lev + n);
    }
    double *r = buf, *rh = buf + n, *p = buf + 2 * n, *v = buf + 3 * n, *s = buf + 4 * n;
    double *t = buf + 5 * n, *ph = buf + 6 * n, *sh = buf + 7 * n, *tmp = buf + 8 * n, *w = buf + 9 * n;
    float iter = 0.0f;
    int conv = ORC_NOT_CONVERGED;
    double iter_res;

    for (int i = 0; i < n; i++) { x[i] = 0.0; v[i] = 0.0; p[i] = 0.0; r[i] = b[i]; rh[i] = b[i]; }
    const double res0 = sqrt(orc_dot(n, r, r, dw));
    const double bnorm = sqrt(orc_dot(n, b, b, dw));
    const double tol_abs = (o->tol_mode == ORC_TOL_RELATIVE) ? o->tol * bnorm : o->tol;
    iter_res = res0;

    /* a zero on U's diagonal is a zero pivot, whatever produced the factors
     * (ParILU sweeps, the caller): x = 0, as after a failed factorization */
    int bad = 0;
    if (pc_kind == ORC_PC_LU)
        for (int i = 0; i < n; i++) bad |= (P1[diag[i]] == 0.0);
    if (bad) {
        conv = ORC_ZERO_PIVOT; iter_res = 0.0;
    } else if (!isfinite(bnorm)) {
        conv = ORC_NONFINITE;
    } else if (bnorm == 0.0) {
        conv = ORC_CONVERGED; iter_res = 0.0;
    } else if (res0 < tol_abs) {
        conv = ORC_CONVERGED;
    } else {
        double rho_old = 1.0, omega_old = 1.0, alpha = 1.0;
        int restarts = 0;
        while (iter < (float)o->max_iter) {
            const double rho_new = orc_dot(n, r, rh, dw);
            if (rho_new == 0.0) break;
            const double beta = (rho_new / rho_old) * (alpha / omega_old);
            for (int i = 0; i < n; i++) p[i] = fma(beta, fma(-omega_old, v[i], p[i]), r[i]);
            apply_pc(n, rp, ci, diag, pc_kind, P1, P2, sw, lev, pats, p, tmp, ph);
            orc_spmv(n, rp, ci, A, ph, v, sw);
            alpha = rho_new / orc_dot(n, rh, v, dw);
            for (int i = 0; i < n; i++) s[i] = fma(-alpha, v[i], r[i]);
            iter_res = sqrt(orc_dot(n, s, s, dw));
            iter += 0.5f;
            if (!isfinite(iter_res)) { conv = ORC_NONFINITE; break; }
            int half_conv = iter_res < tol_abs;
            double omega_new = 0.0;
            if (half_conv) {
                for (int i = 0; i < n; i++) x[i] = fma(alpha, ph[i], x[i]);
            } else {
                apply_pc(n, rp, ci, diag, pc_kind, P1, P2, sw, lev, pats, s, tmp, sh);
                orc_spmv(n, rp, ci, A, sh, t, sw);
                const double ts = orc_dot(n, t, s, dw);
                const double tt = orc_dot(n, t, t, dw);
                omega_new = ts / tt;
                for (int i = 0; i < n; i++) x[i] = fma(omega_new, sh[i], fma(alpha, ph[i], x[i]));
                iter += 0.5f;
                for (int i = 0; i < n; i++) r[i] = fma(-omega_new, t[i], s[i]);
                iter_res = sqrt(orc_dot(n, r, r, dw));
                rho_old = rho_new;
                omega_old = omega_new;
                if (!isfinite(iter_res)) { conv = ORC_NONFINITE; break; }
            }
            if (half_conv || iter_res < tol_abs) {
                conv = ORC_CONVERGED;
                if (!o->true_res_restart || restarts >= o->max_restarts) break;
                /* true-residual check: r_true = b - A x */
                orc_spmv(n, rp, ci, A, x, w, sw);
                for (int i = 0; i < n; i++) w[i] = b[i] - w[i];
                const double tr = sqrt(orc_dot(n, w, w, dw));
                if (tr < tol_abs) break;
                restarts++;
                conv = ORC_NOT_CONVERGED;
                for (int i = 0; i < n; i++) { r[i] = w[i]; rh[i] = w[i]; p[i] = 0.0; v[i] = 0.0; }
                rho_old = 1.0; omega_old = 1.0; alpha = 1.0;
                iter_res = tr;
            }
        }
    }
    /* fused epilogue: absolute true residual */
    orc_spmv(n, rp, ci, A, x, w, sw);
    for (int i = 0; i < n; i++) w[i] = b[i] - w[i];
    *true_res_out = sqrt(orc_dot(n, w, w, dw));
    *iter_out = iter;
    *conv_out = conv;
    *iter_res_out = iter_res;
    free(buf);
    free(lev);
    return 0;
}

/* ---- batch driver (also the timed CPU baseline) ---- */
typedef struct {
    int n, nnz, batch_beg, batch_end, precond, sweeps;
    const int *rp, *ci, *diag;
    const double *A, *b;
    double *x;
    const orc_opts *o;
    float *iter; int *conv; double *iter_res, *true_res;
} batch_job;

static void *batch_worker(void *arg)
{
    batch_job *j = (batch_job *)arg;
    const int n = j->n, nnz = j->nnz;
    double *F = (double *)malloc(sizeof(double) * (size_t)nnz);
    double *P2 = (double *)malloc(sizeof(double) * (size_t)nnz);
    double *P1 = (double *)malloc(sizeof(double) * (size_t)nnz);
    for (int s = j->batch_beg; s < j->batch_end; s++) {
        const double *A = j->A + (size_t)s * nnz;
        const double *b = j->b + (size_t)s * n;
        double *x = j->x + (size_t)s * n;
        int st = 0, kind = ORC_PC_LU;
        const double *p1 = F, *p2 = NULL;
        memcpy(F, A, sizeof(double) * (size_t)nnz);
        switch (j->precond) {
        case 0: st = orc_ilu0(n, j->rp, j->ci, j->diag, F); break;
        case 1: st = orc_parilu0_sweeps(n, j->rp, j->ci, j->diag, A, F, j->sweeps); break;
        case 2: case 3:
            st = (j->precond == 2) ? orc_ilu0(n, j->rp, j->ci, j->diag, F)
                                   : orc_parilu0_sweeps(n, j->rp, j->ci, j->diag, A, F, j->sweeps);
            if (st == 0) st = orc_isai(n, j->rp, j->ci, j->diag, F, ORC_ISAI_LOWER, P1);
            if (st == 0) st = orc_isai(n, j->rp, j->ci, j->diag, F, ORC_ISAI_UPPER, P2);
            kind = ORC_PC_ISAI_LU; p1 = P1; p2 = P2;
            break;
        case 4:
            st = orc_isai(n, j->rp, j->ci, j->diag, A, ORC_ISAI_GENERAL, P1);
            kind = ORC_PC_ISAI_GEN; p1 = P1;
            break;
        default: kind = ORC_PC_NONE; break;
        }
        if (st != 0) {
            for (int i = 0; i < n; i++) x[i] = 0.0;
            j->iter[s] = 0.0f; j->conv[s] = st; j->iter_res[s] = 0.0;
            j->true_res[s] = sqrt(orc_dot(n, b, b, j->o->dot_width));
            continue;
        }
        orc_bicgstab(n, j->rp, j->ci, j->diag, A, b, x, kind, p1, p2, j->o,
                     &j->iter[s], &j->conv[s], &j->iter_res[s], &j->true_res[s]);
    }
    free(F); free(P1); free(P2);
    return NULL;
}

int orc_batch_solve(int n, int nnz, const int *rp, const int *ci, int batch,
                    const double *A, const double *b, double *x,
                    int precond, int sweeps, const orc_opts *o,
                    float *iter, int *conv, double *iter_res, double *true_res,
                    int threads)
{
    int *diag = (int *)malloc(sizeof(int) * (size_t)n);
    if (orc_find_diag(n, rp, ci, diag) > 0) { free(diag); return -5; }
    if (threads < 1) threads = 1;
    if (threads > batch) threads = batch > 0 ? batch : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    batch_job *jobs = (batch_job *)malloc(sizeof(batch_job) * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        batch_job *j = &jobs[t];
        j->n = n; j->nnz = nnz; j->precond = precond; j->sweeps = sweeps;
        j->batch_beg = (int)((long long)batch * t / threads);
        j->batch_end = (int)((long long)batch * (t + 1) / threads);
        j->rp = rp; j->ci = ci; j->diag = diag; j->A = A; j->b = b; j->x = x; j->o = o;
        j->iter = iter; j->conv = conv; j->iter_res = iter_res; j->true_res = true_res;
        if (threads == 1) batch_worker(j);
        else pthread_create(&th[t], NULL, batch_worker, j);
    }
    if (threads > 1)
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th); free(jobs); free(diag);
    return 0;
}

/* ---------------------------------------------------------------------
 * Mixed precision (FP32 inner BiCGSTAB inside FP64 iterative refinement)
 * and ParILU(0) with a warm start and a residual stop.  These paths have no
 * counterpart in the reference; they restate the build's own persistent
 * kernels operation for operation, on the reference's building blocks
 * (BiCGSTAB control flow of merged.cu:296-524, ParILU formula of
 * parILU_0_gpu.cu:18-57).
 * ------------------------------------------------------------------- */

int orc_bicgstab_mixed_pc(int n, const int *rp, const int *ci, const int *diag,
                          const int *lev_l, const int *lev_u, const double *A, const double *b,
                          const float *F32, int isai, int zp, const orc_opts *o, double inner_tol,
                          int max_refine, double *x, float *iter_out, int *conv_out,
                          double *iter_res_out, double *true_res_out, int *refine_out);

/* FP32 dot in the wave64 order (lane-strided fmaf partial sums, then the
 * pairwise tree with offsets 32..1) */
static float dotf64w(int n, const float *a, const float *b)
{
    float t[64];
    for (int l = 0; l < 64; l++) {
        float s = 0.0f;
        for (int idx = l; idx < n; idx += 64) s = fmaf(a[idx], b[idx], s);
        t[l] = s;
    }
    for (int off = 32; off > 0; off /= 2)
        for (int l = 0; l < off; l++) t[l] = t[l] + t[l + off];
    return t[0];
}

/* FP32 SpMV, each row a left-to-right fmaf chain */
static void spmvf(int n, const int *rp, const int *ci, const float *vals, const float *x, float *y)
{
    for (int i = 0; i < n; i++) {
        float s = 0.0f;
        for (int p = rp[i]; p < rp[i + 1]; p++) s = fmaf(vals[p], x[ci[p]], s);
        y[i] = s;
    }
}

/* FP32 twin of orc_lu_solve_avail (availability order, reciprocal of
 * u_ii): each product is formed in FP32 and added into an FP64 row sum,
 * which is rounded to FP32 once before the row is finalised (the kernel's
 * accumulators are FP64 LDS cells). */
static void lu_solve_avail_f(int n, const int *rp, const int *ci, const int *diag,
                             const int *lev_l, const int *lev_u, const float *F,
                             const float *v, float *tmp, float *out)
{
    int pairs[2 * 1024];
    for (int i = 0; i < n; i++) {
        int m = 0;
        for (int p = rp[i]; p < diag[i] && m < 1024; p++, m++) { pairs[2 * m] = ci[p]; pairs[2 * m + 1] = p; }
        sort_avail(pairs, m, lev_l);
        double sum = 0.0;
        for (int q = 0; q < m; q++) sum = sum + (double)(F[pairs[2 * q + 1]] * tmp[pairs[2 * q]]);
        tmp[i] = v[i] - (float)sum;
    }
    for (int i = n - 1; i >= 0; i--) {
        int m = 0;
        for (int p = diag[i] + 1; p < rp[i + 1] && m < 1024; p++, m++) { pairs[2 * m] = ci[p]; pairs[2 * m + 1] = p; }
        sort_avail(pairs, m, lev_u);
        double sum = 0.0;
        for (int q = 0; q < m; q++) sum = sum + (double)(F[pairs[2 * q + 1]] * out[pairs[2 * q]]);
        out[i] = (tmp[i] - (float)sum) * (1.0f / F[diag[i]]);
    }
}

/* ISAI(L) and ISAI(U) of an in-place factor F (power-1 patterns) for the
 * FP32 preconditioner: the column sweeps of orc_isai (isai_gpu.cu:325-389)
 * over the factor rounded to FP32, in FP64 arithmetic, with every finished
 * m_c rounded to FP32 before it is used, and updates only from entries of
 * the pattern.  M[rp[i] .. diag[i]) = ISAI(L) (its unit
 * diagonal implicit), M[diag[i] .. rp[i+1]) = ISAI(U).  Returns
 * ORC_ZERO_PIVOT if a U diagonal is 0. */
int orc_isai_lu_f32(int n, const int *rp, const int *ci, const int *diag, const double *F, float *M)
{
    double sol[64];
    int bad = 0;
    for (int i = 0; i < n; i++) {
        /* lower: J = cols of row i up to the diagonal, rhs 1 at the last position */
        int b = rp[i], m = diag[i] - rp[i] + 1;
        if (m > 64) return -4;
        for (int t = 0; t < m; t++) sol[t] = (t == m - 1) ? 1.0 : 0.0;
        for (int c = m - 1; c >= 0; c--) {
            sol[c] = (double)(float)sol[c];   /* L(J_c, J_c) = 1 */
            for (int t = c - 1; t >= 0; t--) {
                const int q = find_pos(rp, ci, ci[b + c], ci[b + t]);
                if (q >= 0) sol[t] = fma(-sol[c], (double)(float)F[q], sol[t]);
            }
        }
        for (int t = 0; t < m - 1; t++) M[b + t] = (float)sol[t];
        /* upper: J = cols of row i from the diagonal, rhs 1 at position 0 */
        b = diag[i]; m = rp[i + 1] - diag[i];
        if (m > 64) return -4;
        for (int t = 0; t < m; t++) sol[t] = (t == 0) ? 1.0 : 0.0;
        for (int c = 0; c < m; c++) {
            const double dv = (double)(float)F[diag[ci[b + c]]];
            if (dv == 0.0) bad = 1;
            sol[c] = (double)(float)(sol[c] / dv);
            for (int t = c + 1; t < m; t++) {
                const int q = find_pos(rp, ci, ci[b + c], ci[b + t]);
                if (q >= 0) sol[t] = fma(-sol[c], (double)(float)F[q], sol[t]);
            }
        }
        for (int t = 0; t < m; t++) M[b + t] = (float)sol[t];
    }
    return bad ? ORC_ZERO_PIVOT : 0;
}

/* FP32 ISAI apply: out = ISAI(U) (ISAI(L) v), each row a CSR-order fmaf
 * chain over the triangle's own entries (ISAI(L): the strictly lower
 * entries and the unit diagonal; ISAI(U): the diagonal and the strictly
 * upper entries) -- the kernel's triangle slot layouts */
static void isai_apply_f(int n, const int *rp, const int *ci, const int *diag, const float *M,
                         const float *v, float *tmp, float *out)
{
    for (int i = 0; i < n; i++) {
        float s = 0.0f;
        for (int p = rp[i]; p <= diag[i]; p++) s = fmaf(p < diag[i] ? M[p] : 1.0f, v[ci[p]], s);
        tmp[i] = s;
    }
    for (int i = 0; i < n; i++) {
        float s = 0.0f;
        for (int p = diag[i]; p < rp[i + 1]; p++) s = fmaf(M[p], tmp[ci[p]], s);
        out[i] = s;
    }
}

int orc_bicgstab_mixed(int n, const int *rp, const int *ci, const int *diag,
                       const int *lev_l, const int *lev_u, const double *A, const double *b,
                       const float *F32, const orc_opts *o, double inner_tol, int max_refine,
                       double *x, float *iter_out, int *conv_out, double *iter_res_out,
                       double *true_res_out, int *refine_out)
{
    return orc_bicgstab_mixed_pc(n, rp, ci, diag, lev_l, lev_u, A, b, F32, 0, 0, o, inner_tol,
                                 max_refine, x, iter_out, conv_out, iter_res_out, true_res_out,
                                 refine_out);
}

/* isai = 0: P32 is the FP32 in-place LU factor (availability-order solves);
 * isai = 1: P32 is the FP32 ISAI(L)/ISAI(U) array, zp its zero-pivot flag;
 * isai = 2: P32 is the FP32 general ISAI of A on A's pattern, applied as one
 * FP32 SpMV (rows left to right), zp its zero-pivot flag */
int orc_bicgstab_mixed_pc(int n, const int *rp, const int *ci, const int *diag,
                          const int *lev_l, const int *lev_u, const double *A, const double *b,
                          const float *F32, int isai, int zp, const orc_opts *o, double inner_tol,
                          int max_refine, double *x, float *iter_out, int *conv_out,
                          double *iter_res_out, double *true_res_out, int *refine_out)
{
    const int nnz = rp[n];
    float *Af = (float *)malloc(sizeof(float) * (size_t)nnz);
    float *fb = (float *)malloc(sizeof(float) * (size_t)n * 10);
    double *r = (double *)malloc(sizeof(double) * (size_t)n * 2), *w = r + n;
    float *ri = fb, *rh = fb + n, *p = fb + 2 * n, *v = fb + 3 * n, *s = fb + 4 * n;
    float *t = fb + 5 * n, *ph = fb + 6 * n, *sh = fb + 7 * n, *d = fb + 8 * n, *tmp = fb + 9 * n;
    for (int q = 0; q < nnz; q++) Af[q] = (float)A[q];
    int bad = 0;
    if (isai) bad = zp;
    else for (int i = 0; i < n; i++) bad |= (F32[diag[i]] == 0.0f);
    for (int i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; }
    const double bnorm = sqrt(orc_dot(n, r, r, 64));
    const double tol_abs = (o->tol_mode == ORC_TOL_RELATIVE) ? o->tol * bnorm : o->tol;
    double rn = bnorm;
    float iter = 0.0f;
    int steps = 0, conv = ORC_NOT_CONVERGED;
    if (bad) {
        conv = ORC_ZERO_PIVOT;
    } else if (!isfinite(bnorm)) {
        conv = ORC_NONFINITE;
    } else if (bnorm == 0.0) {
        conv = ORC_CONVERGED;
        rn = 0.0;
    } else if (rn < tol_abs) {
        conv = ORC_CONVERGED;
    } else {
        while (steps < max_refine) {
            const double sc = 1.0 / rn;
            for (int i = 0; i < n; i++) {
                ri[i] = (float)(r[i] * sc);
                rh[i] = ri[i];
                d[i] = 0.0f; p[i] = 0.0f; v[i] = 0.0f;
            }
            const float r0 = sqrtf(dotf64w(n, ri, ri));
            /* the inner solve stops at the relative residual inner_tol, or looser
             * when that already takes |r| below the outer tolerance with the
             * margin inner_adapt (the last refinement step of most systems) */
            double want = inner_tol;
            if (o->inner_adapt > 0.0 && o->inner_adapt * tol_abs / rn > want) want = o->inner_adapt * tol_abs / rn;
            const float itol = (float)want * r0;
            float ires = r0, it = 0.0f;
            float rho_old = 1.0f, omega_old = 1.0f, alpha = 1.0f;
            int inner_bad = 0;
            while (it < (float)o->max_iter && ires >= itol) {
                const float rho_new = dotf64w(n, ri, rh);
                if (rho_new == 0.0f) break;
                const float beta = (rho_new / rho_old) * (alpha / omega_old);
                for (int i = 0; i < n; i++) p[i] = fmaf(beta, fmaf(-omega_old, v[i], p[i]), ri[i]);
                if (isai == 2) spmvf(n, rp, ci, F32, p, ph);
                else if (isai) isai_apply_f(n, rp, ci, diag, F32, p, tmp, ph);
                else lu_solve_avail_f(n, rp, ci, diag, lev_l, lev_u, F32, p, tmp, ph);
                spmvf(n, rp, ci, Af, ph, v);
                alpha = rho_new / dotf64w(n, rh, v);
                for (int i = 0; i < n; i++) d[i] = fmaf(alpha, ph[i], d[i]);
                for (int i = 0; i < n; i++) s[i] = fmaf(-alpha, v[i], ri[i]);
                ires = sqrtf(dotf64w(n, s, s));
                it += 0.5f;
                if (!isfinite(ires)) { inner_bad = 1; break; }
                if (ires < itol) break;
                if (isai == 2) spmvf(n, rp, ci, F32, s, sh);
                else if (isai) isai_apply_f(n, rp, ci, diag, F32, s, tmp, sh);
                else lu_solve_avail_f(n, rp, ci, diag, lev_l, lev_u, F32, s, tmp, sh);
                spmvf(n, rp, ci, Af, sh, t);
                const float ts = dotf64w(n, t, s);
                const float tt = dotf64w(n, t, t);
                const float omega_new = ts / tt;
                for (int i = 0; i < n; i++) d[i] = fmaf(omega_new, sh[i], d[i]);
                it += 0.5f;
                for (int i = 0; i < n; i++) ri[i] = fmaf(-omega_new, t[i], s[i]);
                ires = sqrtf(dotf64w(n, ri, ri));
                rho_old = rho_new;
                omega_old = omega_new;
                if (!isfinite(ires)) { inner_bad = 1; break; }
            }
            iter += it;
            steps++;
            if (inner_bad) { conv = ORC_NONFINITE; break; }
            for (int i = 0; i < n; i++) x[i] = fma((double)d[i], rn, x[i]);
            orc_spmv(n, rp, ci, A, x, w, 1);
            for (int i = 0; i < n; i++) r[i] = b[i] - w[i];
            rn = sqrt(orc_dot(n, r, r, 64));
            if (!isfinite(rn)) { conv = ORC_NONFINITE; break; }
            if (rn < tol_abs) { conv = ORC_CONVERGED; break; }
            if (it == 0.0f) break;
        }
    }
    if (conv == ORC_ZERO_PIVOT) {
        for (int i = 0; i < n; i++) x[i] = 0.0;
        rn = bnorm;
    }
    *iter_out = iter;
    *conv_out = conv;
    *iter_res_out = (conv == ORC_ZERO_PIVOT) ? 0.0 : rn;
    *true_res_out = rn;
    *refine_out = steps;
    free(Af); free(fb); free(r);
    return 0;
}

/* Entries with work of the persistent factor kernel (a dependency pair or
 * a pivot division), ordered by level in the entry dependency DAG, then by
 * position.  Returns their number. */
int orc_entry_order(int n, const int *rp, const int *ci, const int *diag, int *order)
{
    const int nnz = rp[n];
    int *elev = (int *)calloc((size_t)nnz, sizeof(int));
    int maxlev = 0;
    for (int i = 0; i < n; i++) {
        for (int e = rp[i]; e < rp[i + 1]; e++) {
            const int j = ci[e];
            const int kmax = (i < j ? i : j) - 1;
            int l = 0;
            for (int p = rp[i]; p < diag[i] && ci[p] <= kmax; p++) {
                const int q = find_pos(rp, ci, ci[p], j);
                if (q < 0) continue;
                const int m = (elev[p] > elev[q] ? elev[p] : elev[q]) + 1;
                if (m > l) l = m;
            }
            if (j < i && elev[diag[j]] + 1 > l) l = elev[diag[j]] + 1;
            elev[e] = l;
            if (l > maxlev) maxlev = l;
        }
    }
    int ne = 0;
    for (int l = 1; l <= maxlev; l++)
        for (int e = 0; e < nnz; e++)
            if (elev[e] == l) order[ne++] = e;
    free(elev);
    return ne;
}

/* The sweep order of the factor kernel's ParILU rounds: the entries with
 * work (orc_entry_order) stably sorted by their number of products
 * l_ik u_kj, fewest first (the entries early in the dependency order come
 * first, which the in-place sweeps need); entry q of it is lane q % 64's in
 * round q / 64.
 * Returns the entry count. */
int orc_sweep_order(int n, const int *rp, const int *ci, const int *diag, int *order)
{
    const int nnz = rp[n];
    int *eo = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    int *cnt = (int *)calloc((size_t)(nnz > 0 ? nnz : 1), sizeof(int));
    const int ne = orc_entry_order(n, rp, ci, diag, eo);
    int maxc = 0;
    for (int i = 0; i < n; i++)
        for (int e = rp[i]; e < rp[i + 1]; e++) {
            const int j = ci[e], kmax = (i < j ? i : j) - 1;
            int c = 0;
            for (int p = rp[i]; p < diag[i] && ci[p] <= kmax; p++)
                if (find_pos(rp, ci, ci[p], j) >= 0) c++;
            cnt[e] = c;
            if (c > maxc) maxc = c;
        }
    int w = 0;   /* by count, ascending; stable in the entry order */
    for (int c = 0; c <= maxc; c++)
        for (int q = 0; q < ne; q++)
            if (cnt[eo[q]] == c) order[w++] = eo[q];
    free(eo); free(cnt);
    return ne;
}

/* ParILU(0) Jacobi sweeps from `init` (NULL: from A; only the entries with
 * work take the guess, the others are exact = a_ij), stopped as soon as the
 * iterate a sweep starts from has ||(LU - A) on the pattern||_F <=
 * stop_tol ||A||_F (that iterate is kept), at most `sweeps` sweeps; when
 * the cap is reached without meeting the tolerance, exact ILU(0) replaces
 * the factors.  stop_tol <= 0: exactly `sweeps` sweeps.  The residual of an
 * iterate is summed in the kernel's order: entry q of the work order
 * (orc_sweep_order) goes to lane q % 64, lanes adding in increasing q, then
 * the wave tree.  *status = sweeps used, or -1 for the fallback.  F[nnz]
 * out. */
int orc_parilu0_stop(int n, const int *rp, const int *ci, const int *diag, const double *A,
                     const double *init, double *F, int sweeps, double stop_tol, int *status, int inplace)
{
    const int nnz = rp[n];
    int *order = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    double *W = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    const int ne = orc_sweep_order(n, rp, ci, diag, order);
    memcpy(F, A, sizeof(double) * (size_t)nnz);
    if (init)
        for (int q = 0; q < ne; q++) F[order[q]] = init[order[q]];
    const double anorm2 = stop_tol > 0.0 ? orc_dot(nnz, A, A, 64) : 0.0;
    const double lim2 = stop_tol * stop_tol * anorm2;
    int used = 0, met = stop_tol <= 0.0;
    int *rowof = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    for (int i = 0; i < n; i++)
        for (int e = rp[i]; e < rp[i + 1]; e++) rowof[e] = i;
    while (used < sweeps) {
        /* Jacobi: every entry reads the iterate the sweep started from (W).
         * In place: the sweep's rounds (64 consecutive entries of the order)
         * run one after the other; a round reads the iterate as the earlier
         * rounds of the sweep left it (W refreshed at every round start) and
         * writes its entries when all of them are computed. */
        memcpy(W, F, sizeof(double) * (size_t)nnz);
        double t[64];
        for (int l = 0; l < 64; l++) t[l] = 0.0;
        for (int q = 0; q < ne; q++) {
            if (inplace && q % 64 == 0 && q > 0)
                for (int qq = q - 64; qq < q; qq++) W[order[qq]] = F[order[qq]];
            const int e = order[q], i = rowof[e], j = ci[e];
            const int kmax = (i < j ? i : j) - 1;
            double sum = 0.0;
            for (int p = rp[i]; p < diag[i] && ci[p] <= kmax; p++) {
                const int qq = find_pos(rp, ci, ci[p], j);
                if (qq >= 0) sum = fma(W[p], W[qq], sum);
            }
            const double tt = A[e] - sum;
            double res;
            if (j < i) {
                const double dv = W[diag[j]];
                res = tt - W[e] * dv;
                F[e] = tt / dv;
            } else {
                res = tt - W[e];
                F[e] = tt;
            }
            t[q % 64] = fma(res, res, t[q % 64]);
        }
        for (int off = 32; off > 0; off /= 2)
            for (int l = 0; l < off; l++) t[l] = t[l] + t[l + off];
        used++;
        if (stop_tol > 0.0 && t[0] <= lim2) {
            /* Jacobi: t is the residual of the iterate the sweep started
             * from, which is kept.  In place: t sums the residuals each
             * round saw; the iterate after the sweep is kept. */
            if (!inplace) {
                memcpy(F, W, sizeof(double) * (size_t)nnz);
                used--;
            }
            met = 1;
            break;
        }
    }
    if (!met) {
        memcpy(F, A, sizeof(double) * (size_t)nnz);
        orc_ilu0(n, rp, ci, diag, F);
        *status = -1;
    } else {
        *status = used;
    }
    free(order); free(W); free(rowof);
    return 0;
}

/* whole-batch driver of the split factor + solve paths: precond 0 = exact
 * ILU(0), 1 = ParILU(0) (with init / stop_tol as orc_parilu0_stop), 2 / 3 =
 * the same followed by ISAI(L), ISAI(U), 4 = the general ISAI of A; precision 0 = FP64 BiCGSTAB
 * (availability-order solves; FP64 ISAI applies, a zero pivot of their
 * generation failing the system as in the batch driver), 1 = FP32 inner /
 * FP64 refinement over the FP32-rounded factors or the FP32 ISAI. */
typedef struct {
    int n, nnz, b0, b1, precond, sweeps, precision, max_refine, inplace;
    double stop_tol, inner_tol;
    const int *rp, *ci, *diag, *lev;
    const double *A, *b, *init;
    double *x;
    const orc_opts *o;
    float *iter; int *conv; double *iter_res, *true_res; int *refine, *fstat;
} split_job;

static void *split_worker(void *arg)
{
    split_job *j = (split_job *)arg;
    const int n = j->n, nnz = j->nnz;
    double *F = (double *)malloc(sizeof(double) * (size_t)nnz);
    double *P1 = (double *)malloc(sizeof(double) * (size_t)nnz);
    double *P2 = (double *)malloc(sizeof(double) * (size_t)nnz);
    float *F32 = (float *)malloc(sizeof(float) * (size_t)nnz);
    for (int s = j->b0; s < j->b1; s++) {
        const double *A = j->A + (size_t)s * nnz;
        const double *b = j->b + (size_t)s * n;
        double *x = j->x + (size_t)s * n;
        int st = 0;
        if (j->precond == 4) {
            /* general ISAI of A (no factorization): FP64 values, then FP32 (precision 1) */
            const int zs = orc_isai(n, j->rp, j->ci, j->diag, A, ORC_ISAI_GENERAL, P1);
            if (j->fstat) j->fstat[s] = 0;
            if (j->precision == 1) {
                for (int q = 0; q < nnz; q++) F32[q] = zs == 0 ? (float)P1[q] : 0.0f;
                orc_bicgstab_mixed_pc(n, j->rp, j->ci, j->diag, j->lev, j->lev + n, A, b, F32, 2, zs != 0, j->o,
                                      j->inner_tol, j->max_refine, x, &j->iter[s], &j->conv[s],
                                      &j->iter_res[s], &j->true_res[s], &j->refine[s]);
            } else if (zs != 0) {
                for (int i = 0; i < n; i++) x[i] = 0.0;
                j->iter[s] = 0.0f; j->conv[s] = zs; j->iter_res[s] = 0.0;
                j->true_res[s] = sqrt(orc_dot(n, b, b, j->o->dot_width));
                j->refine[s] = 0;
            } else {
                orc_bicgstab(n, j->rp, j->ci, j->diag, A, b, x, ORC_PC_ISAI_GEN, P1, NULL, j->o,
                             &j->iter[s], &j->conv[s], &j->iter_res[s], &j->true_res[s]);
                j->refine[s] = 0;
            }
            continue;
        }
        if (j->precond == 1 || j->precond == 3) {
            orc_parilu0_stop(n, j->rp, j->ci, j->diag, A, j->init ? j->init + (size_t)s * nnz : NULL, F,
                             j->sweeps, j->stop_tol, &st, j->inplace);
        } else {
            memcpy(F, A, sizeof(double) * (size_t)nnz);
            orc_ilu0(n, j->rp, j->ci, j->diag, F);
        }
        if (j->fstat) j->fstat[s] = st;
        if (j->precision == 1 && j->precond >= 2) {
            const int zp = orc_isai_lu_f32(n, j->rp, j->ci, j->diag, F, F32) != 0;
            orc_bicgstab_mixed_pc(n, j->rp, j->ci, j->diag, j->lev, j->lev + n, A, b, F32, 1, zp, j->o,
                                  j->inner_tol, j->max_refine, x, &j->iter[s], &j->conv[s],
                                  &j->iter_res[s], &j->true_res[s], &j->refine[s]);
        } else if (j->precision == 1) {
            for (int q = 0; q < nnz; q++) F32[q] = (float)F[q];
            orc_bicgstab_mixed(n, j->rp, j->ci, j->diag, j->lev, j->lev + n, A, b, F32, j->o,
                               j->inner_tol, j->max_refine, x, &j->iter[s], &j->conv[s],
                               &j->iter_res[s], &j->true_res[s], &j->refine[s]);
        } else if (j->precond >= 2) {
            int zs = orc_isai(n, j->rp, j->ci, j->diag, F, ORC_ISAI_LOWER, P1);
            if (zs == 0) zs = orc_isai(n, j->rp, j->ci, j->diag, F, ORC_ISAI_UPPER, P2);
            if (zs != 0) {
                for (int i = 0; i < n; i++) x[i] = 0.0;
                j->iter[s] = 0.0f; j->conv[s] = zs; j->iter_res[s] = 0.0;
                j->true_res[s] = sqrt(orc_dot(n, b, b, j->o->dot_width));
            } else {
                orc_bicgstab(n, j->rp, j->ci, j->diag, A, b, x, ORC_PC_ISAI_LU, P1, P2, j->o,
                             &j->iter[s], &j->conv[s], &j->iter_res[s], &j->true_res[s]);
            }
            j->refine[s] = 0;
        } else {
            orc_bicgstab(n, j->rp, j->ci, j->diag, A, b, x, ORC_PC_LU, F, NULL, j->o,
                         &j->iter[s], &j->conv[s], &j->iter_res[s], &j->true_res[s]);
            j->refine[s] = 0;
        }
    }
    free(F); free(P1); free(P2); free(F32);
    return NULL;
}

int orc_batch_solve_split(int n, int nnz, const int *rp, const int *ci, int batch,
                          const double *A, const double *b, double *x, int precond, int sweeps,
                          const double *init, double stop_tol, int precision, double inner_tol,
                          int max_refine, const orc_opts *o, float *iter, int *conv,
                          double *iter_res, double *true_res, int *refine, int *fstat, int threads,
                          int parilu_inplace)
{
    int *diag = (int *)malloc(sizeof(int) * (size_t)n);
    if (orc_find_diag(n, rp, ci, diag) > 0) { free(diag); return -5; }
    int *lev = (int *)malloc(sizeof(int) * 2 * (size_t)n);
    orc_trsv_levels(n, rp, ci, diag, lev, lev + n);
    if (threads < 1) threads = 1;
    if (threads > batch) threads = batch > 0 ? batch : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    split_job *jobs = (split_job *)malloc(sizeof(split_job) * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        split_job *j = &jobs[t];
        j->n = n; j->nnz = nnz; j->precond = precond; j->sweeps = sweeps; j->precision = precision;
        j->max_refine = max_refine; j->stop_tol = stop_tol; j->inner_tol = inner_tol;
        j->inplace = parilu_inplace;
        j->b0 = (int)((long long)batch * t / threads);
        j->b1 = (int)((long long)batch * (t + 1) / threads);
        j->rp = rp; j->ci = ci; j->diag = diag; j->lev = lev; j->A = A; j->b = b; j->init = init;
        j->x = x; j->o = o; j->iter = iter; j->conv = conv; j->iter_res = iter_res;
        j->true_res = true_res; j->refine = refine; j->fstat = fstat;
        if (threads == 1) split_worker(j);
        else pthread_create(&th[t], NULL, split_worker, j);
    }
    if (threads > 1)
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th); free(jobs); free(diag); free(lev);
    return 0;
}

/* One system through the ISAI solvers with power-k patterns: precond 2 =
 * exact ILU(0) then ISAI(L), ISAI(U) on |L|^power, |U|^power (the
 * reference's ILU-ISAI solver uses power 2, PrecondBiCGSTAB_ilu_isai_gpu.cu:
 * 297-299), 3 = the same after `sweeps` ParILU(0) sweeps, 4 = general ISAI
 * of A on |A|^power; then BiCGSTAB applying them as SpMVs on their own
 * patterns.  A factorization or ISAI zero pivot gives x = 0, flag -2. */
int orc_solve_isai_pow(int n, const int *rp, const int *ci, const double *A, const double *b, double *x,
                       int precond, int sweeps, int power, const orc_opts *o, float *iter, int *conv,
                       double *iter_res, double *true_res)
{
    const int nnz = rp[n];
    int *diag = (int *)malloc(sizeof(int) * (size_t)n);
    if (orc_find_diag(n, rp, ci, diag) > 0) { free(diag); return -5; }
    const int t1 = precond == 4 ? ORC_ISAI_GENERAL : ORC_ISAI_LOWER;
    const int n1 = orc_isai_pattern(n, rp, ci, t1, power, NULL, NULL);
    const int n2 = precond == 4 ? 0 : orc_isai_pattern(n, rp, ci, ORC_ISAI_UPPER, power, NULL, NULL);
    int *lrp = (int *)malloc(sizeof(int) * (size_t)(n + 1)), *lci = (int *)malloc(sizeof(int) * (size_t)(n1 + 1));
    int *urp = (int *)malloc(sizeof(int) * (size_t)(n + 1)), *uci = (int *)malloc(sizeof(int) * (size_t)(n2 + 1));
    orc_isai_pattern(n, rp, ci, t1, power, lrp, lci);
    if (precond != 4) orc_isai_pattern(n, rp, ci, ORC_ISAI_UPPER, power, urp, uci);
    double *F = (double *)malloc(sizeof(double) * (size_t)nnz);
    double *M1 = (double *)malloc(sizeof(double) * (size_t)(n1 + 1));
    double *M2 = (double *)malloc(sizeof(double) * (size_t)(n2 + 1));
    memcpy(F, A, sizeof(double) * (size_t)nnz);
    int st = 0;
    if (precond == 2) st = orc_ilu0(n, rp, ci, diag, F);
    else if (precond == 3) st = orc_parilu0_sweeps(n, rp, ci, diag, A, F, sweeps);
    if (st == 0) st = isai_on_rows(n, rp, ci, F, t1, lrp, lci, M1);
    if (st == 0 && precond != 4) st = isai_on_rows(n, rp, ci, F, ORC_ISAI_UPPER, urp, uci, M2);
    int rc = 0;
    if (st == -4) {
        rc = -4;
    } else if (st != 0) {
        for (int i = 0; i < n; i++) x[i] = 0.0;
        *iter = 0.0f; *conv = st; *iter_res = 0.0;
        *true_res = sqrt(orc_dot(n, b, b, o->dot_width));
    } else {
        const isai_pats pats = {lrp, lci, urp, uci};
        bicgstab_core(n, rp, ci, diag, A, b, x, precond == 4 ? ORC_PC_ISAI_GEN_PAT : ORC_PC_ISAI_LU_PAT,
                      M1, M2, &pats, o, iter, conv, iter_res, true_res);
    }
    free(diag); free(lrp); free(lci); free(urp); free(uci); free(F); free(M1); free(M2);
    return rc;
}
