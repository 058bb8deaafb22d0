// Batched shared-pattern ILU(0) / ParILU(0) / ISAI preconditioned BiCGSTAB
// for gfx950 (MI355X), and the C ABI declared in include/bilu.h.
//
// Execution model: one wave64 workgroup per system (grid = batch).  The
// system's matrix values, factors and the vectors that are gathered from
// (SpMV / triangular-solve inputs) live in LDS; the Krylov vectors live in
// VGPRs in a fixed "row-cyclic" distribution (element i in lane i%64,
// register slot i/64), so axpys are register-only and dot products are a
// per-lane strided sum followed by a 6-step butterfly -- no LDS, no
// barrier.  Scalars (rho, alpha, omega, norms) come out of the butterfly
// identical in every lane, so all control flow is wave-uniform.
// Pattern-derived schedules (level sets, ILU(0) update lists, ParILU
// dependency lists, ISAI gather maps) are computed once on the host
// (pattern.cpp) and read from global memory (L1/L2 resident: a few KB).
//
// Arithmetic order is fixed and mirrors the CPU oracle bit for bit:
// explicit fma() everywhere a product is accumulated, compiled with
// -ffp-contract=off; SpMV row sums and triangular-solve sums run in CSR
// order (L left->right, U right->left as merged.cu:43-51,85-95); dots sum
// idx = lane, lane+64, ... per lane and then combine lanes with a butterfly
// whose lane-0 association equals a shfl_down tree (oracle dot width 64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/bilu.h"
#include "mlp_scaling.hpp"
#include "pattern.hpp"

namespace {

constexpr int WAVE = 64;
constexpr int MAX_NR = 8;        // rows per lane in the register layout: n <= 512
constexpr int ISAI_GROUP = 32;   // lanes per ISAI row (reference row_size_limit 32)
constexpr int ISAI_TRI_MAX = 2 * 64; // longest ISAI(L) / ISAI(U) row: two unknowns per lane

enum PcKind { PC_LU = 0, PC_ISAI_LU = 1, PC_ISAI_GEN = 2, PC_NONE = 3, PC_ISAI_POW = 4, PC_ISAI_GEN_POW = 5 };
// FK_ISAI_LOADED (one-wave solve kernel, ISAI preconditioners): the ISAI values
// come from the generation launches (gF, SolveParams::m2, ::zstatus)
enum FactorKind { FK_ILU = 0, FK_PARILU = 1, FK_LOAD = 2, FK_NONE = 3, FK_ILU_ENTRIES = 4, FK_ILU_ERAS = 5,
                  FK_ISAI_LOADED = 6 };

struct DevIsai {
    const int *beg, *m, *goff, *gather, *rhs;
    const int *order;  // rows by decreasing m: n_pair rows two per wave, then four per wave
    int n_pair;
    int max_m;         // longest row (lower / upper beyond ISAI_GROUP: isai_tri_wide)
};

// an ISAI on a chosen pattern (powers > 1): the pattern and its gather map
struct DevIsaiPow {
    int nnz = 0;
    const int *rp = nullptr, *ci = nullptr;
    DevIsai map{};
};

struct DevPattern {
    int n, nnz, nlev_l, nlev_u;
    const int *rp, *ci, *diag, *erow;
    const int *lev_l_ptr, *lev_l_rows, *lev_u_ptr, *lev_u_rows;
    const int *upd_ptr;
    const int2 *upd;
    const int *dep_ptr;
    const int2 *dep;
    const int *l_src, *u_src;
    int l_nnz, u_nnz;
    DevIsai isai[3];
    // alternative exact ILU(0) schedules: entry levels, pivot eras
    int nlev_e, n_eras;
    const int *lev_e_ptr, *lev_e_ent;
    const int *era_div_ptr, *era_tgt_ptr, *era_tgt, *tgt_pair_ptr;
    const int2 *era_div, *tgt_pair;
};

struct SolveParams {
    double tol;
    int tol_mode, max_iter, restart, max_restarts;
    int factor, sweeps;
    const int *zstatus = nullptr;   // FK_ISAI_LOADED: zero-pivot flags of the generation, per system
    const double *m2 = nullptr;     // FK_ISAI_LOADED, ISAI(L, U) on |L|^k, |U|^k: the ISAI(U) values
};

// ---------------------------------------------------------------- copies
// global -> LDS copy of cnt values by one wave (lane l takes l, l + 64, ...),
// U loads in flight per lane: a plain copy loop waits for each load before
// its store, which leaves a wave idle on one memory latency per 64 values
template <typename TD, typename TS>
__device__ __forceinline__ void wave_copy(TD *dst, const TS *__restrict__ src, int cnt, int lane)
{
    constexpr int U = 8;
    int e0 = lane;
    for (; e0 + (U - 1) * WAVE < cnt; e0 += U * WAVE) {
        TS v[U];
#pragma unroll
        for (int j = 0; j < U; j++) v[j] = src[e0 + j * WAVE];
#pragma unroll
        for (int j = 0; j < U; j++) dst[e0 + j * WAVE] = (TD)v[j];
    }
    if (e0 < cnt) {
        TS v[U];
#pragma unroll
        for (int j = 0; j < U; j++) v[j] = (e0 + j * WAVE < cnt) ? src[e0 + j * WAVE] : TS(0);
#pragma unroll
        for (int j = 0; j < U; j++)
            if (e0 + j * WAVE < cnt) dst[e0 + j * WAVE] = (TD)v[j];
    }
}

// wave_copy with U values per lane in flight: one memory round trip per
// 64 U values (the persistent kernels run few waves per SIMD, so a copy's
// latency is not hidden by other waves)
template <int U, typename TD, typename TS>
__device__ __forceinline__ void wave_copy_wide(TD *dst, const TS *__restrict__ src, int cnt, int lane)
{
    for (int e0 = lane; e0 < cnt; e0 += U * WAVE) {
        TS v[U];
#pragma unroll
        for (int j = 0; j < U; j++) v[j] = (e0 + j * WAVE < cnt) ? src[e0 + j * WAVE] : TS(0);
#pragma unroll
        for (int j = 0; j < U; j++)
            if (e0 + j * WAVE < cnt) dst[e0 + j * WAVE] = (TD)v[j];
    }
}

// ---------------------------------------------------------------- wave ops
// 64-bit lane exchanges without the LDS pipe: DPP within 16-lane rows,
// gfx950 permlane swaps across rows / halves
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// value of lane (lane ^ 32)
__device__ __forceinline__ double lane_xor32(double v)
{
    const unsigned lo = __double2loint(v), hi = __double2hiint(v);
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    const bool low = (__lane_id() & 32) == 0;
    return __hiloint2double(low ? b[1] : b[0], low ? a[1] : a[0]);
}

// value of lane (lane ^ 16)
__device__ __forceinline__ double lane_xor16(double v)
{
    const unsigned lo = __double2loint(v), hi = __double2hiint(v);
    const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const bool even = (__lane_id() & 16) == 0;
    return __hiloint2double(even ? b[1] : b[0], even ? a[1] : a[0]);
}

// Sum over the wave, identical in every lane.  The association is the xor
// butterfly with masks 32,16,8,4,2,1, whose lane-0 value equals a shfl_down
// tree (offsets 32..1).  Masks 8 and 4 use row rotations: after the earlier
// stages the partial sums are periodic, so the rotated partner holds the
// same bits as the xor partner.
__device__ __forceinline__ double wave_sum(double v)
{
    v = v + lane_xor32(v);
    v = v + lane_xor16(v);
    v = v + dpp_mov<0x128>(v);  // row_ror:8
    v = v + dpp_mov<0x124>(v);  // row_ror:4
    v = v + dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
    v = v + dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
    return v;
}

// FP32 versions of the lane exchanges and the wave sum (same association)
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ float lane_xor32(float v)
{
    const unsigned u = __float_as_uint(v);
    const auto a = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    return __uint_as_float((__lane_id() & 32) == 0 ? a[1] : a[0]);
}

__device__ __forceinline__ float lane_xor16(float v)
{
    const unsigned u = __float_as_uint(v);
    const auto a = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    return __uint_as_float((__lane_id() & 16) == 0 ? a[1] : a[0]);
}

__device__ __forceinline__ float wave_sum(float v)
{
    v = v + lane_xor32(v);
    v = v + lane_xor16(v);
    v = v + dpp_mov<0x128>(v);
    v = v + dpp_mov<0x124>(v);
    v = v + dpp_mov<0x4E>(v);
    v = v + dpp_mov<0xB1>(v);
    return v;
}

__device__ __forceinline__ double fmaT(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// a lane's NR rows (lane + 64 k) of a vector, in registers
template <int NR, typename T = double>
struct Vec {
    T e[NR];
};

template <int NR, typename T>
__device__ __forceinline__ T vdot(const Vec<NR, T> &a, const Vec<NR, T> &b, int lane, int n)
{
    T s = T(0);
#pragma unroll
    for (int k = 0; k < NR; k++)
        if (lane + WAVE * k < n) s = fmaT(a.e[k], b.e[k], s);
    return wave_sum(s);
}

template <int NR, typename T>
__device__ __forceinline__ void to_lds(const Vec<NR, T> &v, T *dst, int lane, int n)
{
#pragma unroll
    for (int k = 0; k < NR; k++)
        if (lane + WAVE * k < n) dst[lane + WAVE * k] = v.e[k];
}

template <int NR, typename T>
__device__ __forceinline__ void from_lds(Vec<NR, T> &v, const T *src, int lane, int n)
{
#pragma unroll
    for (int k = 0; k < NR; k++) v.e[k] = (lane + WAVE * k < n) ? src[lane + WAVE * k] : T(0);
}

// y = A x, rows in the register layout; A over positions [rp[i], rp[i+1])
template <int NR>
__device__ __forceinline__ void spmv(const DevPattern &P, const double *A, const double *x,
                                     Vec<NR> &y, int lane)
{
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int i = lane + WAVE * k;
        double acc = 0.0;
        if (i < P.n) {
            const int e = P.rp[i + 1];
            for (int p = P.rp[i]; p < e; p++) acc = fma(A[p], x[P.ci[p]], acc);
        }
        y.e[k] = acc;
    }
}

// out = U^{-1} L^{-1} vin on the in-place factor F, level by level
__device__ void lu_apply(const DevPattern &P, const double *F, const double *vin, double *y,
                         double *out, int lane)
{
    for (int l = 0; l < P.nlev_l; l++) {
        const int b = P.lev_l_ptr[l], e = P.lev_l_ptr[l + 1];
        for (int idx = b + lane; idx < e; idx += WAVE) {
            const int i = P.lev_l_rows[idx];
            double acc = 0.0;
            const int pd = P.diag[i];
            for (int p = P.rp[i]; p < pd; p++) acc = fma(F[p], y[P.ci[p]], acc);
            y[i] = vin[i] - acc;
        }
        __syncthreads();
    }
    for (int l = 0; l < P.nlev_u; l++) {
        const int b = P.lev_u_ptr[l], e = P.lev_u_ptr[l + 1];
        for (int idx = b + lane; idx < e; idx += WAVE) {
            const int i = P.lev_u_rows[idx];
            double acc = 0.0;
            const int pd = P.diag[i];
            for (int p = P.rp[i + 1] - 1; p > pd; p--) acc = fma(F[p], out[P.ci[p]], acc);
            out[i] = (y[i] - acc) / F[pd];
        }
        __syncthreads();
    }
}

// out = aiU (aiL vin); aiL on [rp[i], diag[i]] of isL, aiU on [diag[i], rp[i+1])
// of isU except its diagonal, which is kept in udiag[i]
__device__ void isai_lu_apply(const DevPattern &P, const double *isL, const double *isU,
                              const double *udiag, const double *vin, double *tmp, double *out,
                              int lane)
{
    for (int i = lane; i < P.n; i += WAVE) {
        double acc = 0.0;
        const int pd = P.diag[i];
        for (int p = P.rp[i]; p <= pd; p++) acc = fma(isL[p], vin[P.ci[p]], acc);
        tmp[i] = acc;
    }
    __syncthreads();
    for (int i = lane; i < P.n; i += WAVE) {
        const int pd = P.diag[i];
        double acc = fma(udiag[i], tmp[i], 0.0);
        const int e = P.rp[i + 1];
        for (int p = pd + 1; p < e; p++) acc = fma(isU[p], tmp[P.ci[p]], acc);
        out[i] = acc;
    }
    __syncthreads();
}

__device__ void spmv_lds(const DevPattern &P, const double *M, const double *vin, double *out,
                         int lane)
{
    for (int i = lane; i < P.n; i += WAVE) {
        double acc = 0.0;
        const int e = P.rp[i + 1];
        for (int p = P.rp[i]; p < e; p++) acc = fma(M[p], vin[P.ci[p]], acc);
        out[i] = acc;
    }
    __syncthreads();
}

// out = aiU (aiL vin) with aiL, aiU on their own patterns, rows left to right
__device__ void isai_pow_apply(const DevIsaiPow &QL, const double *isL, const DevIsaiPow &QU, const double *isU,
                               int n, const double *vin, double *tmp, double *out, int lane)
{
    for (int i = lane; i < n; i += WAVE) {
        double acc = 0.0;
        const int e = QL.rp[i + 1];
        for (int p = QL.rp[i]; p < e; p++) acc = fma(isL[p], vin[QL.ci[p]], acc);
        tmp[i] = acc;
    }
    __syncthreads();
    for (int i = lane; i < n; i += WAVE) {
        double acc = 0.0;
        const int e = QU.rp[i + 1];
        for (int p = QU.rp[i]; p < e; p++) acc = fma(isU[p], tmp[QU.ci[p]], acc);
        out[i] = acc;
    }
    __syncthreads();
}

__device__ void spmv_pow(const DevIsaiPow &Q, const double *M, int n, const double *vin, double *out, int lane)
{
    for (int i = lane; i < n; i += WAVE) {
        double acc = 0.0;
        const int e = Q.rp[i + 1];
        for (int p = Q.rp[i]; p < e; p++) acc = fma(M[p], vin[Q.ci[p]], acc);
        out[i] = acc;
    }
    __syncthreads();
}

// exact ILU(0) in place, rows level by level (IKJ order == the reference KIJ
// order per entry, merged.cu:145-246); returns nonzero on a zero pivot
__device__ int ilu_factor(const DevPattern &P, double *F, int lane)
{
    int bad = 0;
    for (int l = 0; l < P.nlev_l; l++) {
        const int b = P.lev_l_ptr[l], e = P.lev_l_ptr[l + 1];
        for (int idx = b + lane; idx < e; idx += WAVE) {
            const int i = P.lev_l_rows[idx];
            const int pd = P.diag[i];
            for (int p = P.rp[i]; p < pd; p++) {
                const double ukk = F[P.diag[P.ci[p]]];
                bad |= (ukk == 0.0);
                const double lik = F[p] / ukk;
                F[p] = lik;
                const int ue = P.upd_ptr[p + 1];
                for (int u = P.upd_ptr[p]; u < ue; u++) {
                    const int2 d = P.upd[u];
                    F[d.x] = fma(-lik, F[d.y], F[d.x]);
                }
            }
        }
        __syncthreads();
    }
    for (int i = lane; i < P.n; i += WAVE) bad |= (F[P.diag[i]] == 0.0);
    return __any(bad);
}

// exact ILU(0) in place by entries, level by level (the dependency-list
// schedule, ILU_0_gpu.cu:693-935, with level barriers instead of ready
// flags): lane per entry, acc = a_ij, acc = fma(-l_ik, u_kj, acc) for k
// increasing, then / u_jj below the diagonal -- the bits of ilu_factor
__device__ int ilu_factor_entries(const DevPattern &P, double *F, int lane)
{
    int bad = 0;
    for (int l = 0; l < P.nlev_e; l++) {
        const int b = P.lev_e_ptr[l], e = P.lev_e_ptr[l + 1];
        for (int idx = b + lane; idx < e; idx += WAVE) {
            const int pos = P.lev_e_ent[idx];
            double acc = F[pos];
            const int de = P.dep_ptr[pos + 1];
            for (int d = P.dep_ptr[pos]; d < de; d++) {
                const int2 pq = P.dep[d];
                acc = fma(-F[pq.x], F[pq.y], acc);
            }
            const int j = P.ci[pos];
            if (j < P.erow[pos]) {
                const double ukk = F[P.diag[j]];
                bad |= (ukk == 0.0);
                acc = acc / ukk;
            }
            F[pos] = acc;
        }
        __syncthreads();
    }
    for (int i = lane; i < P.n; i += WAVE) bad |= (F[P.diag[i]] == 0.0);
    return __any(bad);
}

// exact ILU(0) in place by pivot eras (right-looking, ILU_0_gpu.cu:947-1229
// with the era's pivots spread over the lanes): divide the eras' columns,
// then every updated entry takes its era's updates in increasing k
__device__ int ilu_factor_eras(const DevPattern &P, double *F, int lane)
{
    int bad = 0;
    for (int r = 0; r < P.n_eras; r++) {
        const int db = P.era_div_ptr[r], de = P.era_div_ptr[r + 1];
        for (int idx = db + lane; idx < de; idx += WAVE) {
            const int2 pd = P.era_div[idx];
            const double ukk = F[pd.y];
            bad |= (ukk == 0.0);
            F[pd.x] = F[pd.x] / ukk;
        }
        __syncthreads();
        const int tb = P.era_tgt_ptr[r], te = P.era_tgt_ptr[r + 1];
        for (int idx = tb + lane; idx < te; idx += WAVE) {
            const int t = P.era_tgt[idx];
            double acc = F[t];
            const int pe = P.tgt_pair_ptr[idx + 1];
            for (int d = P.tgt_pair_ptr[idx]; d < pe; d++) {
                const int2 pq = P.tgt_pair[d];
                acc = fma(-F[pq.x], F[pq.y], acc);
            }
            F[t] = acc;
        }
        __syncthreads();
    }
    for (int i = lane; i < P.n; i += WAVE) bad |= (F[P.diag[i]] == 0.0);
    return __any(bad);
}

// zero-pivot status of in-place factors: one wave per system
__global__ __launch_bounds__(WAVE) void k_diag_status(DevPattern P, int batch, const double *__restrict__ gF,
                                                     int *__restrict__ g_status)
{
    const int sys = blockIdx.x;
    if (sys >= batch) return;
    const double *Fp = gF + (size_t)sys * P.nnz;
    int bad = 0;
    for (int i = threadIdx.x; i < P.n; i += WAVE) bad |= (Fp[P.diag[i]] == 0.0);
    bad = __any(bad);
    if (threadIdx.x == 0) g_status[sys] = bad ? BILU_ZERO_PIVOT : 0;
}

// Jacobi ParILU(0) sweeps (parILU_0_gpu.cu:18-57 formula); F holds the
// initial guess on entry, the result on exit; W is scratch of nnz doubles
__device__ void parilu_sweeps(const DevPattern &P, const double *A, double *F, double *W,
                              int sweeps, int lane)
{
    for (int s = 0; s < sweeps; s++) {
        for (int e = lane; e < P.nnz; e += WAVE) W[e] = F[e];
        __syncthreads();
        for (int e = lane; e < P.nnz; e += WAVE) {
            const int i = P.erow[e], j = P.ci[e];
            double sum = 0.0;
            const int de = P.dep_ptr[e + 1];
            for (int d = P.dep_ptr[e]; d < de; d++) {
                const int2 pq = P.dep[d];
                sum = fma(W[pq.x], W[pq.y], sum);
            }
            F[e] = (i > j) ? (A[e] - sum) / W[P.diag[j]] : A[e] - sum;
        }
        __syncthreads();
    }
}

// ISAI (power 1) rows, two rows per wave (32 lanes each, lane t owns dense
// column t of the block: the reference's "implicit transpose",
// isai_gpu.cu:559-626).  Lower/upper: column-sweep triangular solves
// (isai_gpu.cu:325-389); general: Gauss-Jordan with the reference pivot
// choice (isai_gpu.cu:417-554).  vals: typed view input; out[pos] for the
// row's positions (diagonal of UPPER goes to udiag when udiag != nullptr).
__device__ __forceinline__ double view_val(const double *vals, int g)
{
    return g >= 0 ? vals[g] : (g == -2 ? 1.0 : 0.0);
}

// view_val with an unconditional LDS read (loads issue together)
__device__ __forceinline__ double view_val_u(const double *vals, int g)
{
    const double a = vals[g > 0 ? g : 0];
    return g >= 0 ? a : (g == -2 ? 1.0 : 0.0);
}

__device__ __forceinline__ double lane_value(double v, int src)  // src wave-uniform
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), src),
                            __builtin_amdgcn_readlane(__double2loint(v), src));
}

// ISAI(L) / ISAI(U) rows of up to ISAI_TRI_MAX entries (power-k patterns;
// the reference stops at 32 on an assert): one row per wave, lane t holding
// unknowns t and t + 64, each receiving the column sweep's operations in the
// order isai_tri applies them.  The finished unknown of column c is read with
// v_readlane; the gather entries and values of 8 columns are loaded ahead, at
// clamped positions.  The wave takes rows w, w + nw, ...
__device__ int isai_tri_wide(const DevIsai &M, int n, int type, const double *vals, double *out,
                             double *udiag, int lane, int w, int nw)
{
    constexpr int CH = 8;
    const bool lower = type == BILU_ISAI_LOWER;
    const int t0 = lane, t1 = lane + WAVE;
    int bad = 0;
    for (int i = w; i < n; i += nw) {
        const int m = M.m[i];
        const int *g = M.gather + M.goff[i];
        const int last = m * m - 1;
        double s0 = (t0 == (lower ? m - 1 : 0)) ? 1.0 : 0.0;
        double s1 = (t1 == (lower ? m - 1 : 0)) ? 1.0 : 0.0;
        for (int q0 = 0; q0 < m; q0 += CH) {
            // columns of this chunk in sweep order: c = m-1-q (lower) or q (upper)
            double v0[CH], v1[CH], dg[CH];
#pragma unroll
            for (int j = 0; j < CH; j++) {
                const int q = min(q0 + j, m - 1);
                const int c = lower ? m - 1 - q : q;
                const int g0 = g[min(c * m + t0, last)], g1 = g[min(c * m + t1, last)], gd = g[c * m + c];
                v0[j] = view_val_u(vals, g0);
                v1[j] = view_val_u(vals, g1);
                dg[j] = view_val_u(vals, gd);
            }
#pragma unroll
            for (int j = 0; j < CH; j++) {
                const int q = q0 + j;
                if (q < m) {
                    const int c = lower ? m - 1 - q : q;
                    if (lane == (c & (WAVE - 1))) {
                        bad |= (dg[j] == 0.0);
                        if (c < WAVE) s0 = s0 / dg[j];
                        else s1 = s1 / dg[j];
                    }
                    const double f = lane_value(c < WAVE ? s0 : s1, c & (WAVE - 1));
                    if (lower) {
                        if (t0 < c) s0 = fma(-f, v0[j], s0);
                        if (t1 < c) s1 = fma(-f, v1[j], s1);
                    } else {
                        if (t0 > c && t0 < m) s0 = fma(-f, v0[j], s0);
                        if (t1 > c && t1 < m) s1 = fma(-f, v1[j], s1);
                    }
                }
            }
        }
        if (t0 < m) {
            if (udiag != nullptr && t0 == 0) udiag[i] = s0;
            else out[M.beg[i] + t0] = s0;
        }
        if (t1 < m) out[M.beg[i] + t1] = s1;
    }
    __syncthreads();
    return __any(bad);
}

// ISAI(L) / ISAI(U) rows of at most 32 entries: the wave takes row pairs
// (i0, i0 + 1) for i0 = p0, p0 + pstep, ... (several waves of one workgroup
// share a system by taking interleaved pairs), 32 lanes per row.  (Loading 8
// sweep columns ahead and reading the finished unknown with v_readlane was
// slower on the one-wave solve kernel: 15.8 against 14.4 ms at n = 144.)
__device__ int isai_tri(const DevIsai &M, int n, int type, const double *vals, double *out,
                        double *udiag, int lane, int p0 = 0, int pstep = 2)
{
    if (M.max_m > ISAI_GROUP) return isai_tri_wide(M, n, type, vals, out, udiag, lane, p0 / 2, pstep / 2);
    const int grp = lane / ISAI_GROUP, t = lane % ISAI_GROUP;
    int bad = 0;
    for (int i0 = p0; i0 < n; i0 += pstep) {
        const int i = i0 + grp;
        const bool act = i < n;
        const int m = act ? M.m[i] : 0;
        const int mo = __shfl_xor(m, ISAI_GROUP, WAVE);
        const int mmax = m > mo ? m : mo;
        const int *g = act ? M.gather + M.goff[i] : M.gather;
        double s;
        if (type == BILU_ISAI_LOWER) {
            s = (t == m - 1) ? 1.0 : 0.0;
            for (int c = mmax - 1; c >= 0; c--) {
                const bool on = c < m;
                if (on && t == c) {
                    const double d = view_val(vals, g[t * m + t]);
                    bad |= (d == 0.0);
                    s = s / d;
                }
                const double bot = __shfl(s, c, ISAI_GROUP);
                if (on && t < c) s = fma(-bot, view_val(vals, g[c * m + t]), s);
            }
        } else {
            s = (t == 0) ? 1.0 : 0.0;
            for (int c = 0; c < mmax; c++) {
                const bool on = c < m;
                if (on && t == c) {
                    const double d = view_val(vals, g[t * m + t]);
                    bad |= (d == 0.0);
                    s = s / d;
                }
                const double top = __shfl(s, c, ISAI_GROUP);
                if (on && t > c && t < m) s = fma(-top, view_val(vals, g[c * m + t]), s);
            }
        }
        if (act && t < m) {
            const int pos = M.beg[i] + t;
            if (udiag != nullptr && t == 0) udiag[i] = s;
            else out[pos] = s;
        }
    }
    __syncthreads();
    return __any(bad);
}

// DPP row_shl:S -- lane l reads lane l + S of its row of 16 lanes (a lane whose
// source falls outside the row reads 0; callers use only lanes l < 16 - S)
template <int S>
__device__ __forceinline__ int row_shl_i32(int v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x100 | S, 0xF, 0xF, true);
}

template <int S>
__device__ __forceinline__ double row_shl_f64(double v)
{
    return __hiloint2double(row_shl_i32<S>(__double2hiint(v)), row_shl_i32<S>(__double2loint(v)));
}

// DPP row_newbcast:N -- every lane reads lane N of its row of 16 lanes
template <int N>
__device__ __forceinline__ int row_bcast_i32(int v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x150 | N, 0xF, 0xF, false);
}

template <int N>
__device__ __forceinline__ double row_bcast_f64(double v)
{
    return __hiloint2double(row_bcast_i32<N>(__double2hiint(v)), row_bcast_i32<N>(__double2loint(v)));
}

// value of lane N of the lane's group of G lanes (G = 16: a DPP row; G = 32:
// rows 2r and 2r + 1 -- row_newbcast within each row, then the permlane16 swap
// hands the row that holds lane N to both rows)
template <int G, int N>
__device__ __forceinline__ int group_bcast_i32(int v)
{
    const int b = row_bcast_i32<N % 16>(v);
    if constexpr (G == 16) {
        return b;
    } else {
        const auto sw = __builtin_amdgcn_permlane16_swap(b, b, false, false);
        return N < 16 ? (int)sw[0] : (int)sw[1];
    }
}

template <int G, int N>
__device__ __forceinline__ double group_bcast(double v)
{
    return __hiloint2double(group_bcast_i32<G, N>(__double2hiint(v)), group_bcast_i32<G, N>(__double2loint(v)));
}

// int value of lane (lane ^ 16)
__device__ __forceinline__ int lane_xor16_i32(int v)
{
    const auto a = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (__lane_id() & 16) == 0 ? (int)a[1] : (int)a[0];
}

// f(std::integral_constant<int, I>()) for I = B .. E - 1 (an unrolled loop with constant indices)
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>());
        static_for<B + 1, E>(f);
    }
}

// General ISAI rows by Gauss-Jordan, one row per group of G lanes (lane t
// holding row t of the m x m block, m <= G): rows order[r0 .. r1) taken
// WAVE / G at a time, batch q = item0, item0 + istep, ...; steps and columns
// stop at the largest m of the wave's rows (beyond a row's m nothing changes
// its solution).  Every loop over the block's columns is unrolled with
// constant indices, so the block stays in registers.  UNR: the G steps are
// unrolled too, and every broadcast from the pivot lane is a DPP / permlane
// move (group_bcast) -- no LDS traffic.  Otherwise (the one-wave solve
// kernels, G = 32) the step loop is rolled, broadcasts go through ds_bpermute
// and the lane's entry in the pivot column is carried in cdp (picked up as the
// previous step's update passes it).  The pivot search is the reference's
// 32-lane shfl_down tree: only lanes below the distance are read by later
// steps, and each of them reads a source inside its own row of 16 lanes at
// distances 8, 4, 2, 1 (DPP row_shl).  At distance 16 a 32-lane group reads
// the other row (lanes t < 16 read t + 16: ds_bpermute, or the permlane16
// swap); a 16-lane group (m <= 16) would read lanes t + 16 >= m, which all
// hold the stand-in (dval, dp): a local comparison.
template <int G, bool UNR, typename TO = double>
__device__ __forceinline__ int isai_general_g(const DevIsai &M, const double *vals, TO *out, int lane, int r0, int r1,
                              int item0, int istep, double *pv = nullptr)
{
    static_assert(UNR || G == 32, "the rolled step loop is the 32-lane one");
    constexpr int RPW = WAVE / G;
    constexpr int CH = (UNR && G == 32) ? 4 : 8;  // columns per batch of loads and pivot-row broadcasts
    const int grp = lane / G, t = lane % G;
    int bad = 0;
    for (int q = item0; r0 + q * RPW < r1; q += istep) {
        const int ri = r0 + q * RPW + grp;
        const bool act = ri < r1;
        const int i = act ? M.order[ri] : 0;
        const int m = act ? M.m[i] : 0;
        int mx = m;
#pragma unroll
        for (int o = G; o < WAVE; o <<= 1) {
            const int mo = __shfl_xor(mx, o, WAVE);
            mx = mx > mo ? mx : mo;
        }
        const int mmax = __builtin_amdgcn_readfirstlane(mx);  // same on every lane
        const int *g = M.gather + (act ? M.goff[i] : 0);
        // unconditional loads at clamped positions (a batch in flight at once)
        const int last = m > 0 ? m * m - 1 : 0;
        double row[G];
#pragma unroll
        for (int k0 = 0; k0 < G; k0 += CH) {
            if (k0 < mmax) {
                int gi[CH];
#pragma unroll
                for (int j = 0; j < CH; j++) gi[j] = g[min((k0 + j) * m + t, last)];
#pragma unroll
                for (int j = 0; j < CH; j++) {
                    const double v = vals[max(gi[j], 0)];
                    row[k0 + j] = (t < m && k0 + j < m) ? (gi[j] >= 0 ? v : (gi[j] == -2 ? 1.0 : 0.0)) : 0.0;
                }
            } else {
#pragma unroll
                for (int j = 0; j < CH; j++) row[k0 + j] = 0.0;
            }
        }
        double rhs = (act && t == M.rhs[i]) ? 1.0 : 0.0;
        if constexpr (!UNR) {
            double cdp = row[0];
            for (int dp = 0; dp < mmax; dp++) {
                const bool on = dp < m;  // uniform within the group
                const double dval = __shfl(cdp, dp, G);
                double val = (t >= m || t < dp) ? dval : cdp;
                int piv = (t >= m || t < dp) ? dp : t;
                {
                    const double vo = __shfl_down(val, 16, G);
                    const int po = __shfl_down(piv, 16, G);
                    if (fabs(vo) > fabs(val)) { val = vo; piv = po; }
                }
#define BILU_PIVOT_DPP(S)                                                                          \
                {                                                                                  \
                    const double vo = row_shl_f64<S>(val);                                         \
                    const int po = row_shl_i32<S>(piv);                                            \
                    if (fabs(vo) > fabs(val)) { val = vo; piv = po; }                              \
                }
                BILU_PIVOT_DPP(8)
                BILU_PIVOT_DPP(4)
                BILU_PIVOT_DPP(2)
                BILU_PIVOT_DPP(1)
#undef BILU_PIVOT_DPP
                {
                    const int pa = __builtin_amdgcn_readlane(piv, 0), pb = __builtin_amdgcn_readlane(piv, G);
                    piv = grp ? pb : pa;
                }
                if (!on) piv = dp;
                double de = dval;
                if (piv != dp) {
#pragma unroll
                    for (int c = 0; c < G; c++) {
                        if (c < mmax) {
                            const double a = __shfl(row[c], dp, G);
                            const double b = __shfl(row[c], piv, G);
                            if (t == dp) row[c] = b;
                            if (t == piv) row[c] = a;
                        }
                    }
                    const double a = __shfl(rhs, dp, G), b = __shfl(rhs, piv, G);
                    if (t == dp) rhs = b;
                    if (t == piv) rhs = a;
                    const double ca = __shfl(cdp, dp, G), cb = __shfl(cdp, piv, G);
                    if (t == dp) cdp = cb;
                    if (t == piv) cdp = ca;
                    de = cb;  // the pivot row's entry, now on lane dp
                }
                if (on) {
                    bad |= (de == 0.0);
                    const double mult = cdp / de;
                    const bool upd = t != dp;
                    const int nx = dp + 1;
#pragma unroll
                    for (int c0 = 0; c0 < G; c0 += CH) {
                        if (c0 < mmax) {
                            double key[CH];
#pragma unroll
                            for (int j = 0; j < CH; j++) key[j] = __shfl(row[c0 + j], dp, G);
                            // columns at and beyond m never reach this row's solution, the
                            // pivot search or another row's columns below m: updated as well
                            if (upd) {
#pragma unroll
                                for (int j = 0; j < CH; j++) row[c0 + j] = fma(-mult, key[j], row[c0 + j]);
                            }
                            // the lane's entry in the next pivot column, after this step
                            if (nx >= c0 && nx < c0 + CH) {
#pragma unroll
                                for (int j = 0; j < CH; j++)
                                    if (nx == c0 + j) cdp = row[c0 + j];
                            }
                        }
                    }
                    const double rk = __shfl(rhs, dp, G);
                    if (upd) rhs = fma(-mult, rk, rhs);
                }
            }
        } else {
            // steps unrolled (dp a constant): every broadcast from lane dp is DPP
            // row_newbcast, across the two rows of a 32-lane group with a
            // permlane16 swap (group_bcast)
            static_for<0, G>([&](auto DP) {
                constexpr int dp = decltype(DP)::value;
                if (dp >= mmax) return;
                const bool on = dp < m;  // uniform within the group
                const double dval = group_bcast<G, dp>(row[dp]);
                double val = (t >= m || t < dp) ? dval : row[dp];
                int piv = (t >= m || t < dp) ? dp : t;
                if constexpr (G == 32) {
                    const double vo = lane_xor16(val);  // lanes t < 16 read t + 16
                    const int po = lane_xor16_i32(piv);
                    if (fabs(vo) > fabs(val)) { val = vo; piv = po; }
                } else {
                    if (fabs(dval) > fabs(val)) { val = dval; piv = dp; }
                }
#define BILU_PIVOT_DPP(S)                                                                          \
                {                                                                                  \
                    const double vo = row_shl_f64<S>(val);                                         \
                    const int po = row_shl_i32<S>(piv);                                            \
                    if (fabs(vo) > fabs(val)) { val = vo; piv = po; }                              \
                }
                BILU_PIVOT_DPP(8)
                BILU_PIVOT_DPP(4)
                BILU_PIVOT_DPP(2)
                BILU_PIVOT_DPP(1)
#undef BILU_PIVOT_DPP
                piv = group_bcast_i32<G, 0>(piv);
                if (!on) piv = dp;
                double de = dval;
                if (piv != dp) {
#pragma unroll
                    for (int c = 0; c < G; c++) {
                        if (c < mmax) {
                            const double a = group_bcast<G, dp>(row[c]);
                            const double b = __shfl(row[c], piv, G);
                            if (t == dp) row[c] = b;
                            if (t == piv) row[c] = a;
                        }
                    }
                    const double a = group_bcast<G, dp>(rhs), b = __shfl(rhs, piv, G);
                    if (t == dp) rhs = b;
                    if (t == piv) rhs = a;
                    de = group_bcast<G, dp>(row[dp]);  // the pivot row's entry, now on lane dp
                }
                if (on) {
                    bad |= (de == 0.0);
                    const double mult = row[dp] / de;
                    const bool upd = t != dp;
                    // the pivot row's columns from CL on go through the wave's LDS cells
                    // pv (stored by the pivot lane, read back as a broadcast), the rest by
                    // DPP -- the two pipes share the broadcasts (each alone sets the pace;
                    // CL measured: 32-lane rows 16 / 8 / 4 -> 10.4 / 8.6 / 8.5 ms, 16-lane
                    // rows 8 / 4 -> 4.15 / 4.56 ms at n = 144, 65536 systems)
                    constexpr int CL = G == 32 ? 4 : 8;
                    if constexpr (CL < G) {
                        if (t == dp) {
#pragma unroll
                            for (int c = CL; c < G; c++)
                                if (c < mmax) pv[grp * G + c] = row[c];
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    }
#pragma unroll
                    for (int c0 = 0; c0 < G; c0 += CH) {
                        if (c0 < mmax) {
                            double key[CH];
#pragma unroll
                            for (int j = 0; j < CH; j++)
                                key[j] = c0 + j >= CL ? pv[grp * G + c0 + j] : group_bcast<G, dp>(row[c0 + j]);
                            if (upd) {
#pragma unroll
                                for (int j = 0; j < CH; j++) row[c0 + j] = fma(-mult, key[j], row[c0 + j]);
                            }
                        }
                    }
                    const double rk = group_bcast<G, dp>(rhs);
                    if (upd) rhs = fma(-mult, rk, rhs);
                }
            });
        }
        double d = 1.0;
#pragma unroll
        for (int k = 0; k < G; k++)
            if (k == t) d = row[k];
        if (act && t < m) out[M.beg[i] + t] = (TO)(rhs / d);
    }
    return bad;
}

// every row of the map, two per wave in the map's row order (the fused
// one-wave solve kernels and the power-k kernel): wave w of nw (of one
// workgroup, or w = 0, nw = 1) takes every nw-th pair.  The batch ISAI
// (k_isai_gen) runs the rows with m <= 16 four per wave in a launch of its own.
__device__ __forceinline__ int isai_general(const DevIsai &M, int n, const double *vals, double *out, int lane, int w = 0,
                            int nw = 1)
{
    const int bad = isai_general_g<32, false>(M, vals, out, lane, 0, n, w, nw);
    __syncthreads();
    return __any(bad);
}

// ------------------------------------------------------------- the solver
// LDS layout (doubles): sA[nnz] | sF[nnz] | sW[nnz or 0] | sU[n or 0] | e0[n] e1[n] e2[n]
// (powers > 1: ... | sML[QL.nnz] | sMU[QU.nnz] | e0 e1 e2, the ISAI arrays on their own patterns)
template <int NR, int PC>
__global__ __launch_bounds__(WAVE) void k_solve(DevPattern P, int batch, const double *__restrict__ gA,
                                               const double *__restrict__ gb, double *__restrict__ gx,
                                               const double *__restrict__ gF, SolveParams sp,
                                               float *__restrict__ g_iter, int *__restrict__ g_conv,
                                               double *__restrict__ g_ires, double *__restrict__ g_tres,
                                               DevIsaiPow QL, DevIsaiPow QU)
{
    extern __shared__ double lds[];
    const int sys = blockIdx.x;
    if (sys >= batch) return;
    const int lane = threadIdx.x;
    const int n = P.n, nnz = P.nnz;
    const bool need_w = (sp.factor == FK_PARILU) || PC == PC_ISAI_LU || PC == PC_ISAI_GEN;
    double *sA = lds;
    double *sF = sA + nnz;
    double *sW = sF + nnz;
    double *sU = sW + (need_w ? nnz : 0);
    double *sML = sU + (PC == PC_ISAI_LU ? n : 0);
    double *sMU = sML + ((PC == PC_ISAI_POW || PC == PC_ISAI_GEN_POW) ? QL.nnz : 0);
    double *e0 = sMU + (PC == PC_ISAI_POW ? QU.nnz : 0);
    double *e1 = e0 + n;
    double *e2 = e1 + n;

    const double *Ap = gA + (size_t)sys * nnz;
    wave_copy(sA, Ap, nnz, lane);
    // the ISAI values generated by their own launches (gF: ISAI(A) on A's pattern,
    // or ISAI(L) / ISAI(A) on its chosen pattern; sp.m2: ISAI(U) on its pattern)
    constexpr bool ISAI_LOADABLE = PC == PC_ISAI_GEN || PC == PC_ISAI_POW || PC == PC_ISAI_GEN_POW || PC == PC_ISAI_LU;
    const bool loaded = ISAI_LOADABLE && sp.factor == FK_ISAI_LOADED;
    if (loaded) {
        if (PC == PC_ISAI_GEN) {
            wave_copy(sW, gF + (size_t)sys * nnz, nnz, lane);
        } else if (PC == PC_ISAI_LU) {
            // ISAI(L) strictly lower, ISAI(U) from the diagonal (the factor kernel's
            // layout): U's diagonal moves to sU, ISAI(L)'s unit diagonal takes its place
            wave_copy(sW, gF + (size_t)sys * nnz, nnz, lane);
            __syncthreads();
            for (int i = lane; i < n; i += WAVE) {
                sU[i] = sW[P.diag[i]];
                sW[P.diag[i]] = 1.0;
            }
        } else {
            wave_copy(sML, gF + (size_t)sys * QL.nnz, QL.nnz, lane);
            if (PC == PC_ISAI_POW) wave_copy(sMU, sp.m2 + (size_t)sys * QU.nnz, QU.nnz, lane);
        }
    } else if (sp.factor == FK_LOAD) {
        const double *Fp = gF + (size_t)sys * nnz;
        wave_copy(sF, Fp, nnz, lane);
    } else if (PC != PC_NONE) {
        wave_copy(sF, Ap, nnz, lane);
    }
    Vec<NR> b, x, r, rh, p, v, s, t, ph, sh;
    const double *bp = gb + (size_t)sys * n;
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int i = lane + WAVE * k;
        b.e[k] = (i < n) ? bp[i] : 0.0;
    }
    __syncthreads();

    // ---- preconditioner generation (fused: no HBM round trip) ----
    int bad = 0;
    if (loaded) {
        bad = sp.zstatus[sys] != 0;
    } else if (PC == PC_LU || PC == PC_ISAI_LU || PC == PC_ISAI_POW) {
        if (sp.factor == FK_ILU) {
            bad = ilu_factor(P, sF, lane);
        } else {
            if (sp.factor == FK_PARILU) parilu_sweeps(P, sA, sF, sW, sp.sweeps, lane);
            // a zero on U's diagonal is a zero pivot, whatever produced the factors
            __syncthreads();
            for (int i = lane; i < P.n; i += WAVE) bad |= (sF[P.diag[i]] == 0.0);
            bad = __any(bad);
        }
    }
    if (loaded) {
        // generated before this launch
    } else if (PC == PC_ISAI_LU && !bad) {
        bad |= isai_tri(P.isai[BILU_ISAI_LOWER], P.n, BILU_ISAI_LOWER, sF, sW, nullptr, lane);
        bad |= isai_tri(P.isai[BILU_ISAI_UPPER], P.n, BILU_ISAI_UPPER, sF, sW, sU, lane);
    } else if (PC == PC_ISAI_GEN) {
        bad |= isai_general(P.isai[BILU_ISAI_GENERAL], P.n, sA, sW, lane);
    } else if (PC == PC_ISAI_POW && !bad) {
        bad |= isai_tri(QL.map, n, BILU_ISAI_LOWER, sF, sML, nullptr, lane);
        bad |= isai_tri(QU.map, n, BILU_ISAI_UPPER, sF, sMU, nullptr, lane);
    } else if (PC == PC_ISAI_GEN_POW) {
        bad |= isai_general(QL.map, n, sA, sML, lane);
    }

    float iter = 0.0f;
    int conv = BILU_NOT_CONVERGED;
    double iter_res = 0.0;
#pragma unroll
    for (int k = 0; k < NR; k++) {
        x.e[k] = 0.0; v.e[k] = 0.0; p.e[k] = 0.0; r.e[k] = b.e[k]; rh.e[k] = b.e[k];
        ph.e[k] = 0.0; sh.e[k] = 0.0; s.e[k] = 0.0; t.e[k] = 0.0;
    }
    const double bnorm = sqrt(vdot(b, b, lane, n));
    const double res0 = bnorm;  // r0 = b - A*0
    const double tol_abs = (sp.tol_mode == BILU_TOL_RELATIVE) ? sp.tol * bnorm : sp.tol;
    iter_res = res0;

    auto apply = [&](const Vec<NR> &in, Vec<NR> &out) {
        if (PC == PC_NONE) {
            to_lds(in, e2, lane, n);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NR; k++) out.e[k] = in.e[k];
            return;
        }
        to_lds(in, e0, lane, n);
        __syncthreads();
        if (PC == PC_LU) lu_apply(P, sF, e0, e1, e2, lane);
        else if (PC == PC_ISAI_LU) isai_lu_apply(P, sW, sW, sU, e0, e1, e2, lane);
        else if (PC == PC_ISAI_POW) isai_pow_apply(QL, sML, QU, sMU, n, e0, e1, e2, lane);
        else if (PC == PC_ISAI_GEN_POW) spmv_pow(QL, sML, n, e0, e2, lane);
        else spmv_lds(P, sW, e0, e2, lane);
        from_lds(out, e2, lane, n);
    };

    if (bad) {
        conv = BILU_ZERO_PIVOT;
        iter_res = 0.0;
    } else if (!isfinite(bnorm)) {
        conv = BILU_NONFINITE;
    } else if (bnorm == 0.0) {
        conv = BILU_CONVERGED;
        iter_res = 0.0;
    } else if (res0 < tol_abs) {
        conv = BILU_CONVERGED;
    } else {
        double rho_old = 1.0, omega_old = 1.0, alpha = 1.0;
        int restarts = 0;
        while (iter < (float)sp.max_iter) {
            const double rho_new = vdot(r, rh, lane, n);
            if (rho_new == 0.0) break;
            const double beta = (rho_new / rho_old) * (alpha / omega_old);
#pragma unroll
            for (int k = 0; k < NR; k++) p.e[k] = fma(beta, fma(-omega_old, v.e[k], p.e[k]), r.e[k]);
            apply(p, ph);            // ph also sits in e2 for the gather
            spmv(P, sA, e2, v, lane);
            alpha = rho_new / vdot(rh, v, lane, n);
#pragma unroll
            for (int k = 0; k < NR; k++) s.e[k] = fma(-alpha, v.e[k], r.e[k]);
            iter_res = sqrt(vdot(s, s, lane, n));
            iter += 0.5f;
            if (!isfinite(iter_res)) { conv = BILU_NONFINITE; break; }
            const bool half_conv = iter_res < tol_abs;
            if (half_conv) {
#pragma unroll
                for (int k = 0; k < NR; k++) x.e[k] = fma(alpha, ph.e[k], x.e[k]);
            } else {
                __syncthreads();     // e2 (ph) reads by spmv done before reuse
                apply(s, sh);
                spmv(P, sA, e2, t, lane);
                const double ts = vdot(t, s, lane, n);
                const double tt = vdot(t, t, lane, n);
                const double omega_new = ts / tt;
#pragma unroll
                for (int k = 0; k < NR; k++) x.e[k] = fma(omega_new, sh.e[k], fma(alpha, ph.e[k], x.e[k]));
                iter += 0.5f;
#pragma unroll
                for (int k = 0; k < NR; k++) r.e[k] = fma(-omega_new, t.e[k], s.e[k]);
                iter_res = sqrt(vdot(r, r, lane, n));
                rho_old = rho_new;
                omega_old = omega_new;
                if (!isfinite(iter_res)) { conv = BILU_NONFINITE; break; }
            }
            __syncthreads();
            if (half_conv || iter_res < tol_abs) {
                conv = BILU_CONVERGED;
                if (!sp.restart || restarts >= sp.max_restarts) break;
                // true residual b - A x; restart from it if it misses tol
                to_lds(x, e0, lane, n);
                __syncthreads();
                Vec<NR> w;
                spmv(P, sA, e0, w, lane);
#pragma unroll
                for (int k = 0; k < NR; k++) w.e[k] = b.e[k] - w.e[k];
                const double tr = sqrt(vdot(w, w, lane, n));
                __syncthreads();
                if (tr < tol_abs) break;
                restarts++;
                conv = BILU_NOT_CONVERGED;
#pragma unroll
                for (int k = 0; k < NR; k++) { r.e[k] = w.e[k]; rh.e[k] = w.e[k]; p.e[k] = 0.0; v.e[k] = 0.0; }
                rho_old = 1.0; omega_old = 1.0; alpha = 1.0;
                iter_res = tr;
            }
        }
    }
    if (conv == BILU_ZERO_PIVOT) {
#pragma unroll
        for (int k = 0; k < NR; k++) x.e[k] = 0.0;
    }
    // fused epilogue: true residual ||b - A x|| from the LDS-resident A
    __syncthreads();
    to_lds(x, e0, lane, n);
    __syncthreads();
    Vec<NR> w;
    spmv(P, sA, e0, w, lane);
#pragma unroll
    for (int k = 0; k < NR; k++) w.e[k] = b.e[k] - w.e[k];
    const double tres = sqrt(vdot(w, w, lane, n));
    double *xp = gx + (size_t)sys * n;
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int i = lane + WAVE * k;
        if (i < n) xp[i] = x.e[k];
    }
    if (lane == 0) {
        g_iter[sys] = iter;
        g_conv[sys] = conv;
        if (g_ires) g_ires[sys] = iter_res;
        if (g_tres) g_tres[sys] = tres;
    }
}

// factor-only kernel (segregated path and the factor ABI):
// F <- ILU(0)(A) or ParILU sweeps from F_init (or A)
__global__ __launch_bounds__(WAVE) void k_factor(DevPattern P, int batch, const double *__restrict__ gA,
                                                const double *__restrict__ gInit, double *__restrict__ gF,
                                                int kind, int sweeps, int *__restrict__ g_status)
{
    extern __shared__ double lds[];
    const int sys = blockIdx.x;
    if (sys >= batch) return;
    const int lane = threadIdx.x, nnz = P.nnz;
    double *sA = lds, *sF = lds + nnz, *sW = sF + nnz;
    const double *Ap = gA + (size_t)sys * nnz;
    const double *Ip = gInit ? gInit + (size_t)sys * nnz : Ap;
    wave_copy(sA, Ap, nnz, lane);
    wave_copy(sF, Ip, nnz, lane);
    __syncthreads();
    int bad = 0;
    if (kind == FK_ILU) bad = ilu_factor(P, sF, lane);
    else if (kind == FK_ILU_ENTRIES) bad = ilu_factor_entries(P, sF, lane);
    else if (kind == FK_ILU_ERAS) bad = ilu_factor_eras(P, sF, lane);
    else parilu_sweeps(P, sA, sF, sW, sweeps, lane);
    __syncthreads();
    double *Fp = gF + (size_t)sys * nnz;
    for (int e = lane; e < nnz; e += WAVE) Fp[e] = sF[e];
    if (g_status && lane == 0) g_status[sys] = bad ? BILU_ZERO_PIVOT : 0;
}

// ISAI(L) or ISAI(U) of a batch of typed views; out positions outside the
// view untouched (the general ISAI: k_isai_gen).  One system per workgroup of
// ISAI_WAVES waves sharing the values in LDS; wave w takes row pairs
// (2w, 2w + 1), (2w + 2 ISAI_WAVES, ...), ...
constexpr int ISAI_WAVES = 4;

__global__ __launch_bounds__(WAVE * ISAI_WAVES) void k_isai(DevPattern P, int batch, const double *__restrict__ gV,
                                                           int type, double *__restrict__ gOut,
                                                           int *__restrict__ g_status)
{
    extern __shared__ double lds[];
    __shared__ int s_bad;
    const int sys = blockIdx.x;
    if (sys >= batch) return;
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE, nnz = P.nnz;
    constexpr int NT = WAVE * ISAI_WAVES;
    double *sV = lds, *sO = lds + nnz;
    const double *Vp = gV + (size_t)sys * nnz;
    double *Op = gOut + (size_t)sys * nnz;
    if (tid == 0) s_bad = 0;
    for (int e = tid; e < nnz; e += NT) sV[e] = Vp[e];
    for (int e = tid; e < nnz; e += NT) sO[e] = Op[e];
    __syncthreads();
    const int bad = isai_tri(P.isai[type], P.n, type, sV, sO, nullptr, lane, 2 * w, 2 * ISAI_WAVES);
    for (int e = tid; e < nnz; e += NT) Op[e] = sO[e];
    if (lane == 0 && bad) atomicOr(&s_bad, 1);
    __syncthreads();
    if (g_status && tid == 0) g_status[sys] = s_bad ? BILU_ZERO_PIVOT : 0;
}

// General ISAI of a batch, one launch per row group: rows order[r0 .. r1) of
// the map with G lanes per row (r0 = 0, r1 = n_pair with G = 32; then the
// rest with G = 16, a launch of its own so that its smaller register block
// sets its occupancy).  Workgroups as in k_isai.  The status of the second
// launch (or_status = 1) keeps a zero pivot the first one found.  TO = float:
// the FP64 values rounded once, as the FP32 preconditioner stores them.
template <int G, typename TO>
__global__ __launch_bounds__(WAVE * ISAI_WAVES) void k_isai_gen(DevIsai M, int nnz, int batch,
                                                               const double *__restrict__ gV,
                                                               TO *__restrict__ gOut, int ostride,
                                                               int *__restrict__ g_status, int r0, int r1,
                                                               int or_status)
{
    extern __shared__ double lds[];
    __shared__ int s_bad;
    const int sys = blockIdx.x;
    if (sys >= batch) return;
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    constexpr int NT = WAVE * ISAI_WAVES;
    const double *Vp = gV + (size_t)sys * nnz;
    if (tid == 0) s_bad = 0;
    for (int e = tid; e < nnz; e += NT) lds[e] = Vp[e];
    __syncthreads();
    // after the values: each wave's pivot-row cells (G per row group), zeroed
    // once (a cell a step does not store holds an earlier step's value)
    double *pv = lds + nnz + (size_t)w * WAVE;
    pv[lane] = 0.0;
    const int bad = isai_general_g<G, true, TO>(M, lds, gOut + (size_t)sys * ostride, lane, r0, r1, w, ISAI_WAVES, pv);
    if (__any(bad) && lane == 0) atomicOr(&s_bad, 1);
    __syncthreads();
    if (g_status && tid == 0) {
        const bool prev = or_status && g_status[sys] != 0;
        g_status[sys] = (s_bad || prev) ? BILU_ZERO_PIVOT : 0;
    }
}

// ISAI on a chosen pattern: sV[nnz] | sO[Q.nnz], workgroups as in k_isai
__global__ __launch_bounds__(WAVE * ISAI_WAVES) void k_isai_pow(DevPattern P, DevIsaiPow Q, int batch,
                                                               const double *__restrict__ gV, int type,
                                                               double *__restrict__ gOut, int *__restrict__ g_status)
{
    extern __shared__ double lds[];
    __shared__ int s_bad;
    const int sys = blockIdx.x;
    if (sys >= batch) return;
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE, nnz = P.nnz;
    constexpr int NT = WAVE * ISAI_WAVES;
    double *sV = lds, *sO = lds + nnz;
    const double *Vp = gV + (size_t)sys * nnz;
    double *Op = gOut + (size_t)sys * Q.nnz;
    if (tid == 0) s_bad = 0;
    for (int e = tid; e < nnz; e += NT) sV[e] = Vp[e];
    __syncthreads();
    int bad;
    if (type == BILU_ISAI_GENERAL) bad = isai_general(Q.map, P.n, sV, sO, lane, w, ISAI_WAVES);
    else bad = isai_tri(Q.map, P.n, type, sV, sO, nullptr, lane, 2 * w, 2 * ISAI_WAVES);
    for (int e = tid; e < Q.nnz; e += NT) Op[e] = sO[e];
    if (lane == 0 && bad) atomicOr(&s_bad, 1);
    __syncthreads();
    if (g_status && tid == 0) g_status[sys] = s_bad ? BILU_ZERO_PIVOT : 0;
}

// ||(L U - A)|_P||_F per system, F in place (unit-diagonal L)
__global__ __launch_bounds__(WAVE) void k_factor_residual(DevPattern P, int batch, const double *__restrict__ gA,
                                                         const double *__restrict__ gF, double *__restrict__ gRes)
{
    extern __shared__ double lds[];
    const int sys = blockIdx.x;
    if (sys >= batch) return;
    const int lane = threadIdx.x, nnz = P.nnz;
    double *sF = lds;
    const double *Fp = gF + (size_t)sys * nnz;
    const double *Ap = gA + (size_t)sys * nnz;
    for (int e = lane; e < nnz; e += WAVE) sF[e] = Fp[e];
    __syncthreads();
    double acc = 0.0;
    for (int e = lane; e < nnz; e += WAVE) {
        const int i = P.erow[e], j = P.ci[e];
        double sum = 0.0;
        const int de = P.dep_ptr[e + 1];
        for (int d = P.dep_ptr[e]; d < de; d++) {
            const int2 pq = P.dep[d];
            sum = fma(sF[pq.x], sF[pq.y], sum);
        }
        const double lu = (i > j) ? fma(sF[e], sF[P.diag[j]], sum) : sum + sF[e];
        const double d = lu - Ap[e];
        acc = fma(d, d, acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) gRes[sys] = sqrt(acc);
}

// values of a batch on the diagonal-augmented pattern: a gather with zero
// fill, HBM-bound (8 bytes read + 8 written per output entry; src is read
// from L2).  Grid-stride over batch*nnz_out entries, 64-bit indices.
__global__ void k_expand_values(int batch, int nnz, int nnz_out, const int *__restrict__ src,
                                const double *__restrict__ v, double *__restrict__ out)
{
    const size_t tot = (size_t)batch * (size_t)nnz_out;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const size_t s = i / (size_t)nnz_out;
        const int j = src[i - s * (size_t)nnz_out];
        out[i] = j < 0 ? 0.0 : v[s * (size_t)nnz + (size_t)j];
    }
}

__global__ void k_split_lu(DevPattern P, int batch, const double *__restrict__ gF,
                           double *__restrict__ gL, double *__restrict__ gU)
{
    const size_t tot_l = (size_t)batch * P.l_nnz, tot_u = (size_t)batch * P.u_nnz;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < tot_l || q < tot_u; q += stride) {
        if (q < tot_l && gL) {
            const size_t s = q / P.l_nnz;
            const int src = P.l_src[q % P.l_nnz];
            gL[q] = src < 0 ? 1.0 : gF[s * P.nnz + src];
        }
        if (q < tot_u && gU) {
            const size_t s = q / P.u_nnz;
            gU[q] = gF[s * P.nnz + P.u_src[q % P.u_nnz]];
        }
    }
}

template <int NR>
__global__ __launch_bounds__(WAVE) void k_true_residual(DevPattern P, int batch, const double *__restrict__ gA,
                                                       const double *__restrict__ gb, const double *__restrict__ gx,
                                                       double *__restrict__ gRes)
{
    extern __shared__ double lds[];
    const int sys = blockIdx.x;
    if (sys >= batch) return;
    const int lane = threadIdx.x, n = P.n, nnz = P.nnz;
    double *sA = lds, *sx = lds + nnz;
    wave_copy(sA, gA + (size_t)sys * nnz, nnz, lane);
    for (int i = lane; i < n; i += WAVE) sx[i] = gx[(size_t)sys * n + i];
    __syncthreads();
    Vec<NR> w;
    spmv(P, sA, sx, w, lane);
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int i = lane + WAVE * k;
        w.e[k] = (i < n) ? gb[(size_t)sys * n + i] - w.e[k] : 0.0;
    }
    const double r = sqrt(vdot(w, w, lane, n));
    if (lane == 0) gRes[sys] = r;
}

#include "fast_kernels.hpp"
#include "mixed_kernels.hpp"

}  // namespace

// ====================================================================== host
struct bilu_pattern_s {
    bilu::PatternAnalysis pa;
    DevPattern dev{};
    int *dbuf = nullptr;
    int device = 0;
    size_t max_lds = 65536;      // per workgroup
    size_t lds_per_cu = 65536;
    int num_cus = 1;
    // persistent-kernel tables
    uint8_t *d_fast = nullptr;     // all persistent-kernel tables, one allocation
    uint32_t *d_slot_meta = nullptr, *d_trL = nullptr, *d_trU = nullptr;
    int32_t *d_slot_pos = nullptr;
    uint16_t *d_cntL = nullptr, *d_cntU = nullptr;
    uint16_t *d_rlevL = nullptr, *d_rlevU = nullptr;
    uint16_t *d_etab = nullptr, *d_eptr = nullptr, *d_dep = nullptr;
    uint32_t *d_isL0 = nullptr, *d_isL1 = nullptr, *d_isU0 = nullptr, *d_isU1 = nullptr;
    uint32_t *d_isl_meta = nullptr, *d_isu_meta = nullptr;
    int32_t *d_isl_pos = nullptr, *d_isu_pos = nullptr;
    uint32_t *d_ilu0 = nullptr, *d_ilu1 = nullptr;
    uint32_t *d_parp = nullptr, *d_parf = nullptr;
    uint16_t *d_wcode = nullptr;
    // the FP32 mixed kernel's slot records (A, then ISAI(L), ISAI(U) layouts)
    // in their LDS layout, for the launches that read them from global memory
    // (ISAI templates of 48 and 72 slots; null otherwise)
    uint32_t *d_rec32 = nullptr;
    int *d_counter = nullptr;      // 256 bytes of work counters: [0] solve kernels, [32] factor kernel
    int ilu_sched = 0;             // exact ILU(0) schedule chosen by bilu_ilu0_tune (0: none yet)
    // solve workspace (factors between the factor and solve launches), grown on demand
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // ISAI maps of chosen patterns (powers > 1), built and uploaded on first use
    struct PowIsai {
        int type, power, max_m;
        std::vector<int> rp, ci;
        int *dbuf = nullptr;
        DevIsaiPow dev;
    };
    std::deque<PowIsai> pow_isai;    // stable addresses: callers keep pointers
};

namespace {

#define HIP_TRY(x)                                                                       \
    do {                                                                                 \
        hipError_t err_ = (x);                                                           \
        if (err_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "bilu: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(err_), \
                         __FILE__, __LINE__);                                            \
            return BILU_ERR_HIP;                                                         \
        }                                                                                \
    } while (0)

// (A slots, triangle slots) template of the mixed ISAI kernel for a pattern's
// slot counts (A's S, ISAI(L)'s SL, ISAI(U)'s SU); false when none fits
bool isai_tier(int S, int SL, int SU, int &sm, int &slu)
{
    const int tl = std::max(SL, SU);
    if (tl <= 0) return false;
    for (const auto &t : {std::make_pair(24, 20), std::make_pair(36, 24), std::make_pair(48, 32),
                          std::make_pair(72, 40)})
        if (S <= t.first && tl <= t.second) { sm = t.first; slu = t.second; return true; }
    return false;
}

// the mixed ISAI kernel's slot records as its LDS copy holds them: A's
// (FP32 x offset | FP32 output offset << 16, the scratch cell n + 1 for a
// slot that ends no row) padded to sm slots, then ISAI(L)'s and ISAI(U)'s
// without their row-end bit, padded to slu slots with zero-cell records
void mixed_isai_records(int n, const bilu::FastPlan &fp, int sm, int slu, std::vector<uint32_t> &rec)
{
    rec.assign((size_t)(sm + 2 * slu) * WAVE, 0);
    const uint32_t scratch = (uint32_t)(n + 1) * 4u;
    for (size_t i = 0; i < (size_t)sm * WAVE; i++) {
        const uint32_t m = i < fp.slot_meta.size() ? fp.slot_meta[i] : 0xFFFF0000u;
        const bool end = (m >> 16) != 0xFFFFu;
        const uint32_t yo = end ? (m >> 16) >> 1 : scratch;
        rec[i] = ((m & 0xFFFFu) >> 1) | (yo << 16);
    }
    const uint32_t pad = (uint32_t)n * 4u | (scratch << 16);
    for (size_t i = 0; i < (size_t)slu * WAVE; i++) {
        rec[(size_t)sm * WAVE + i] = i < fp.isl_meta.size() ? fp.isl_meta[i] & 0x7FFFFFFFu : pad;
        rec[(size_t)(sm + slu) * WAVE + i] = i < fp.isu_meta.size() ? fp.isu_meta[i] & 0x7FFFFFFFu : pad;
    }
}

int upload(bilu_pattern_s &h)
{
    const auto &pa = h.pa;
    std::vector<int> buf;
    auto put = [&](const std::vector<int> &v) {
        size_t off = buf.size();
        buf.insert(buf.end(), v.begin(), v.end());
        while (buf.size() % 4) buf.push_back(0);  // keep int2 arrays 8-byte aligned
        return off;
    };
    std::vector<int> erow(pa.nnz);
    for (int i = 0; i < pa.n; i++)
        for (int p = pa.rp[i]; p < pa.rp[i + 1]; p++) erow[p] = i;
    size_t o_rp = put(pa.rp), o_ci = put(pa.ci), o_diag = put(pa.diag), o_erow = put(erow);
    size_t o_llp = put(pa.lev_l_ptr), o_llr = put(pa.lev_l_rows), o_lup = put(pa.lev_u_ptr),
           o_lur = put(pa.lev_u_rows);
    size_t o_upp = put(pa.upd_ptr), o_upd = put(pa.upd.empty() ? std::vector<int>{0, 0} : pa.upd);
    size_t o_dpp = put(pa.dep_ptr), o_dep = put(pa.dep.empty() ? std::vector<int>{0, 0} : pa.dep);
    size_t o_ls = put(pa.l_src), o_us = put(pa.u_src);
    auto nz = [](const std::vector<int> &v) { return v.empty() ? std::vector<int>{0, 0} : v; };
    size_t o_lep = put(pa.lev_e_ptr), o_lee = put(nz(pa.lev_e_ent)), o_edp = put(pa.era_div_ptr),
           o_ed = put(nz(pa.era_div)), o_etp = put(pa.era_tgt_ptr), o_et = put(nz(pa.era_tgt)),
           o_tpp = put(pa.tgt_pair_ptr), o_tp = put(nz(pa.tgt_pair));
    size_t o_is[3][6];
    for (int t = 0; t < 3; t++) {
        const auto &M = pa.isai[t];
        o_is[t][0] = put(M.beg); o_is[t][1] = put(M.m); o_is[t][2] = put(M.goff);
        o_is[t][3] = put(M.gather.empty() ? std::vector<int>{0} : M.gather); o_is[t][4] = put(M.rhs_pos);
        o_is[t][5] = put(M.order);
    }
    HIP_TRY(hipMalloc((void **)&h.dbuf, buf.size() * sizeof(int)));
    HIP_TRY(hipMemcpy(h.dbuf, buf.data(), buf.size() * sizeof(int), hipMemcpyHostToDevice));
    DevPattern &d = h.dev;
    const int *B = h.dbuf;
    d.n = pa.n; d.nnz = pa.nnz;
    d.nlev_l = int(pa.lev_l_ptr.size()) - 1;
    d.nlev_u = int(pa.lev_u_ptr.size()) - 1;
    d.rp = B + o_rp; d.ci = B + o_ci; d.diag = B + o_diag; d.erow = B + o_erow;
    d.lev_l_ptr = B + o_llp; d.lev_l_rows = B + o_llr; d.lev_u_ptr = B + o_lup; d.lev_u_rows = B + o_lur;
    d.upd_ptr = B + o_upp; d.upd = reinterpret_cast<const int2 *>(B + o_upd);
    d.dep_ptr = B + o_dpp; d.dep = reinterpret_cast<const int2 *>(B + o_dep);
    d.l_src = B + o_ls; d.u_src = B + o_us;
    d.l_nnz = int(pa.l_ci.size()); d.u_nnz = int(pa.u_ci.size());
    d.nlev_e = int(pa.lev_e_ptr.size()) - 1;
    d.n_eras = int(pa.era_div_ptr.size()) - 1;
    d.lev_e_ptr = B + o_lep; d.lev_e_ent = B + o_lee;
    d.era_div_ptr = B + o_edp; d.era_div = reinterpret_cast<const int2 *>(B + o_ed);
    d.era_tgt_ptr = B + o_etp; d.era_tgt = B + o_et;
    d.tgt_pair_ptr = B + o_tpp; d.tgt_pair = reinterpret_cast<const int2 *>(B + o_tp);
    for (int t = 0; t < 3; t++) {
        d.isai[t].beg = B + o_is[t][0]; d.isai[t].m = B + o_is[t][1]; d.isai[t].goff = B + o_is[t][2];
        d.isai[t].gather = B + o_is[t][3]; d.isai[t].rhs = B + o_is[t][4];
        d.isai[t].order = B + o_is[t][5]; d.isai[t].n_pair = pa.isai[t].n_pair;
        d.isai[t].max_m = pa.isai[t].max_m;
    }
    HIP_TRY(hipMalloc((void **)&h.d_counter, 256));
    HIP_TRY(hipMemset(h.d_counter, 0, 256));
    const auto &fp = pa.fast;
    if (fp.ok) {
        std::vector<uint8_t> blob;
        auto put = [&](const void *src, size_t bytes) {
            while (blob.size() % 256) blob.push_back(0);
            const size_t off = blob.size();
            blob.insert(blob.end(), (const uint8_t *)src, (const uint8_t *)src + bytes);
            return off;
        };
        const size_t o_meta = put(fp.slot_meta.data(), fp.slot_meta.size() * 4);
        const size_t o_pos = put(fp.slot_pos.data(), fp.slot_pos.size() * 4);
        const size_t o_trL = put(fp.trL.data(), fp.trL.size() * 4);
        const size_t o_trU = put(fp.trU.data(), fp.trU.size() * 4);
        const size_t o_sL = put(fp.cntL.data(), fp.cntL.size() * 2);
        const size_t o_sU = put(fp.cntU.data(), fp.cntU.size() * 2);
        const size_t o_rL = put(fp.rlevL.data(), fp.rlevL.size() * 2);
        const size_t o_rU = put(fp.rlevU.data(), fp.rlevU.size() * 2);
        const size_t o_et = put(fp.etab.data(), fp.etab.size() * 2);
        const size_t o_ep = put(fp.eptr.data(), fp.eptr.size() * 2);
        const size_t o_dp = put(fp.dep.data(), fp.dep.size() * 2);
        const size_t o_iL0 = put(fp.isL0.data(), fp.isL0.size() * 4);
        const size_t o_iL1 = put(fp.isL1.data(), fp.isL1.size() * 4);
        const size_t o_iU0 = put(fp.isU0.data(), fp.isU0.size() * 4);
        const size_t o_iU1 = put(fp.isU1.data(), fp.isU1.size() * 4);
        const size_t o_lm = put(fp.isl_meta.data(), fp.isl_meta.size() * 4);
        const size_t o_um = put(fp.isu_meta.data(), fp.isu_meta.size() * 4);
        const size_t o_lp = put(fp.isl_pos.data(), fp.isl_pos.size() * 4);
        const size_t o_up = put(fp.isu_pos.data(), fp.isu_pos.size() * 4);
        const size_t o_i0 = put(fp.ilu0.data(), fp.ilu0.size() * 4);
        const size_t o_i1 = put(fp.ilu1.data(), fp.ilu1.size() * 4);
        const size_t o_p0 = put(fp.parp.data(), fp.parp.size() * 4);
        const size_t o_p1 = put(fp.parf.data(), fp.parf.size() * 4);
        const size_t o_wc = put(fp.wcode.data(), fp.wcode.size() * 2);
        // the mixed ISAI kernel's records in global memory (its LDS copy's layout)
        int sm = 0, slu = 0;
        size_t o_r32 = 0;
        const bool r32 = isai_tier(fp.S, fp.SL, fp.SU, sm, slu) && sm >= 48;
        if (r32) {
            std::vector<uint32_t> rec;
            mixed_isai_records(pa.n, fp, sm, slu, rec);
            o_r32 = put(rec.data(), rec.size() * 4);
        }
        HIP_TRY(hipMalloc((void **)&h.d_fast, blob.size()));
        HIP_TRY(hipMemcpy(h.d_fast, blob.data(), blob.size(), hipMemcpyHostToDevice));
        uint8_t *b = h.d_fast;
        h.d_slot_meta = (uint32_t *)(b + o_meta); h.d_slot_pos = (int32_t *)(b + o_pos);
        h.d_trL = (uint32_t *)(b + o_trL); h.d_trU = (uint32_t *)(b + o_trU);
        h.d_cntL = (uint16_t *)(b + o_sL); h.d_cntU = (uint16_t *)(b + o_sU);
        h.d_rlevL = (uint16_t *)(b + o_rL); h.d_rlevU = (uint16_t *)(b + o_rU);
        h.d_etab = (uint16_t *)(b + o_et); h.d_eptr = (uint16_t *)(b + o_ep); h.d_dep = (uint16_t *)(b + o_dp);
        h.d_isL0 = (uint32_t *)(b + o_iL0); h.d_isL1 = (uint32_t *)(b + o_iL1);
        h.d_isU0 = (uint32_t *)(b + o_iU0); h.d_isU1 = (uint32_t *)(b + o_iU1);
        h.d_isl_meta = (uint32_t *)(b + o_lm); h.d_isu_meta = (uint32_t *)(b + o_um);
        h.d_isl_pos = (int32_t *)(b + o_lp); h.d_isu_pos = (int32_t *)(b + o_up);
        h.d_ilu0 = (uint32_t *)(b + o_i0); h.d_ilu1 = (uint32_t *)(b + o_i1);
        h.d_parp = (uint32_t *)(b + o_p0); h.d_parf = (uint32_t *)(b + o_p1);
        h.d_wcode = (uint16_t *)(b + o_wc);
        h.d_rec32 = r32 ? (uint32_t *)(b + o_r32) : nullptr;
    }
    return BILU_OK;
}

void free_handle(bilu_pattern_s *h)
{
    if (!h) return;
    if (h->dbuf) (void)hipFree(h->dbuf);
    if (h->d_counter) (void)hipFree(h->d_counter);
    if (h->ws) (void)hipFree(h->ws);
    if (h->d_fast) (void)hipFree(h->d_fast);
    for (auto &q : h->pow_isai)
        if (q.dbuf) (void)hipFree(q.dbuf);
    delete h;
}

// the ISAI map of |P|^power for `type`, built and uploaded at its first use
int get_pow_isai(bilu_pattern_s &h, int type, int power, const bilu_pattern_s::PowIsai *&out)
{
    for (auto &q : h.pow_isai)
        if (q.type == type && q.power == power) { out = &q; return BILU_OK; }
    bilu_pattern_s::PowIsai q;
    q.type = type;
    q.power = power;
    bilu::isai_power_pattern(h.pa, type, power, q.rp, q.ci);
    bilu::IsaiMap M;
    bilu::build_isai_on(h.pa, type, q.rp, q.ci, M);
    q.max_m = M.max_m;
    const int n = h.pa.n;
    std::vector<int> buf;
    auto put = [&](const std::vector<int> &v) { const size_t o = buf.size(); buf.insert(buf.end(), v.begin(), v.end()); return o; };
    const size_t o_rp = put(q.rp), o_ci = put(q.ci), o_beg = put(M.beg), o_m = put(M.m), o_goff = put(M.goff),
                 o_g = put(M.gather), o_rhs = put(M.rhs_pos), o_ord = put(M.order);
    HIP_TRY(hipSetDevice(h.device));
    HIP_TRY(hipMalloc((void **)&q.dbuf, std::max<size_t>(buf.size(), 1) * sizeof(int)));
    HIP_TRY(hipMemcpy(q.dbuf, buf.data(), buf.size() * sizeof(int), hipMemcpyHostToDevice));
    q.dev.nnz = q.rp[n];
    q.dev.rp = q.dbuf + o_rp; q.dev.ci = q.dbuf + o_ci;
    q.dev.map.beg = q.dbuf + o_beg; q.dev.map.m = q.dbuf + o_m; q.dev.map.goff = q.dbuf + o_goff;
    q.dev.map.gather = q.dbuf + o_g; q.dev.map.rhs = q.dbuf + o_rhs;
    q.dev.map.order = q.dbuf + o_ord; q.dev.map.n_pair = M.n_pair; q.dev.map.max_m = M.max_m;
    h.pow_isai.push_back(std::move(q));
    out = &h.pow_isai.back();
    return BILU_OK;
}

// table pointers of the persistent kernels (no per-launch sizing)
void fast_tables(const bilu_pattern_s &h, FastParams &fp)
{
    const auto &pl = h.pa.fast;
    fp.trL = h.d_trL; fp.trU = h.d_trU;
    fp.trL_words = int(pl.trL.size()); fp.trU_words = int(pl.trU.size());
    fp.cntL = h.d_cntL; fp.cntU = h.d_cntU;
    fp.rlevL = h.d_rlevL; fp.rlevU = h.d_rlevU;
    fp.etab = reinterpret_cast<const ushort4 *>(h.d_etab); fp.eptr = h.d_eptr;
    fp.wcode = h.d_wcode;
    fp.dep = reinterpret_cast<const ushort2 *>(h.d_dep);
    fp.diag = h.dev.diag;
    fp.nlev_l = h.dev.nlev_l; fp.nlev_u = h.dev.nlev_u; fp.nlev_e = pl.nlev_e; fp.ne = pl.ne;
    fp.S = pl.S; fp.slot_meta = h.d_slot_meta; fp.slot_pos = h.d_slot_pos;
    fp.counter = h.d_counter;
    fp.isL0 = h.d_isL0; fp.isL1 = h.d_isL1; fp.isU0 = h.d_isU0; fp.isU1 = h.d_isU1;
    fp.isL_steps = pl.isL_steps; fp.isU_steps = pl.isU_steps;
    fp.isl_meta = h.d_isl_meta; fp.isu_meta = h.d_isu_meta; fp.isl_pos = h.d_isl_pos; fp.isu_pos = h.d_isu_pos;
    fp.SL = pl.SL; fp.SU = pl.SU;
    fp.ilu0 = h.d_ilu0; fp.ilu1 = h.d_ilu1; fp.ilu_steps = pl.ilu_steps;
    fp.parp = h.d_parp; fp.parf = h.d_parf; fp.par_rounds = pl.par_rounds; fp.par_psteps = pl.par_psteps;
    fp.rec32 = h.d_rec32;
}

// Kernel timing (bilu_timing_begin / _end): while a window is open, the
// launches of one kernel kind are bracketed by HIP events recorded on their
// stream, from a pool created when the window opens.
struct Timing {
    std::mutex mu;
    int kind = 0;                                  // BILU_TIMING_* or 0 (closed)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t used = 0;
};
Timing g_timing;

// record the start (before = true) or the stop event of a kernel of `kind` on st
void timing_mark(int kind, hipStream_t st, bool before)
{
    if (g_timing.kind != kind) return;
    std::lock_guard<std::mutex> lock(g_timing.mu);
    if (g_timing.kind != kind) return;
    if (before) {
        if (g_timing.used < g_timing.ev.size()) (void)hipEventRecord(g_timing.ev[g_timing.used].first, st);
    } else {
        if (g_timing.used < g_timing.ev.size()) (void)hipEventRecord(g_timing.ev[g_timing.used].second, st);
        g_timing.used++;
    }
}

template <int NR, int SM, int PC>
int launch_fast_t(bilu_pattern_s &h, const FastParams &fp, int grid, int threads, size_t lds, int batch,
                  const double *A, const double *b, double *x, const double *F, const SolveParams &sp,
                  float *it, int *cv, double *ir, double *tr, const int *Z, hipStream_t st)
{
    auto kern = k_fast<NR, SM, PC>;
    if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, fp.counter);
    timing_mark(BILU_TIMING_SOLVE, st, true);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, st, h.dev, fp, batch, A, b, x, F, sp, it, cv, ir, tr, Z);
    timing_mark(BILU_TIMING_SOLVE, st, false);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

// the persistent FP64 solve; pc = FPC_LU (F: factors, or null to factor in
// the kernel), FPC_ISAI / FPC_GEN (F: ISAI values on the pattern, Z: their
// zero-pivot flags).  Returns -1 when the kernel cannot take this
// pattern/option set; dry: only that check, nothing launched.
int launch_fast(bilu_pattern_s &h, int batch, const double *A, const double *b, double *x, const double *F,
                const SolveParams &sp, float *it, int *cv, double *ir, double *tr, hipStream_t st,
                int pc = FPC_LU, const int *Z = nullptr, bool dry = false)
{
    const auto &pl = h.pa.fast;
    const int n = h.pa.n, nnz = h.pa.nnz;
    const int nr = (n + WAVE - 1) / WAVE;
    if (!pl.ok || nr > 5) return -1;   // the instantiated row counts
    int sm;
    if (pl.S <= 24) sm = 24;
    else if (pl.S <= 36) sm = 36;
    else if (pl.S <= 48) sm = 48;
    else return -1;
    if (pc == FPC_ISAI && (pl.SL <= 0 || pl.SU <= 0)) return -1;
    const bool par = pc == FPC_LU && sp.factor == FK_PARILU;
    if (pl.nr != nr) return -1;
    FastParams fp{};
    fast_tables(h, fp);
    // per wave: LU: F (+2 cells), (ParILU: the previous sweep's factors); ISAI: the
    // coefficients slot-major; then e0, e1 (+2 cells each); LU: the row
    // accumulators of the solves (64 NR)
    fp.f_cells = pc == FPC_LU ? (par ? 2 : 1) * (nnz + 2) : pc == FPC_ISAI ? (pl.SL + pl.SU) * WAVE : pl.S * WAVE;
    fp.wave_bytes = int(((size_t)fp.f_cells + 2 * (n + 2) + (pc == FPC_LU ? 64 * nr : 0)) * sizeof(double));
    fp.wave_bytes = (fp.wave_bytes + 15) / 16 * 16;
    fp.par_w = par ? 1 : 0;
    fp.counter = h.d_counter;
    auto a16 = [](size_t x) { return (x + 15) / 16 * 16; };
    fp.o_trL = fp.o_trU = 0;
    fp.shared_bytes = int(a16((size_t)(sm + (pc == FPC_ISAI ? pl.SL + pl.SU : 0)) * WAVE * 4));
    if ((size_t)fp.shared_bytes >= h.max_lds) return -1;
    int waves = (int)std::min<size_t>(KF_MAX_WAVES, (h.max_lds - fp.shared_bytes) / fp.wave_bytes);
    if (waves < 1) return -1;
    waves = std::min(waves, std::max(1, batch));
    const size_t lds = fp.shared_bytes + (size_t)waves * fp.wave_bytes;
    const int per_cu = std::max<int>(1, (int)(h.lds_per_cu / lds));
    int grid = h.num_cus * per_cu;
    grid = std::max(1, std::min(grid, (batch + waves - 1) / waves));
    const int threads = waves * WAVE;
    if (dry) return BILU_OK;
#define FAST_CASE(R, M)                                                                                       \
    if (nr == R && sm == M) {                                                                                 \
        if (pc == FPC_ISAI)                                                                                   \
            return launch_fast_t<R, M, FPC_ISAI>(h, fp, grid, threads, lds, batch, A, b, x, F, sp, it, cv, ir, \
                                                 tr, Z, st);                                                  \
        if (pc == FPC_GEN)                                                                                    \
            return launch_fast_t<R, M, FPC_GEN>(h, fp, grid, threads, lds, batch, A, b, x, F, sp, it, cv, ir,  \
                                                tr, Z, st);                                                   \
        return launch_fast_t<R, M, FPC_LU>(h, fp, grid, threads, lds, batch, A, b, x, F, sp, it, cv, ir, tr,   \
                                           Z, st);                                                            \
    }
#define FAST_ROW(R) FAST_CASE(R, 24) FAST_CASE(R, 36) FAST_CASE(R, 48)
    FAST_ROW(1) FAST_ROW(2) FAST_ROW(3) FAST_ROW(4) FAST_ROW(5)
#undef FAST_ROW
#undef FAST_CASE
    return -1;
}

__global__ void k_fill_int(int *p, int v, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int slot_template(int S)
{
    return S <= 24 ? 24 : S <= 36 ? 36 : S <= 48 ? 48 : -1;
}

// grid of a persistent kernel: every CU filled with as many workgroups as
// its LDS holds, never more waves than systems
int persistent_grid(const bilu_pattern_s &h, size_t lds, int waves, int batch)
{
    const int per_cu = std::max<int>(1, (int)(h.lds_per_cu / lds));
    return std::max(1, std::min(h.num_cus * per_cu, (batch + waves - 1) / waves));
}

// workspace of at least `bytes` (grown by reallocation; the handle is used by
// one host thread at a time, so growing waits for the device first).  A
// stream under graph capture cannot wait or allocate: growing is refused
// there (bilu_pattern_reserve sizes the workspace beforehand).
int workspace(bilu_pattern_s &h, size_t bytes, void **out, hipStream_t st)
{
    if (h.ws_bytes < bytes) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIP_TRY(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone) {
            std::fprintf(stderr, "bilu: the solve workspace needs %zu bytes (%zu reserved) inside a graph "
                                 "capture; call bilu_pattern_reserve first\n", bytes, h.ws_bytes);
            return BILU_ERR_INVALID_ARG;
        }
        if (h.ws) {
            HIP_TRY(hipDeviceSynchronize());
            HIP_TRY(hipFree(h.ws));
            h.ws = nullptr;
            h.ws_bytes = 0;
        }
        HIP_TRY(hipMalloc(&h.ws, bytes));
        h.ws_bytes = bytes;
    }
    *out = h.ws;
    return BILU_OK;
}

// workspace bytes a solve of `batch` systems needs on any path: FP32 factors
// or ISAI values + per-system flags (mixed precision, the learned solve's
// chunks), or FP64 factors (the FP64 split and segregated paths)
size_t workspace_need(const bilu_pattern_s &h, int batch)
{
    const size_t nnz = (size_t)h.pa.nnz, b = (size_t)std::max(batch, 0);
    return ((b * nnz * sizeof(double) + 255) / 256) * 256 + b * sizeof(int);
}

// persistent factorization: exact ILU(0) or ParILU(0) sweeps (optionally
// from F_init, optionally stopped on the factor residual) -> F (TO)
// isai: write ISAI(L) / ISAI(U) of the factors (TO) to M and zero-pivot
// flags to Z instead of the factors
template <typename TO>
int launch_factor_fast(bilu_pattern_s &h, int batch, const double *A, const double *F_init, TO *F,
                       int mode, int sweeps, double stop_tol, int *status, hipStream_t st,
                       int isai = 0, void *M = nullptr, int *Z = nullptr, const int *sysidx = nullptr,
                       int max_waves = KF_MAX_WAVES, const _Float16 *Y = nullptr, int P3 = 0, int inplace = 0)
{
    const auto &pl = h.pa.fast;
    if (!pl.ok) return BILU_ERR_TOO_LARGE;
    const int nnz = h.pa.nnz;
    FastParams fp{};
    fast_tables(h, fp);
    fp.counter = h.d_counter + 32;
    // F (FP64; FP32 ISAI values share it once the factors are rounded to FP32), then
    // the Jacobi ParILU buffer (FP64), which FP64 ISAI values take afterwards
    const bool isai64 = isai && sizeof(TO) == 8;
    const size_t second = isai64 ? (size_t)(nnz + 4) * sizeof(double)
                                 : mode == FS_PARILU && !inplace ? (size_t)(nnz + 2) * sizeof(double) : 0;
    // the fused warm start's FP32 diagonal at the end of the region
    fp.dg_floats = Y ? (h.pa.n + 3) / 4 * 4 : 0;
    fp.wave_bytes = int(((size_t)(nnz + 4) * sizeof(double) + second + (size_t)fp.dg_floats * 4 + 15) / 16 * 16);
    const int wcap = std::min(max_waves, KF_MAX_WAVES);
    int waves = (int)std::min<size_t>(wcap, h.max_lds / fp.wave_bytes);
    if (waves < 1) return BILU_ERR_TOO_LARGE;
    if (mode == FS_PARILU && sweeps > 0 && pl.par_rounds > 0) {
        // the sweep records in a shared LDS region when that costs no wave
        const size_t rec = ((size_t)(pl.par_psteps + pl.par_rounds) * WAVE * 4 + 15) / 16 * 16;
        if (rec < h.max_lds && (int)std::min<size_t>(wcap, (h.max_lds - rec) / fp.wave_bytes) >= waves)
            fp.shared_bytes = int(rec);
    }
    waves = std::min(waves, std::max(1, batch));
    const size_t lds = fp.shared_bytes + (size_t)waves * fp.wave_bytes;
    const int grid = persistent_grid(h, lds, waves, batch);
    auto kern = k_factor_fast<TO>;
    if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, fp.counter);
    timing_mark(BILU_TIMING_FACTOR, st, true);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(waves * WAVE), lds, st, fp, h.pa.n, nnz, batch, A, F_init, F,
                       mode, sweeps, stop_tol, status, isai, M, Z, sysidx, Y, P3, inplace);
    timing_mark(BILU_TIMING_FACTOR, st, false);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

// indirection of the jobs of one launch (uneven request lists): job j uses
// system sysidx[j] and writes results at outidx[j]; null = identity
struct Jobs {
    const int *sysidx = nullptr, *outidx = nullptr;
};

template <int NR, int SM, int PCK, int SLU, int RG = 0>
int launch_mixed_t(bilu_pattern_s &h, const FastParams &fp, int grid, int waves, size_t lds, int batch,
                   const double *A, const double *b, double *x, const float *F, const int *Z,
                   const MixedParams &mp, float *it, int *cv, double *ir, double *tr, int *refine,
                   hipStream_t st, Jobs jobs, const FactorArgs *fa)
{
    auto kern = k_mixed<NR, SM, PCK, SLU, 0, RG>;
    if constexpr (PCK != PCK_GEN) {
        if (fa) kern = k_mixed<NR, SM, PCK, SLU, 1, RG>;
    }
    if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, fp.counter);
    timing_mark(BILU_TIMING_SOLVE, st, true);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(waves * WAVE), lds, st, h.dev, fp, batch, A, b, x, F, Z, mp, it, cv,
                       ir, tr, refine, jobs.sysidx, jobs.outidx, fa ? *fa : FactorArgs{});
    timing_mark(BILU_TIMING_SOLVE, st, false);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

// the ISAI instantiations: records in global memory only for the 48- and
// 72-slot templates (the smaller ones are bound by registers, not LDS)
template <int NR, int SM, int SLU>
int launch_mixed_isai(bool rg, bilu_pattern_s &h, const FastParams &fp, int grid, int waves, size_t lds, int batch,
                      const double *A, const double *b, double *x, const float *F, const int *Z,
                      const MixedParams &mp, float *it, int *cv, double *ir, double *tr, int *refine,
                      hipStream_t st, Jobs jobs, const FactorArgs *fa)
{
    if constexpr (SM >= 48) {
        if (rg)
            return launch_mixed_t<NR, SM, PCK_ISAI, SLU, 1>(h, fp, grid, waves, lds, batch, A, b, x, F, Z, mp, it, cv,
                                                             ir, tr, refine, st, jobs, fa);
    }
    return launch_mixed_t<NR, SM, PCK_ISAI, SLU, 0>(h, fp, grid, waves, lds, batch, A, b, x, F, Z, mp, it, cv, ir,
                                                     tr, refine, st, jobs, fa);
}

// SpMV slot templates of the mixed kernel (FP32 A: up to 72 slots per lane)
int slot_template_mixed(int S)
{
    return S <= 24 ? 24 : S <= 36 ? 36 : S <= 48 ? 48 : S <= 72 ? 72 : -1;
}

// FP32 inner / FP64 refinement solve over FP32 factors F (pck = PCK_LU) or
// FP32 ISAI values F with zero-pivot flags Z (pck = PCK_ISAI); at most
// max_waves waves per workgroup.  fa non-null: the merged kernel, which
// factors each system itself (F and Z unused)
int launch_mixed(bilu_pattern_s &h, int batch, const double *A, const double *b, double *x, const float *F,
                 const int *Z, int pck, const MixedParams &mp, float *it, int *cv, double *ir, double *tr,
                 int *refine, hipStream_t st, Jobs jobs = Jobs(), int max_waves = 64, const FactorArgs *fa = nullptr,
                 int rec_global = 0)
{
    const auto &pl = h.pa.fast;
    const int n = h.pa.n, nnz = h.pa.nnz;
    const int nr = (n + WAVE - 1) / WAVE;
    int sm = slot_template_mixed(pl.S), slu = 1;
    if (!pl.ok || nr > MAX_NR || pl.nr != nr || sm < 0) return BILU_ERR_TOO_LARGE;
    if (pck == PCK_ISAI && !isai_tier(pl.S, pl.SL, pl.SU, sm, slu)) return BILU_ERR_TOO_LARGE;
    if (pck == PCK_GEN && fa) return BILU_ERR_INVALID_ARG;   // generated by its own launch
    // records in global memory: the ISAI templates of 48 and 72 slots, when asked
    const bool rg = rec_global && pck == PCK_ISAI && sm >= 48 && h.d_rec32;
    FastParams fp{};
    fast_tables(h, fp);
    // per wave: LU: FP32 factors (+2 cells, 16-aligned), two (n + 2)-double vector regions,
    // FP64 row accumulators of the solves (64 NR); ISAI: A (FP32, slot-major), the two vector regions
    fp.wave_bytes = pck == PCK_LU
        ? int(((size_t)(nnz + 2) * 4 + 15) / 16 * 16 + (size_t)2 * (n + 2) * 8 + (size_t)64 * nr * 8)
        : int((size_t)sm * WAVE * 4 + (size_t)2 * (n + 2) * 8);
    if (fa) {
        // merged: the factorization's region (FP64 factors, the Jacobi buffer, the
        // warm start's FP32 diagonal at the end) when it is the larger
        fp.dg_floats = fa->Y ? (n + 3) / 4 * 4 : 0;
        const size_t fr = (size_t)(nnz + 4) * 8 +
                          (fa->mode == FS_PARILU && !fa->inplace ? (size_t)(nnz + 2) * 8 : 0) +
                          (size_t)fp.dg_floats * 4;
        fp.wave_bytes = std::max<int>(fp.wave_bytes, (int)fr);
    }
    fp.wave_bytes = (fp.wave_bytes + 15) / 16 * 16;
    // shared LDS: A's slot records (ISAI: then the two triangle layouts' records),
    // unless they are read from global memory
    const size_t rec = rg ? 0 : (size_t)(sm + (pck == PCK_ISAI ? 2 * slu : 0)) * WAVE * 4;
    if (rec >= h.max_lds) return BILU_ERR_TOO_LARGE;
    const int wcap = std::min(max_waves, km_waves(sm, pck, fa ? 1 : 0));
    fp.shared_bytes = int(rec);
    int waves = (int)std::min<size_t>(wcap, (h.max_lds - rec) / fp.wave_bytes);
    if (waves < 1) return BILU_ERR_TOO_LARGE;
    waves = std::min(waves, std::max(1, batch));
    const size_t lds = fp.shared_bytes + (size_t)waves * fp.wave_bytes;
    const int grid = persistent_grid(h, lds, waves, batch);
#define MIXED_CASE(R, M, T) \
    if (nr == R && sm == M) \
        return pck == PCK_LU \
            ? launch_mixed_t<R, M, PCK_LU, 1>(h, fp, grid, waves, lds, batch, A, b, x, F, Z, mp, it, cv, ir, tr, refine, st, jobs, fa) \
            : pck == PCK_GEN \
            ? launch_mixed_t<R, M, PCK_GEN, 1>(h, fp, grid, waves, lds, batch, A, b, x, F, Z, mp, it, cv, ir, tr, refine, st, jobs, nullptr) \
            : launch_mixed_isai<R, M, T>(rg, h, fp, grid, waves, lds, batch, A, b, x, F, Z, mp, it, cv, ir, tr, refine, st, jobs, fa);
#define MIXED_ROW(R) MIXED_CASE(R, 24, 20) MIXED_CASE(R, 36, 24) MIXED_CASE(R, 48, 32) MIXED_CASE(R, 72, 40)
    MIXED_ROW(1) MIXED_ROW(2) MIXED_ROW(3) MIXED_ROW(4) MIXED_ROW(5)
#undef MIXED_ROW
#undef MIXED_CASE
    return BILU_ERR_TOO_LARGE;
}

template <int NR, int PC>
int launch_solve_nr(bilu_pattern_s &h, int batch, const double *A, const double *b, double *x,
                    const double *F, const SolveParams &sp, float *it, int *cv, double *ir,
                    double *tr, size_t lds, hipStream_t st, const DevIsaiPow &QL, const DevIsaiPow &QU)
{
    auto kern = k_solve<NR, PC>;
    if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    timing_mark(BILU_TIMING_SOLVE, st, true);
    hipLaunchKernelGGL(kern, dim3(batch), dim3(WAVE), lds, st, h.dev, batch, A, b, x, F, sp, it, cv, ir, tr, QL, QU);
    timing_mark(BILU_TIMING_SOLVE, st, false);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

template <int PC>
int launch_solve(bilu_pattern_s &h, int batch, const double *A, const double *b, double *x,
                 const double *F, const SolveParams &sp, float *it, int *cv, double *ir, double *tr,
                 size_t lds, hipStream_t st, const DevIsaiPow &QL = DevIsaiPow{}, const DevIsaiPow &QU = DevIsaiPow{})
{
    const int nr = (h.pa.n + WAVE - 1) / WAVE;
    switch (nr) {
    case 1: return launch_solve_nr<1, PC>(h, batch, A, b, x, F, sp, it, cv, ir, tr, lds, st, QL, QU);
    case 2: return launch_solve_nr<2, PC>(h, batch, A, b, x, F, sp, it, cv, ir, tr, lds, st, QL, QU);
    case 3: return launch_solve_nr<3, PC>(h, batch, A, b, x, F, sp, it, cv, ir, tr, lds, st, QL, QU);
    case 4: return launch_solve_nr<4, PC>(h, batch, A, b, x, F, sp, it, cv, ir, tr, lds, st, QL, QU);
    case 5: return launch_solve_nr<5, PC>(h, batch, A, b, x, F, sp, it, cv, ir, tr, lds, st, QL, QU);
    case 6: return launch_solve_nr<6, PC>(h, batch, A, b, x, F, sp, it, cv, ir, tr, lds, st, QL, QU);
    case 7: return launch_solve_nr<7, PC>(h, batch, A, b, x, F, sp, it, cv, ir, tr, lds, st, QL, QU);
    case 8: return launch_solve_nr<8, PC>(h, batch, A, b, x, F, sp, it, cv, ir, tr, lds, st, QL, QU);
    default: return BILU_ERR_TOO_LARGE;
    }
}

// ---- scratch of the allocation-free solve (bilu_solve_ex_noalloc) ----
// A fixed device array that the library never allocates or frees: the
// factors or ISAI values that pass between the factor and solve launches,
// for one chunk of systems.  Each GPU the library runs on holds its own copy
// until the process ends.  The array's name is a host-side stand-in: host
// code uses only the per-device address from hipGetSymbolAddress.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};

// the array's address on `dev` (current), looked up on the first call on that device
int noalloc_scratch(int dev, void **out)
{
    if (dev < 0 || dev >= 64) return BILU_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return BILU_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return BILU_OK;
}

}  // namespace

extern "C" {

void bilu_opts_default(bilu_opts *o)
{
    if (!o) return;
    o->precond = BILU_PRECOND_ILU0;
    o->parilu_sweeps = 5;
    o->tol = 1e-11;
    o->tol_mode = BILU_TOL_ABSOLUTE;
    o->max_iter = 500;
    o->merge = BILU_MERGE_AUTO;
    o->true_res_restart = 0;
    o->max_restarts = 3;
    o->kernel = BILU_KERNEL_AUTO;
    o->precision = BILU_PRECISION_FP64;
    o->inner_tol = 1e-5;
    o->inner_adapt = 0.5;
    o->max_refine = 10;
    o->parilu_stop_tol = 0.0;
    o->isai_power = 1;
    o->parilu_order = BILU_PARILU_JACOBI;
}

const char *bilu_version(void) { return "bilu 0.1 gfx950"; }

const char *bilu_status_string(int s)
{
    switch (s) {
    case BILU_OK: return "ok";
    case BILU_ERR_INVALID_ARG: return "invalid argument";
    case BILU_ERR_INVALID_PATTERN: return "invalid pattern";
    case BILU_ERR_MISSING_DIAGONAL: return "missing diagonal";
    case BILU_ERR_TOO_LARGE: return "problem too large for the kernels";
    case BILU_ERR_HIP: return "HIP runtime error";
    case BILU_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
    }
}

int bilu_pattern_create(int n, int nnz, const int *row_ptrs, const int *col_idx, bilu_pattern_t *out)
{
    if (!out) return BILU_ERR_INVALID_ARG;
    *out = nullptr;
    auto *h = new (std::nothrow) bilu_pattern_s();
    if (!h) return BILU_ERR_INVALID_ARG;
    std::string why;
    int st = bilu::analyse_pattern(n, nnz, row_ptrs, col_idx, h->pa, why);
    if (st != BILU_OK) {
        std::fprintf(stderr, "bilu_pattern_create: %s\n", why.c_str());
        delete h;
        return st;
    }
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) { delete h; return BILU_ERR_NO_DEVICE; }
    if (hipGetDevice(&h->device) != hipSuccess) { delete h; return BILU_ERR_HIP; }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device) == hipSuccess && lds > 0)
        h->max_lds = (size_t)lds;
    h->lds_per_cu = h->max_lds;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, h->device) == hipSuccess && lds > 0)
        h->lds_per_cu = (size_t)lds;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess && cus > 0)
        h->num_cus = cus;
    st = upload(*h);
    if (st != BILU_OK) { free_handle(h); return st; }
    *out = h;
    return BILU_OK;
}

int bilu_pattern_destroy(bilu_pattern_t p)
{
    if (!p) return BILU_ERR_INVALID_ARG;
    free_handle(p);
    return BILU_OK;
}

int bilu_pattern_get_info(bilu_pattern_t p, bilu_pattern_info *info)
{
    if (!p || !info) return BILU_ERR_INVALID_ARG;
    info->n = p->pa.n;
    info->nnz = p->pa.nnz;
    info->l_nnz = int(p->pa.l_ci.size());
    info->u_nnz = int(p->pa.u_ci.size());
    info->levels_lower = p->dev.nlev_l;
    info->levels_upper = p->dev.nlev_u;
    info->ilu_updates = int(p->pa.upd.size() / 2);
    info->max_row_nnz = p->pa.max_row;
    info->levels_entries = p->dev.nlev_e;
    info->ilu_eras = p->dev.n_eras;
    return BILU_OK;
}

int bilu_timing_begin(int kernel, int max_launches)
{
    if ((kernel != BILU_TIMING_FACTOR && kernel != BILU_TIMING_SOLVE) || max_launches < 1) return BILU_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(g_timing.mu);
    if (g_timing.kind) return BILU_ERR_INVALID_ARG;   // one window at a time
    for (auto &e : g_timing.ev) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    g_timing.ev.assign((size_t)max_launches, {nullptr, nullptr});
    for (auto &e : g_timing.ev) {
        HIP_TRY(hipEventCreate(&e.first));
        HIP_TRY(hipEventCreate(&e.second));
    }
    g_timing.used = 0;
    g_timing.kind = kernel;
    return BILU_OK;
}

int bilu_timing_end(float *avg_ms, int *launches)
{
    std::lock_guard<std::mutex> lock(g_timing.mu);
    if (!g_timing.kind) return BILU_ERR_INVALID_ARG;
    g_timing.kind = 0;
    const size_t n = std::min(g_timing.used, g_timing.ev.size());
    double sum = 0.0;
    for (size_t k = 0; k < n; k++) {
        HIP_TRY(hipEventSynchronize(g_timing.ev[k].second));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, g_timing.ev[k].first, g_timing.ev[k].second));
        sum += ms;
    }
    if (avg_ms) *avg_ms = n ? (float)(sum / (double)n) : 0.0f;
    if (launches) *launches = (int)n;
    return BILU_OK;
}

int bilu_pattern_reserve(bilu_pattern_t h, int batch)
{
    if (!h || batch < 0) return BILU_ERR_INVALID_ARG;
    HIP_TRY(hipSetDevice(h->device));
    void *ws = nullptr;
    return workspace(*h, workspace_need(*h, batch), &ws, nullptr);
}

int bilu_pattern_lu_split(bilu_pattern_t p, int *lrp, int *lci, int *urp, int *uci)
{
    if (!p) return BILU_ERR_INVALID_ARG;
    const auto &pa = p->pa;
    if (lrp) std::memcpy(lrp, pa.l_rp.data(), pa.l_rp.size() * sizeof(int));
    if (lci) std::memcpy(lci, pa.l_ci.data(), pa.l_ci.size() * sizeof(int));
    if (urp) std::memcpy(urp, pa.u_rp.data(), pa.u_rp.size() * sizeof(int));
    if (uci) std::memcpy(uci, pa.u_ci.data(), pa.u_ci.size() * sizeof(int));
    return BILU_OK;
}

// the workspace of one solve: the handle's (grown on demand) or, for the
// allocation-free entry point, a fixed buffer of `fixed_cap` bytes
struct SolveWs {
    void *fixed = nullptr;
    size_t fixed_cap = 0;
};

int solve_ws(bilu_pattern_s &h, size_t bytes, void **out, hipStream_t st, const SolveWs &w)
{
    if (!w.fixed) return workspace(h, bytes, out, st);
    if (bytes > w.fixed_cap) return BILU_ERR_TOO_LARGE;
    *out = w.fixed;
    return BILU_OK;
}

// general ISAI of a batch (values V on the pattern, output on the map's
// pattern: ostride values per system): the paired rows, then the
// four-per-wave rows, two launches on the stream
static int launch_isai_gen(const bilu_pattern_s &h, const DevIsai &M, int np, int ostride, int batch,
                           const double *V, void *out, bool f32, int *status, hipStream_t st)
{
    const int n = h.pa.n, nnz = h.pa.nnz;
    const size_t lds = (size_t)nnz * sizeof(double) + (size_t)ISAI_WAVES * WAVE * sizeof(double);  // + pivot-row cells
    if (lds > h.max_lds) return BILU_ERR_TOO_LARGE;
    const void *k32 = f32 ? (const void *)k_isai_gen<32, float> : (const void *)k_isai_gen<32, double>;
    const void *k16 = f32 ? (const void *)k_isai_gen<16, float> : (const void *)k_isai_gen<16, double>;
    if (lds > 65536) {
        HIP_TRY(hipFuncSetAttribute(k32, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        HIP_TRY(hipFuncSetAttribute(k16, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    float *o32 = static_cast<float *>(out);
    double *o64 = static_cast<double *>(out);
    if (np > 0) {
        if (f32) hipLaunchKernelGGL((k_isai_gen<32, float>), dim3(batch), dim3(WAVE * ISAI_WAVES), lds, st, M, nnz,
                                    batch, V, o32, ostride, status, 0, np, 0);
        else hipLaunchKernelGGL((k_isai_gen<32, double>), dim3(batch), dim3(WAVE * ISAI_WAVES), lds, st, M, nnz,
                                batch, V, o64, ostride, status, 0, np, 0);
    }
    if (n > np) {
        if (f32) hipLaunchKernelGGL((k_isai_gen<16, float>), dim3(batch), dim3(WAVE * ISAI_WAVES), lds, st, M, nnz,
                                    batch, V, o32, ostride, status, np, n, np > 0 ? 1 : 0);
        else hipLaunchKernelGGL((k_isai_gen<16, double>), dim3(batch), dim3(WAVE * ISAI_WAVES), lds, st, M, nnz,
                                batch, V, o64, ostride, status, np, n, np > 0 ? 1 : 0);
    }
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

// the general ISAI on the pattern itself
static int launch_isai_general(const bilu_pattern_s &h, int batch, const double *V, void *out, bool f32,
                               int *status, hipStream_t st)
{
    return launch_isai_gen(h, h.dev.isai[BILU_ISAI_GENERAL], h.pa.isai[BILU_ISAI_GENERAL].n_pair, h.pa.nnz, batch,
                           V, out, f32, status, st);
}

// FP64 ISAI solvers on the persistent kernel: a generation launch writes
// the ISAI values and their zero-pivot flags into the workspace -- ILU(0) or
// ParILU(0) factors then ISAI(L), ISAI(U) (k_factor_fast; FP64 throughout),
// or the general ISAI of A (k_isai_gen) -- then the persistent solve applies them
// (ilu_isai_solve.cu, gisai_solve.cu).  -1, with nothing launched, when the
// persistent solve cannot take the pattern.
static int solve_isai_fast(bilu_pattern_s &h, int batch, const double *A, const double *b, double *x, int pc,
                           int mode, int sweeps, double stop, int inplace, const double *F_init, SolveParams sp,
                           float *it, int *cv, double *ir, double *tr, int *fstatus, hipStream_t st,
                           const SolveWs &wsrc)
{
    const int fpc = pc == PC_ISAI_GEN ? FPC_GEN : FPC_ISAI;
    if (launch_fast(h, batch, A, b, x, nullptr, sp, it, cv, ir, tr, st, fpc, nullptr, true) != BILU_OK) return -1;
    const int nnz = h.pa.nnz;
    if (pc == PC_ISAI_GEN && (h.pa.isai[BILU_ISAI_GENERAL].max_m > ISAI_GROUP ||
                              (size_t)(nnz + ISAI_WAVES * WAVE) * sizeof(double) > h.max_lds))
        return -1;
    const size_t mbytes = ((size_t)batch * nnz * sizeof(double) + 255) / 256 * 256;
    void *ws = nullptr;
    int rc = solve_ws(h, mbytes + (size_t)batch * sizeof(int), &ws, st, wsrc);
    if (rc != BILU_OK) return rc;
    double *M = static_cast<double *>(ws);
    int *Z = reinterpret_cast<int *>(static_cast<char *>(ws) + mbytes);
    if (pc == PC_ISAI_GEN) {
        timing_mark(BILU_TIMING_FACTOR, st, true);
        rc = launch_isai_general(h, batch, A, M, false, Z, st);
        timing_mark(BILU_TIMING_FACTOR, st, false);
        if (rc != BILU_OK) return rc;
        if (fstatus) hipLaunchKernelGGL(k_fill_int, dim3((batch + 255) / 256), dim3(256), 0, st, fstatus, 0, batch);
    } else {
        rc = launch_factor_fast<double>(h, batch, A, F_init, M, mode, sweeps, stop, fstatus, st, 1, M, Z, nullptr,
                                        KF_MAX_WAVES, nullptr, 0, inplace);
        if (rc != BILU_OK) return rc;
    }
    sp.factor = FK_LOAD;
    return launch_fast(h, batch, A, b, x, M, sp, it, cv, ir, tr, st, fpc, Z);
}

static int solve_fp64(bilu_pattern_t h, int batch, const double *d_A, const double *d_b, double *d_x,
                      const bilu_opts &o, float *d_iter, int *d_conv, double *d_iter_res,
                      double *d_true_res, void *stream, const SolveWs &wsrc = SolveWs())
{
    const auto &pa = h->pa;
    const int n = pa.n, nnz = pa.nnz;
    if ((n + WAVE - 1) / WAVE > MAX_NR) return BILU_ERR_TOO_LARGE;
    int pc;
    SolveParams sp{o.tol, o.tol_mode, o.max_iter, o.true_res_restart, o.max_restarts, FK_NONE, o.parilu_sweeps};
    switch (o.precond) {
    case BILU_PRECOND_ILU0: pc = PC_LU; sp.factor = FK_ILU; break;
    case BILU_PRECOND_PARILU0: pc = PC_LU; sp.factor = FK_PARILU; break;
    case BILU_PRECOND_ILU0_ISAI: pc = PC_ISAI_LU; sp.factor = FK_ILU; break;
    case BILU_PRECOND_PARILU0_ISAI: pc = PC_ISAI_LU; sp.factor = FK_PARILU; break;
    case BILU_PRECOND_GENERAL_ISAI: pc = PC_ISAI_GEN; sp.factor = FK_NONE; break;
    default: pc = PC_NONE; sp.factor = FK_NONE; break;
    }
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(h->device));
    if (o.isai_power > 1 && (pc == PC_ISAI_LU || pc == PC_ISAI_GEN)) {
        // ISAI on |L|^k, |U|^k (|A|^k): the wave-per-system kernel with both maps
        const bilu_pattern_s::PowIsai *q1 = nullptr, *q2 = nullptr;
        const bool gen = pc == PC_ISAI_GEN;
        int rc = get_pow_isai(*h, gen ? BILU_ISAI_GENERAL : BILU_ISAI_LOWER, o.isai_power, q1);
        if (rc == BILU_OK && !gen) rc = get_pow_isai(*h, BILU_ISAI_UPPER, o.isai_power, q2);
        if (rc != BILU_OK) return rc;
        if (q1->max_m > (gen ? ISAI_GROUP : ISAI_TRI_MAX) || (q2 && q2->max_m > ISAI_TRI_MAX)) return BILU_ERR_TOO_LARGE;
        const size_t lds = sizeof(double) * ((size_t)nnz * (2 + (sp.factor == FK_PARILU ? 1 : 0)) +
                                             (size_t)q1->dev.nnz + (q2 ? (size_t)q2->dev.nnz : 0) + (size_t)n * 3);
        if (lds > h->max_lds) return BILU_ERR_TOO_LARGE;
        // the factors (one wave per system) and the ISAI values (k_isai_pow, four
        // waves per system) by launches of their own, through the workspace, when
        // their LDS fits; the solve kernel then loads them
        const size_t glds = sizeof(double) * ((size_t)nnz + std::max<size_t>(q1->dev.nnz, q2 ? q2->dev.nnz : 0));
        const size_t flds = sizeof(double) * (size_t)nnz * (sp.factor == FK_PARILU ? 3 : 2);
        if (glds <= h->max_lds && flds <= h->max_lds) {
            auto al = [](size_t b) { return (b + 255) / 256 * 256; };
            const size_t bF = gen ? 0 : al((size_t)batch * nnz * sizeof(double));
            const size_t b1 = al((size_t)batch * q1->dev.nnz * sizeof(double));
            const size_t b2 = gen ? 0 : al((size_t)batch * q2->dev.nnz * sizeof(double));
            void *ws = nullptr;
            rc = solve_ws(*h, bF + b1 + b2 + (size_t)batch * sizeof(int), &ws, st, wsrc);
            if (rc != BILU_OK) return rc;
            char *wb = static_cast<char *>(ws);
            double *Fw = reinterpret_cast<double *>(wb);
            double *M1 = reinterpret_cast<double *>(wb + bF);
            double *M2 = reinterpret_cast<double *>(wb + bF + b1);
            int *Z = reinterpret_cast<int *>(wb + bF + b1 + b2);
            if (glds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)k_isai_pow, hipFuncAttributeMaxDynamicSharedMemorySize, (int)glds));
            timing_mark(BILU_TIMING_FACTOR, st, true);
            if (gen) {
                rc = launch_isai_gen(*h, q1->dev.map, q1->dev.map.n_pair, q1->dev.nnz, batch, d_A, M1, false, Z, st);
                if (rc != BILU_OK) return rc;
            } else {
                if (flds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)k_factor, hipFuncAttributeMaxDynamicSharedMemorySize, (int)flds));
                hipLaunchKernelGGL(k_factor, dim3(batch), dim3(WAVE), flds, st, h->dev, batch, d_A,
                                   (const double *)nullptr, Fw, sp.factor, sp.sweeps, (int *)nullptr);
                // ISAI(L) has a unit diagonal: a zero pivot shows on U's diagonal
                hipLaunchKernelGGL(k_isai_pow, dim3(batch), dim3(WAVE * ISAI_WAVES), glds, st, h->dev, q1->dev, batch,
                                   (const double *)Fw, (int)BILU_ISAI_LOWER, M1, (int *)nullptr);
                hipLaunchKernelGGL(k_isai_pow, dim3(batch), dim3(WAVE * ISAI_WAVES), glds, st, h->dev, q2->dev, batch,
                                   (const double *)Fw, (int)BILU_ISAI_UPPER, M2, Z);
            }
            timing_mark(BILU_TIMING_FACTOR, st, false);
            HIP_TRY(hipGetLastError());
            sp.factor = FK_ISAI_LOADED;
            sp.zstatus = Z;
            sp.m2 = gen ? nullptr : M2;
            if (gen) return launch_solve<PC_ISAI_GEN_POW>(*h, batch, d_A, d_b, d_x, M1, sp, d_iter, d_conv,
                                                          d_iter_res, d_true_res, lds, st, q1->dev, DevIsaiPow{});
            return launch_solve<PC_ISAI_POW>(*h, batch, d_A, d_b, d_x, M1, sp, d_iter, d_conv, d_iter_res,
                                             d_true_res, lds, st, q1->dev, q2->dev);
        }
        if (gen) return launch_solve<PC_ISAI_GEN_POW>(*h, batch, d_A, d_b, d_x, nullptr, sp, d_iter, d_conv,
                                                      d_iter_res, d_true_res, lds, st, q1->dev, DevIsaiPow{});
        return launch_solve<PC_ISAI_POW>(*h, batch, d_A, d_b, d_x, nullptr, sp, d_iter, d_conv, d_iter_res,
                                         d_true_res, lds, st, q1->dev, q2->dev);
    }
    if (pc == PC_ISAI_LU && (pa.isai[BILU_ISAI_LOWER].max_m > ISAI_GROUP || pa.isai[BILU_ISAI_UPPER].max_m > ISAI_GROUP))
        return BILU_ERR_TOO_LARGE;
    if (pc == PC_ISAI_GEN && pa.isai[BILU_ISAI_GENERAL].max_m > ISAI_GROUP) return BILU_ERR_TOO_LARGE;
    const double *F = nullptr;
    // BILU_MERGE_AUTO splits ParILU(0): the factor kernel's lane-synchronous
    // sweeps beat the merged kernel's register-lean ones (n = 144, batch
    // 65536, 5 sweeps: 13.6 vs 25.6 ms), with the same results; and exact
    // ILU(0) when the solve runs on the one-wave kernel (the persistent factor
    // kernel's lane streams beat its row levels: n = 400, batch 8192: 51.2 vs
    // 55.3 ms)
    const bool fast_solve = pc == PC_LU && o.kernel != BILU_KERNEL_WAVE &&
                            launch_fast(*h, batch, d_A, d_b, d_x, nullptr, sp, d_iter, d_conv, d_iter_res,
                                        d_true_res, st, FPC_LU, nullptr, true) == BILU_OK;
    const bool seg = o.merge == BILU_SEGREGATED ||
                     (o.merge == BILU_MERGE_AUTO && pa.fast.ok && (sp.factor == FK_PARILU || !fast_solve));
    if (seg && pc == PC_LU) {
        // factor launch, factors through HBM (the handle's workspace), then the
        // solve launch (segregated.cu:429-490)
        const size_t flds = (size_t)(sp.factor == FK_PARILU ? 3 : 2) * nnz * sizeof(double);
        if (!pa.fast.ok && flds > h->max_lds) return BILU_ERR_TOO_LARGE;
        void *ws = nullptr;
        const int wrc = solve_ws(*h, (size_t)batch * nnz * sizeof(double), &ws, st, wsrc);
        if (wrc != BILU_OK) return wrc;
        double *Fbuf = static_cast<double *>(ws);
        if (pa.fast.ok) {
            // the persistent factorization (exact ILU(0) lane streams or Jacobi sweeps)
            const int frc = launch_factor_fast<double>(*h, batch, d_A, nullptr, Fbuf,
                                                       sp.factor == FK_PARILU ? FS_PARILU : FS_ILU, sp.sweeps, 0.0,
                                                       nullptr, st);
            if (frc != BILU_OK) return frc;
        } else {
            if (flds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)k_factor, hipFuncAttributeMaxDynamicSharedMemorySize, (int)flds));
            hipLaunchKernelGGL(k_factor, dim3(batch), dim3(WAVE), flds, st, h->dev, batch, d_A,
                               (const double *)nullptr, Fbuf, sp.factor, sp.sweeps, (int *)nullptr);
            HIP_TRY(hipGetLastError());
        }
        F = Fbuf;
        sp.factor = FK_LOAD;
    }
    int rc = -1;
    if (pc == PC_LU && o.kernel != BILU_KERNEL_WAVE) {
        rc = launch_fast(*h, batch, d_A, d_b, d_x, F, sp, d_iter, d_conv, d_iter_res, d_true_res, st);
        if (rc == -1 && o.kernel == BILU_KERNEL_PERSISTENT) rc = BILU_ERR_TOO_LARGE;
    } else if ((pc == PC_ISAI_LU || pc == PC_ISAI_GEN) && o.kernel != BILU_KERNEL_WAVE) {
        const int mode = sp.factor == FK_PARILU ? FS_PARILU : FS_ILU;
        rc = solve_isai_fast(*h, batch, d_A, d_b, d_x, pc, mode, sp.sweeps, 0.0, 0, nullptr, sp, d_iter, d_conv,
                             d_iter_res, d_true_res, nullptr, st, wsrc);
        if (rc == -1 && o.kernel == BILU_KERNEL_PERSISTENT) rc = BILU_ERR_TOO_LARGE;
    } else if (o.kernel == BILU_KERNEL_PERSISTENT) {
        rc = BILU_ERR_INVALID_ARG;   // no persistent kernel without a preconditioner
    }
    if (rc != -1) return rc;
    const bool need_w = (sp.factor == FK_PARILU) || pc == PC_ISAI_LU || pc == PC_ISAI_GEN;
    const size_t lds = sizeof(double) * ((size_t)nnz * (2 + (need_w ? 1 : 0)) + (size_t)n * (3 + (pc == PC_ISAI_LU ? 1 : 0)));
    if (lds > h->max_lds) rc = BILU_ERR_TOO_LARGE;
    else if (pc == PC_LU) rc = launch_solve<PC_LU>(*h, batch, d_A, d_b, d_x, F, sp, d_iter, d_conv, d_iter_res, d_true_res, lds, st);
    else if (pc == PC_ISAI_LU) {
        // the factors and ISAI(L), ISAI(U) by the persistent factor kernel (FP64),
        // through the workspace, when its plan exists; else inside the solve kernel
        if (pa.fast.ok) {
            const size_t mb = ((size_t)batch * nnz * sizeof(double) + 255) / 256 * 256;
            void *ws = nullptr;
            rc = solve_ws(*h, mb + (size_t)batch * sizeof(int), &ws, st, wsrc);
            if (rc != BILU_OK) return rc;
            double *M = static_cast<double *>(ws);
            int *Z = reinterpret_cast<int *>(static_cast<char *>(ws) + mb);
            rc = launch_factor_fast<double>(*h, batch, d_A, nullptr, M, sp.factor == FK_PARILU ? FS_PARILU : FS_ILU,
                                            sp.sweeps, 0.0, nullptr, st, 1, M, Z);
            if (rc != BILU_OK) return rc;
            sp.factor = FK_ISAI_LOADED;
            sp.zstatus = Z;
            F = M;
        }
        rc = launch_solve<PC_ISAI_LU>(*h, batch, d_A, d_b, d_x, F, sp, d_iter, d_conv, d_iter_res, d_true_res, lds, st);
    }
    else if (pc == PC_ISAI_GEN) {
        // the general ISAI by its multi-wave launches (k_isai_gen), through the
        // workspace, when their LDS fits; else generated inside the solve kernel
        if ((size_t)(nnz + ISAI_WAVES * WAVE) * sizeof(double) <= h->max_lds) {
            const size_t mb = ((size_t)batch * nnz * sizeof(double) + 255) / 256 * 256;
            void *ws = nullptr;
            rc = solve_ws(*h, mb + (size_t)batch * sizeof(int), &ws, st, wsrc);
            if (rc != BILU_OK) return rc;
            double *M = static_cast<double *>(ws);
            int *Z = reinterpret_cast<int *>(static_cast<char *>(ws) + mb);
            timing_mark(BILU_TIMING_FACTOR, st, true);
            rc = launch_isai_general(*h, batch, d_A, M, false, Z, st);
            timing_mark(BILU_TIMING_FACTOR, st, false);
            if (rc != BILU_OK) return rc;
            sp.factor = FK_ISAI_LOADED;
            sp.zstatus = Z;
            F = M;
        }
        rc = launch_solve<PC_ISAI_GEN>(*h, batch, d_A, d_b, d_x, F, sp, d_iter, d_conv, d_iter_res, d_true_res, lds, st);
    }
    else rc = launch_solve<PC_NONE>(*h, batch, d_A, d_b, d_x, F, sp, d_iter, d_conv, d_iter_res, d_true_res, lds, st);
    return rc;
}

static int solve_ex_impl(bilu_pattern_t h, int batch, const double *d_A, const double *d_b, double *d_x,
                         const bilu_opts *opts, const double *d_F_init, float *d_iter, int *d_conv,
                         double *d_iter_res, double *d_true_res, int *d_refine, int *d_factor_status,
                         void *stream, const SolveWs &wsrc)
{
    if (!h || batch < 0 || (batch > 0 && (!d_A || !d_b || !d_x || !d_iter || !d_conv)))
        return BILU_ERR_INVALID_ARG;
    if (batch == 0) return BILU_OK;
    bilu_opts o;
    if (opts) o = *opts; else bilu_opts_default(&o);
    if (o.precond < BILU_PRECOND_ILU0 || o.precond > BILU_PRECOND_FACTORS || o.max_iter < 0 ||
        o.parilu_sweeps < 0 || !(o.tol >= 0.0) ||
        (o.precond == BILU_PRECOND_FACTORS && (!d_F_init || o.precision != BILU_PRECISION_FP64)) ||
        (o.precision != BILU_PRECISION_FP64 && o.precision != BILU_PRECISION_FP32_IR) ||
        !(o.parilu_stop_tol >= 0.0) || o.isai_power < 1 || o.isai_power > 16 ||
        (o.parilu_order != BILU_PARILU_JACOBI && o.parilu_order != BILU_PARILU_INPLACE))
        return BILU_ERR_INVALID_ARG;
    if (o.precision == BILU_PRECISION_FP32_IR &&
        (!(o.inner_tol > 0.0 && o.inner_tol < 1.0) || o.max_refine < 1 || !(o.inner_adapt >= 0.0 && o.inner_adapt < 1.0)))
        return BILU_ERR_INVALID_ARG;
    const bool isai_any = o.precond == BILU_PRECOND_ILU0_ISAI || o.precond == BILU_PRECOND_PARILU0_ISAI ||
                          o.precond == BILU_PRECOND_GENERAL_ISAI;
    if (o.isai_power > 1 && isai_any) {
        // chosen-pattern ISAI: FP64, fixed sweeps, one wave per system
        if (o.precision != BILU_PRECISION_FP64 || d_F_init || o.parilu_stop_tol > 0.0 ||
            o.kernel == BILU_KERNEL_PERSISTENT)
            return BILU_ERR_INVALID_ARG;
        const int rc = solve_fp64(h, batch, d_A, d_b, d_x, o, d_iter, d_conv, d_iter_res, d_true_res, stream, wsrc);
        if (rc == BILU_OK && (d_refine || d_factor_status)) {
            hipStream_t st0 = (hipStream_t)stream;
            const int fs = o.precond == BILU_PRECOND_PARILU0_ISAI ? o.parilu_sweeps : 0;
            if (d_refine) hipLaunchKernelGGL(k_fill_int, dim3((batch + 255) / 256), dim3(256), 0, st0, d_refine, 0, batch);
            if (d_factor_status) hipLaunchKernelGGL(k_fill_int, dim3((batch + 255) / 256), dim3(256), 0, st0, d_factor_status, fs, batch);
            HIP_TRY(hipGetLastError());
        }
        return rc;
    }
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(h->device));
    const bool lu = o.precond == BILU_PRECOND_ILU0 || o.precond == BILU_PRECOND_PARILU0;
    const bool isai = o.precond == BILU_PRECOND_ILU0_ISAI || o.precond == BILU_PRECOND_PARILU0_ISAI;
    const bool par = o.precond == BILU_PRECOND_PARILU0 || o.precond == BILU_PRECOND_PARILU0_ISAI;
    const bool inplace = par && o.parilu_order == BILU_PARILU_INPLACE;
    const bool warm = par && (d_F_init || o.parilu_stop_tol > 0.0);
    const bool split = o.precision == BILU_PRECISION_FP32_IR || warm || d_factor_status || inplace;
    if (o.precond == BILU_PRECOND_FACTORS) {
        // caller's factors: the solve launch only (persistent kernel when the pattern fits)
        SolveParams sp{o.tol, o.tol_mode, o.max_iter, o.true_res_restart, o.max_restarts, FK_LOAD, 0};
        int rc = -1;
        if (o.kernel != BILU_KERNEL_WAVE)
            rc = launch_fast(*h, batch, d_A, d_b, d_x, d_F_init, sp, d_iter, d_conv, d_iter_res, d_true_res, st);
        if (rc == -1) {
            if (o.kernel == BILU_KERNEL_PERSISTENT) return BILU_ERR_TOO_LARGE;
            const size_t lds = sizeof(double) * ((size_t)h->pa.nnz * 2 + (size_t)h->pa.n * 3);
            if (lds > h->max_lds) return BILU_ERR_TOO_LARGE;
            rc = launch_solve<PC_LU>(*h, batch, d_A, d_b, d_x, d_F_init, sp, d_iter, d_conv, d_iter_res,
                                     d_true_res, lds, st);
        }
        if (rc == BILU_OK && (d_refine || d_factor_status)) {
            if (d_refine) hipLaunchKernelGGL(k_fill_int, dim3((batch + 255) / 256), dim3(256), 0, st, d_refine, 0, batch);
            if (d_factor_status) hipLaunchKernelGGL(k_fill_int, dim3((batch + 255) / 256), dim3(256), 0, st, d_factor_status, 0, batch);
            HIP_TRY(hipGetLastError());
        }
        return rc;
    }
    if (!split) {
        const int rc = solve_fp64(h, batch, d_A, d_b, d_x, o, d_iter, d_conv, d_iter_res, d_true_res, stream, wsrc);
        if (rc == BILU_OK && d_refine) {
            hipLaunchKernelGGL(k_fill_int, dim3((batch + 255) / 256), dim3(256), 0, st, d_refine, 0, batch);
            HIP_TRY(hipGetLastError());
        }
        return rc;
    }
    if (o.precond == BILU_PRECOND_GENERAL_ISAI && o.precision == BILU_PRECISION_FP32_IR) {
        // the general ISAI of A (FP64 Gauss-Jordan, stored FP32) through the
        // workspace, then the FP32 inner / FP64 refinement solve applying it
        // as one SpMV; always the two launches
        if (o.merge == BILU_MERGED) return BILU_ERR_INVALID_ARG;
        const auto &pa = h->pa;
        const int nr = (pa.n + WAVE - 1) / WAVE;
        if (!pa.fast.ok || slot_template_mixed(pa.fast.S) < 0 || nr > MAX_NR || pa.fast.nr != nr ||
            pa.isai[BILU_ISAI_GENERAL].max_m > ISAI_GROUP ||
            (size_t)(pa.nnz + ISAI_WAVES * WAVE) * sizeof(double) > h->max_lds)
            return BILU_ERR_TOO_LARGE;
        MixedParams mp{o.tol, o.tol_mode, (float)o.inner_tol, o.max_iter, o.max_refine, o.inner_tol, o.inner_adapt};
        const size_t fbytes = ((size_t)batch * pa.nnz * sizeof(float) + 255) / 256 * 256;
        void *gws = nullptr;
        int rc = solve_ws(*h, fbytes + (size_t)batch * sizeof(int), &gws, st, wsrc);
        if (rc != BILU_OK) return rc;
        float *M = static_cast<float *>(gws);
        int *Z = reinterpret_cast<int *>(static_cast<char *>(gws) + fbytes);
        timing_mark(BILU_TIMING_FACTOR, st, true);
        rc = launch_isai_general(*h, batch, d_A, M, true, Z, st);
        timing_mark(BILU_TIMING_FACTOR, st, false);
        if (rc != BILU_OK) return rc;
        if (d_factor_status) {
            hipLaunchKernelGGL(k_fill_int, dim3((batch + 255) / 256), dim3(256), 0, st, d_factor_status, 0, batch);
            HIP_TRY(hipGetLastError());
        }
        return launch_mixed(*h, batch, d_A, d_b, d_x, M, Z, PCK_GEN, mp, d_iter, d_conv, d_iter_res, d_true_res,
                            d_refine, st);
    }
    // factor launch (persistent), factors (or their ISAI(L, U)) through the
    // handle's workspace, then the solve; these paths serve ILU(0) / ParILU(0)
    // and their ISAI(L, U)
    if (!lu && !isai) return BILU_ERR_INVALID_ARG;
    if (!h->pa.fast.ok) return BILU_ERR_TOO_LARGE;
    const int nnz = h->pa.nnz;
    const int mode = par ? FS_PARILU : FS_ILU;
    const double stop = mode == FS_PARILU ? o.parilu_stop_tol : 0.0;
    void *ws = nullptr;
    int rc;
    if (o.precision == BILU_PRECISION_FP32_IR) {
        const int nr = (h->pa.n + WAVE - 1) / WAVE;
        if (slot_template_mixed(h->pa.fast.S) < 0 || nr > MAX_NR || h->pa.fast.nr != nr) return BILU_ERR_TOO_LARGE;
        MixedParams mp{o.tol, o.tol_mode, (float)o.inner_tol, o.max_iter, o.max_refine, o.inner_tol, o.inner_adapt};
        if (o.merge == BILU_MERGED) {
            // one launch: each wave factors its system in LDS, then solves it
            FactorArgs fa{mode, o.parilu_sweeps, inplace ? 1 : 0, stop, mode == FS_PARILU ? d_F_init : nullptr,
                          nullptr, 0, d_factor_status};
            return launch_mixed(*h, batch, d_A, d_b, d_x, nullptr, nullptr, isai ? PCK_ISAI : PCK_LU, mp, d_iter,
                                d_conv, d_iter_res, d_true_res, d_refine, st, Jobs(), 64, &fa);
        }
        const size_t fbytes = ((size_t)batch * nnz * sizeof(float) + 255) / 256 * 256;
        rc = solve_ws(*h, fbytes + (size_t)batch * sizeof(int), &ws, st, wsrc);
        if (rc != BILU_OK) return rc;
        float *F = static_cast<float *>(ws);
        int *Z = reinterpret_cast<int *>(static_cast<char *>(ws) + fbytes);
        rc = launch_factor_fast<float>(*h, batch, d_A, mode == FS_PARILU ? d_F_init : nullptr, F, mode,
                                       o.parilu_sweeps, stop, d_factor_status, st, isai ? 1 : 0, F, Z, nullptr,
                                       KF_MAX_WAVES, nullptr, 0, inplace ? 1 : 0);
        if (rc != BILU_OK) return rc;
        return launch_mixed(*h, batch, d_A, d_b, d_x, F, Z, isai ? PCK_ISAI : PCK_LU, mp, d_iter, d_conv,
                            d_iter_res, d_true_res, d_refine, st);
    }
    if (isai) {
        SolveParams sp{o.tol, o.tol_mode, o.max_iter, o.true_res_restart, o.max_restarts, FK_LOAD, 0};
        rc = solve_isai_fast(*h, batch, d_A, d_b, d_x, PC_ISAI_LU, mode, o.parilu_sweeps, stop, inplace ? 1 : 0,
                             mode == FS_PARILU ? d_F_init : nullptr, sp, d_iter, d_conv, d_iter_res, d_true_res,
                             d_factor_status, st, wsrc);
        if (rc == -1) return BILU_ERR_TOO_LARGE;
        if (rc == BILU_OK && d_refine) {
            hipLaunchKernelGGL(k_fill_int, dim3((batch + 255) / 256), dim3(256), 0, st, d_refine, 0, batch);
            HIP_TRY(hipGetLastError());
        }
        return rc;
    }
    rc = solve_ws(*h, (size_t)batch * nnz * sizeof(double), &ws, st, wsrc);
    if (rc != BILU_OK) return rc;
    double *F = static_cast<double *>(ws);
    rc = launch_factor_fast<double>(*h, batch, d_A, mode == FS_PARILU ? d_F_init : nullptr, F, mode,
                                    o.parilu_sweeps, stop, d_factor_status, st, 0, nullptr, nullptr, nullptr,
                                    KF_MAX_WAVES, nullptr, 0, inplace ? 1 : 0);
    if (rc != BILU_OK) return rc;
    SolveParams sp{o.tol, o.tol_mode, o.max_iter, o.true_res_restart, o.max_restarts, FK_LOAD, 0};
    rc = -1;
    if (o.kernel != BILU_KERNEL_WAVE) rc = launch_fast(*h, batch, d_A, d_b, d_x, F, sp, d_iter, d_conv, d_iter_res, d_true_res, st);
    if (rc == -1) {
        if (o.kernel == BILU_KERNEL_PERSISTENT) return BILU_ERR_TOO_LARGE;
        const size_t lds = sizeof(double) * ((size_t)nnz * 2 + (size_t)h->pa.n * 3);
        if (lds > h->max_lds) return BILU_ERR_TOO_LARGE;
        rc = launch_solve<PC_LU>(*h, batch, d_A, d_b, d_x, F, sp, d_iter, d_conv, d_iter_res, d_true_res, lds, st);
    }
    if (rc == BILU_OK && d_refine) {
        hipLaunchKernelGGL(k_fill_int, dim3((batch + 255) / 256), dim3(256), 0, st, d_refine, 0, batch);
        HIP_TRY(hipGetLastError());
    }
    return rc;
}

int bilu_solve_ex(bilu_pattern_t h, int batch, const double *d_A, const double *d_b, double *d_x,
                  const bilu_opts *opts, const double *d_F_init, float *d_iter, int *d_conv,
                  double *d_iter_res, double *d_true_res, int *d_refine, int *d_factor_status,
                  void *stream)
{
    return solve_ex_impl(h, batch, d_A, d_b, d_x, opts, d_F_init, d_iter, d_conv, d_iter_res, d_true_res, d_refine,
                         d_factor_status, stream, SolveWs());
}

int bilu_solve_ex_noalloc(bilu_pattern_t h, int batch, const double *d_A, const double *d_b, double *d_x,
                          const bilu_opts *opts, const double *d_F_init, float *d_iter, int *d_conv,
                          double *d_iter_res, double *d_true_res, int *d_refine, int *d_factor_status,
                          void *stream)
{
    if (!h || batch < 0 || (batch > 0 && (!d_A || !d_b || !d_x || !d_iter || !d_conv)))
        return BILU_ERR_INVALID_ARG;
    if (batch == 0) return BILU_OK;
    bilu_opts o;
    if (opts) o = *opts; else bilu_opts_default(&o);
    // power-k ISAI patterns are built and uploaded at their first use (an allocation)
    if (o.isai_power != 1) return BILU_ERR_INVALID_ARG;
    HIP_TRY(hipSetDevice(h->device));
    void *ws = nullptr;
    const int rc0 = noalloc_scratch(h->device, &ws);
    if (rc0 != BILU_OK) return rc0;
    // systems per chunk: every split path's layout fits (FP32 values + flags
    // with the flags 256-byte aligned, or FP64 factors)
    const size_t nnz = (size_t)h->pa.nnz, n = (size_t)h->pa.n;
    const size_t per = nnz * sizeof(double) + sizeof(int);
    if (NOALLOC_SCRATCH_BYTES < 256 + per) {
        std::fprintf(stderr, "bilu: bilu_solve_ex_noalloc: one system (nnz %zu) exceeds the %zu-byte scratch\n", nnz,
                     NOALLOC_SCRATCH_BYTES);
        return BILU_ERR_TOO_LARGE;
    }
    const int cb = (int)std::min<size_t>((size_t)batch, (NOALLOC_SCRATCH_BYTES - 256) / per);
    SolveWs w;
    w.fixed = ws;
    w.fixed_cap = NOALLOC_SCRATCH_BYTES;
    for (int s0 = 0; s0 < batch; s0 += cb) {
        const int c = std::min(cb, batch - s0);
        auto off = [&](auto *p, size_t stride) { return p ? p + (size_t)s0 * stride : p; };
        const int rc = solve_ex_impl(h, c, d_A + (size_t)s0 * nnz, d_b + (size_t)s0 * n, d_x + (size_t)s0 * n, &o,
                                     off(d_F_init, nnz), d_iter + s0, d_conv + s0, off(d_iter_res, 1),
                                     off(d_true_res, 1), off(d_refine, 1), off(d_factor_status, 1), stream, w);
        if (rc != BILU_OK) return rc;
    }
    return BILU_OK;
}

int bilu_mlp_solve(bilu_mlp_t m, bilu_pattern_t h, int batch, const double *d_A, const double *d_b, double *d_x,
                   const bilu_opts *opts, float *d_iter, int *d_conv, double *d_iter_res, double *d_true_res,
                   int *d_refine, int *d_factor_status, void *stream)
{
    if (!m || !h || batch < 0 || (batch > 0 && (!d_A || !d_b || !d_x || !d_iter || !d_conv)))
        return BILU_ERR_INVALID_ARG;
    if (batch == 0) return BILU_OK;
    bilu_opts o;
    if (opts) o = *opts; else bilu_opts_default(&o);
    const bool isai = o.precond == BILU_PRECOND_PARILU0_ISAI;
    if ((o.precond != BILU_PRECOND_PARILU0 && !isai) || o.precision != BILU_PRECISION_FP32_IR ||
        o.parilu_sweeps < 0 || !(o.tol >= 0.0) || o.max_iter < 0 || !(o.parilu_stop_tol >= 0.0) ||
        !(o.inner_tol > 0.0 && o.inner_tol < 1.0) || o.max_refine < 1 || o.isai_power != 1 ||
        !(o.inner_adapt >= 0.0 && o.inner_adapt < 1.0) ||
        (o.parilu_order != BILU_PARILU_JACOBI && o.parilu_order != BILU_PARILU_INPLACE))
        return BILU_ERR_INVALID_ARG;
    if (!h->pa.fast.ok) return BILU_ERR_TOO_LARGE;
    const int n = h->pa.n, nnz = h->pa.nnz, nr = (n + WAVE - 1) / WAVE;
    if (slot_template_mixed(h->pa.fast.S) < 0 || nr > MAX_NR || h->pa.fast.nr != nr) return BILU_ERR_TOO_LARGE;
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(h->device));
    const int chunk = std::min(batch, bilu_mlp_forward_rows());
    const bool merged = o.merge == BILU_MERGED;
    const size_t fbytes = ((size_t)chunk * nnz * sizeof(float) + 255) / 256 * 256;
    void *ws = nullptr;
    int rc = merged ? BILU_OK : workspace(*h, fbytes + (size_t)chunk * sizeof(int), &ws, st);
    if (rc != BILU_OK) return rc;
    float *F = static_cast<float *>(ws);
    int *Z = merged ? nullptr : reinterpret_cast<int *>(static_cast<char *>(ws) + fbytes);
    MixedParams mp{o.tol, o.tol_mode, (float)o.inner_tol, o.max_iter, o.max_refine, o.inner_tol, o.inner_adapt};
    for (int s0 = 0; s0 < batch; s0 += chunk) {
        const int cb = std::min(chunk, batch - s0);
        const double *A = d_A + (size_t)s0 * nnz;
        const _Float16 *Y = nullptr;
        int P3 = 0;
        rc = bilu_mlp_forward_raw(m, h, cb, A, st, &Y, &P3);
        if (rc != BILU_OK) return rc;
        if (merged) {
            // one launch: the warm start, sweeps and ISAI in each wave's LDS, then the solve
            FactorArgs fa{FS_PARILU, o.parilu_sweeps, o.parilu_order == BILU_PARILU_INPLACE ? 1 : 0,
                          o.parilu_stop_tol, nullptr, Y, P3, d_factor_status ? d_factor_status + s0 : nullptr};
            rc = launch_mixed(*h, cb, A, d_b + (size_t)s0 * n, d_x + (size_t)s0 * n, nullptr, nullptr,
                              isai ? PCK_ISAI : PCK_LU, mp, d_iter + s0, d_conv + s0,
                              d_iter_res ? d_iter_res + s0 : nullptr, d_true_res ? d_true_res + s0 : nullptr,
                              d_refine ? d_refine + s0 : nullptr, st, Jobs(), 64, &fa);
            if (rc != BILU_OK) return rc;
            continue;
        }
        rc = launch_factor_fast<float>(*h, cb, A, nullptr, F, FS_PARILU, o.parilu_sweeps, o.parilu_stop_tol,
                                       d_factor_status ? d_factor_status + s0 : nullptr, st, isai ? 1 : 0, F, Z,
                                       nullptr, KF_MAX_WAVES, Y, P3, o.parilu_order == BILU_PARILU_INPLACE ? 1 : 0);
        if (rc != BILU_OK) return rc;
        rc = launch_mixed(*h, cb, A, d_b + (size_t)s0 * n, d_x + (size_t)s0 * n, F, Z, isai ? PCK_ISAI : PCK_LU, mp,
                          d_iter + s0, d_conv + s0, d_iter_res ? d_iter_res + s0 : nullptr,
                          d_true_res ? d_true_res + s0 : nullptr, d_refine ? d_refine + s0 : nullptr, st);
        if (rc != BILU_OK) return rc;
    }
    return BILU_OK;
}

int bilu_solve(bilu_pattern_t h, int batch, const double *d_A, const double *d_b, double *d_x,
               const bilu_opts *opts, float *d_iter, int *d_conv, double *d_iter_res,
               double *d_true_res, void *stream)
{
    return bilu_solve_ex(h, batch, d_A, d_b, d_x, opts, nullptr, d_iter, d_conv, d_iter_res, d_true_res,
                         nullptr, nullptr, stream);
}

int bilu_ilu0_ex(bilu_pattern_t h, int batch, const double *d_A, double *d_F, int schedule, int *d_status,
                 void *stream)
{
    if (!h || batch < 0 || (batch > 0 && (!d_A || !d_F)) || schedule < BILU_ILU_SCHED_AUTO ||
        schedule > BILU_ILU_SCHED_STREAMS)
        return BILU_ERR_INVALID_ARG;
    if (schedule == BILU_ILU_SCHED_AUTO) schedule = h->ilu_sched ? h->ilu_sched : BILU_ILU_SCHED_ROWS;
    if (batch == 0) return BILU_OK;
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t st = (hipStream_t)stream;
    if (schedule == BILU_ILU_SCHED_STREAMS) {
        if (!h->pa.fast.ok) return BILU_ERR_TOO_LARGE;
        const int rc = launch_factor_fast<double>(*h, batch, d_A, nullptr, d_F, FS_ILU, 0, 0.0, nullptr, st);
        if (rc != BILU_OK) return rc;
        if (d_status) hipLaunchKernelGGL(k_diag_status, dim3(batch), dim3(WAVE), 0, st, h->dev, batch, d_F, d_status);
        HIP_TRY(hipGetLastError());
        return BILU_OK;
    }
    const size_t lds = 2 * (size_t)h->pa.nnz * sizeof(double);
    if (lds > h->max_lds) return BILU_ERR_TOO_LARGE;
    if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)k_factor, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int kind = schedule == BILU_ILU_SCHED_ENTRIES ? FK_ILU_ENTRIES
                     : schedule == BILU_ILU_SCHED_ERAS  ? FK_ILU_ERAS
                                                        : FK_ILU;
    hipLaunchKernelGGL(k_factor, dim3(batch), dim3(WAVE), lds, st, h->dev, batch, d_A,
                       (const double *)nullptr, d_F, kind, 0, d_status);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_ilu0(bilu_pattern_t h, int batch, const double *d_A, double *d_F, int *d_status, void *stream)
{
    return bilu_ilu0_ex(h, batch, d_A, d_F, BILU_ILU_SCHED_AUTO, d_status, stream);
}

int bilu_ilu0_tune(bilu_pattern_t h, int batch, const double *d_A, double *d_F, float *ms, int *best,
                   void *stream)
{
    if (!h || batch <= 0 || !d_A || !d_F) return BILU_ERR_INVALID_ARG;
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int choice = 0;
    float tbest = 0.0f;
    int rc = BILU_OK;
    for (int s = BILU_ILU_SCHED_ROWS; s <= BILU_ILU_SCHED_STREAMS && rc == BILU_OK; s++) {
        float t = -1.0f;
        int r = bilu_ilu0_ex(h, batch, d_A, d_F, s, nullptr, stream);   // warm-up; also the support check
        if (r == BILU_OK) {
            if (hipEventRecord(e0, st) != hipSuccess) { rc = BILU_ERR_HIP; break; }
            for (int k = 0; k < 2 && r == BILU_OK; k++) r = bilu_ilu0_ex(h, batch, d_A, d_F, s, nullptr, stream);
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess) { rc = BILU_ERR_HIP; break; }
            t *= 0.5f;
            if (r != BILU_OK) rc = r;
            else if (!choice || t < tbest) { choice = s; tbest = t; }
        } else if (r != BILU_ERR_TOO_LARGE) {
            rc = r;
        }
        if (ms) ms[s - 1] = r == BILU_OK ? t : -1.0f;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != BILU_OK) return rc;
    if (!choice) return BILU_ERR_TOO_LARGE;
    h->ilu_sched = choice;
    if (best) *best = choice;
    return BILU_OK;
}

int bilu_parilu0(bilu_pattern_t h, int batch, const double *d_A, const double *d_F_init, double *d_F,
                 int sweeps, void *stream)
{
    if (!h || batch < 0 || sweeps < 0 || (batch > 0 && (!d_A || !d_F))) return BILU_ERR_INVALID_ARG;
    if (batch == 0) return BILU_OK;
    const size_t lds = 3 * (size_t)h->pa.nnz * sizeof(double);
    if (lds > h->max_lds) return BILU_ERR_TOO_LARGE;
    HIP_TRY(hipSetDevice(h->device));
    if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)k_factor, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_factor, dim3(batch), dim3(WAVE), lds, (hipStream_t)stream, h->dev, batch, d_A,
                       d_F_init, d_F, (int)FK_PARILU, sweeps, (int *)nullptr);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_factor_residual(bilu_pattern_t h, int batch, const double *d_A, const double *d_F, double *d_res,
                         void *stream)
{
    if (!h || batch < 0 || (batch > 0 && (!d_A || !d_F || !d_res))) return BILU_ERR_INVALID_ARG;
    if (batch == 0) return BILU_OK;
    const size_t lds = (size_t)h->pa.nnz * sizeof(double);
    if (lds > h->max_lds) return BILU_ERR_TOO_LARGE;
    HIP_TRY(hipSetDevice(h->device));
    if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)k_factor_residual, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_factor_residual, dim3(batch), dim3(WAVE), lds, (hipStream_t)stream, h->dev, batch, d_A, d_F, d_res);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_split_lu(bilu_pattern_t h, int batch, const double *d_F, double *d_L, double *d_U, void *stream)
{
    if (!h || batch < 0 || (batch > 0 && !d_F)) return BILU_ERR_INVALID_ARG;
    if (batch == 0) return BILU_OK;
    HIP_TRY(hipSetDevice(h->device));
    const size_t tot = (size_t)batch * (size_t)std::max(h->dev.l_nnz, h->dev.u_nnz);
    const int grid = (int)std::min<size_t>((tot + 255) / 256, 65536);
    hipLaunchKernelGGL(k_split_lu, dim3(grid), dim3(256), 0, (hipStream_t)stream, h->dev, batch, d_F, d_L, d_U);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_expand_values(int batch, int nnz, int nnz_out, const int *d_src, const double *d_vals, double *d_out,
                       void *stream)
{
    if (batch < 0 || nnz < 0 || nnz_out < nnz || (batch > 0 && nnz_out > 0 && (!d_src || !d_out || (nnz > 0 && !d_vals))))
        return BILU_ERR_INVALID_ARG;
    const size_t tot = (size_t)batch * (size_t)nnz_out;
    if (tot == 0) return BILU_OK;
    const int grid = (int)std::min<size_t>((tot + 255) / 256, 8192);
    hipLaunchKernelGGL(k_expand_values, dim3(grid), dim3(256), 0, (hipStream_t)stream, batch, nnz, nnz_out, d_src,
                       d_vals, d_out);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_isai(bilu_pattern_t h, int batch, const double *d_vals, int type, double *d_out, int *d_status,
              void *stream)
{
    if (!h || batch < 0 || type < 0 || type > 2 || (batch > 0 && (!d_vals || !d_out))) return BILU_ERR_INVALID_ARG;
    if (batch == 0) return BILU_OK;
    if (h->pa.isai[type].max_m > ISAI_GROUP) return BILU_ERR_TOO_LARGE;
    const size_t lds = (type == BILU_ISAI_GENERAL ? 1 : 2) * (size_t)h->pa.nnz * sizeof(double);
    if (lds > h->max_lds) return BILU_ERR_TOO_LARGE;
    HIP_TRY(hipSetDevice(h->device));
    if (type == BILU_ISAI_GENERAL) return launch_isai_general(*h, batch, d_vals, d_out, false, d_status, (hipStream_t)stream);
    if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)k_isai, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_isai, dim3(batch), dim3(WAVE * ISAI_WAVES), lds, (hipStream_t)stream, h->dev, batch, d_vals, type, d_out, d_status);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_isai_pattern(bilu_pattern_t h, int type, int power, int *nnz, int *row_ptrs, int *col_idx)
{
    if (!h || type < 0 || type > 2 || power < 1 || power > 16 || !nnz || (row_ptrs && !col_idx))
        return BILU_ERR_INVALID_ARG;
    std::vector<int> rp, ci;
    bilu::isai_power_pattern(h->pa, type, power, rp, ci);
    *nnz = rp[h->pa.n];
    if (row_ptrs) {
        std::copy(rp.begin(), rp.end(), row_ptrs);
        std::copy(ci.begin(), ci.end(), col_idx);
    }
    return BILU_OK;
}

int bilu_isai_power(bilu_pattern_t h, int batch, const double *d_vals, int type, int power, double *d_out,
                    int *d_status, void *stream)
{
    if (!h || batch < 0 || type < 0 || type > 2 || power < 1 || power > 16 || (batch > 0 && (!d_vals || !d_out)))
        return BILU_ERR_INVALID_ARG;
    if (batch == 0) return BILU_OK;
    const bilu_pattern_s::PowIsai *q = nullptr;
    const int rc = get_pow_isai(*h, type, power, q);
    if (rc != BILU_OK) return rc;
    if (q->max_m > (type == BILU_ISAI_GENERAL ? ISAI_GROUP : ISAI_TRI_MAX)) return BILU_ERR_TOO_LARGE;
    const size_t lds = ((size_t)h->pa.nnz + (size_t)q->dev.nnz) * sizeof(double);
    if (lds > h->max_lds) return BILU_ERR_TOO_LARGE;
    HIP_TRY(hipSetDevice(h->device));
    if (type == BILU_ISAI_GENERAL)   // the general ISAI's own kernels, on the chosen pattern's map
        return launch_isai_gen(*h, q->dev.map, q->dev.map.n_pair, q->dev.nnz, batch, d_vals, d_out, false, d_status,
                               (hipStream_t)stream);
    if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)k_isai_pow, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_isai_pow, dim3(batch), dim3(WAVE * ISAI_WAVES), lds, (hipStream_t)stream, h->dev, q->dev, batch, d_vals,
                       type, d_out, d_status);
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_true_residual(bilu_pattern_t h, int batch, const double *d_A, const double *d_b, const double *d_x,
                       double *d_res, void *stream)
{
    if (!h || batch < 0 || (batch > 0 && (!d_A || !d_b || !d_x || !d_res))) return BILU_ERR_INVALID_ARG;
    if (batch == 0) return BILU_OK;
    const int nr = (h->pa.n + WAVE - 1) / WAVE;
    const size_t lds = ((size_t)h->pa.nnz + h->pa.n) * sizeof(double);
    if (lds > h->max_lds || nr > MAX_NR) return BILU_ERR_TOO_LARGE;
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t st = (hipStream_t)stream;
#define TR_CASE(R)                                                                                   \
    case R:                                                                                          \
        if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)k_true_residual<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
        hipLaunchKernelGGL(k_true_residual<R>, dim3(batch), dim3(WAVE), lds, st, h->dev, batch, d_A, d_b, d_x, d_res); \
        break;
    switch (nr) {
        TR_CASE(1) TR_CASE(2) TR_CASE(3) TR_CASE(4) TR_CASE(5) TR_CASE(6) TR_CASE(7) TR_CASE(8)
    default: return BILU_ERR_TOO_LARGE;
    }
#undef TR_CASE
    HIP_TRY(hipGetLastError());
    return BILU_OK;
}

}  // extern "C"

#include "uneven.hpp"
