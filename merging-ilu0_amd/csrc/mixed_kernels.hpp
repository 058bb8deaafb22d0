// Mixed-precision path and the stand-alone persistent factorization kernel.
// Included by solver.hip after fast_kernels.hpp (same anonymous namespace).
//
//   k_factor_fast<TO>  one wave per system at a time (persistent): exact
//                      ILU(0) (pull form by entry levels) or ParILU(0) Jacobi
//                      sweeps, optionally from a caller-supplied initial guess
//                      and stopped on the factor residual, with exact ILU(0)
//                      as the fallback; factors in FP64 in LDS, written out as
//                      TO (float for the mixed solve, double otherwise).
//   k_mixed<NR, SM>    FP64 iterative refinement around an FP32 BiCGSTAB:
//                      x += |r| d with A d = r / |r| solved in FP32 (FP32 A in
//                      VGPRs, FP32 factors in LDS, FP32 vectors and dots),
//                      then r = b - A x in FP64 (A re-read from global memory,
//                      FP64 dot), until |r| < tol.

constexpr int FS_ILU = 0, FS_PARILU = 1;

// ISAI of the L (or U) factor in F (the in-place factor layout, T = float:
// rounded to FP32, the records' 8-byte offsets halved here; T = double:
// FP64) into M (type T on the pattern: strictly lower positions = ISAI(L),
// diagonal and upper = ISAI(U)), by the lane streams of the pattern plan:
// each step is acc = fma(-m_c, f_q, acc) in FP64; a finishing step writes
// the target (divided by its diagonal for U) rounded to T and restarts acc.  A
// target's m_c were written by earlier steps of the same lane, and a wave's
// LDS accesses execute in program order, so no synchronisation is needed
// inside the stream.  Returns nonzero when a divisor (a U diagonal) is 0.
// DIV = false for the ISAI(L) streams, whose records never divide (unit
// diagonal): the division and the divisor's load are left out.
template <bool DIV, typename T = float>
__device__ __forceinline__ int isai_gen(const uint32_t *__restrict__ r0p, const uint32_t *__restrict__ r1p,
                                        int steps, const T *F, T *M, T *mscratch, int lane)
{
    constexpr int SH = sizeof(T) == 4 ? 1 : 0;
    constexpr int K = bilu::ISAI_BATCH;
    const char *Fb = reinterpret_cast<const char *>(F);
    char *Mb = reinterpret_cast<char *>(M);
    double acc = 0.0;
    int bad = 0;
    // records arrive in batches of K steps (streams are padded to whole
    // batches), the next batch in flight during the current one; operands
    // are read in the step (a step may read a value the lane wrote just before)
    uint32_t ra[K], rb[K];
#pragma unroll
    for (int j = 0; j < K; j++) { ra[j] = r0p[j * WAVE + lane]; rb[j] = r1p[j * WAVE + lane]; }
    for (int s0 = 0; s0 < steps; s0 += K) {
        const int nx = (s0 + K < steps) ? s0 + K : s0;
        uint32_t na[K], nb[K];
#pragma unroll
        for (int j = 0; j < K; j++) { na[j] = r0p[(nx + j) * WAVE + lane]; nb[j] = r1p[(nx + j) * WAVE + lane]; }
#pragma unroll
        for (int j = 0; j < K; j++) {
            const uint32_t a = ra[j], b = rb[j];
            const double mv = (double)*reinterpret_cast<const T *>(Mb + ((a >> 16) >> SH));
            const double fv = (double)*reinterpret_cast<const T *>(Fb + ((a & 0xFFFFu) >> SH));
            acc = fma(-mv, fv, acc);
            const uint32_t p = b & 0xFFFFu, d = b >> 16;
            const bool fin = p != 0xFFFFu;
            double val = acc;
            if constexpr (DIV) {
                const bool div = d != 0xFFFFu;
                const double dv = (double)*reinterpret_cast<const T *>(Fb + ((div ? d : 0u) >> SH));
                bad |= (fin && div && dv == 0.0);
                val = div ? acc / dv : acc;
            }
            T *dst = fin ? reinterpret_cast<T *>(Mb + (p >> SH)) : mscratch;
            *dst = (T)val;
            acc = fin ? 0.0 : acc;
        }
#pragma unroll
        for (int j = 0; j < K; j++) { ra[j] = na[j]; rb[j] = nb[j]; }
    }
    return bad;
}

// F[0 .. cnt) (FP64) rounded to FP32 in place (the first half of the
// region), 64 values at a time: a chunk's values are all read before its
// writes, which cover only the halves of chunks already read
__device__ __forceinline__ void round_in_place(double *F, int cnt, int lane)
{
    float *F32 = reinterpret_cast<float *>(F);
    for (int e0 = 0; e0 < cnt; e0 += WAVE) {
        const int e = e0 + lane;
        const double v = e < cnt ? F[e] : 0.0;
        asm volatile("" ::: "memory");
        if (e < cnt) F32[e] = (float)v;
        asm volatile("" ::: "memory");
    }
}

// One system's factorization in the wave's LDS region: A (FP64, Ap) into
// F[nnz + 4] (F[nnz] = 0, F[nnz + 1] = 1, F[nnz + 2] scratch), then exact
// ILU(0) (FS_ILU) or ParILU(0) sweeps (FS_PARILU: warm start from the
// network's raw outputs Yp or from the guess Ip, in place or Jacobi between F
// and W, residual stop with exact ILU(0) as the fallback).  isai: ISAI(L)
// and ISAI(U) of the factors into the ISAI region (TO = float: the factors
// rounded to FP32 in place, M = the second half of F's region; TO = double:
// FP64 factors, M = W's region).  Returns the ISAI generation's zero-divisor
// flag (wave-uniform); status = sweeps used, -1 on the fallback (0 for FS_ILU).
template <typename TO>
__device__ __forceinline__ int factor_system(const FastParams &fp, int n, int nnz, const double *__restrict__ Ap,
                                             const double *__restrict__ Ip, const _Float16 *__restrict__ Yp,
                                             int mode, int sweeps, double stop_tol, int isai, int inplace,
                                             const uint32_t *par_l, bool par_lds, double *F, double *W, float *Dg,
                                             int lane, int &status)
{
    float *F32 = reinterpret_cast<float *>(F);
    float *M = F32 + (nnz + 4);
    status = 0;
    wave_copy_wide<24>(F, Ap, nnz, lane);
    // the zero and one cells (the previous system's ISAI values may cover them)
    if (lane == 0) { F[nnz] = 0.0; F[nnz + 1] = 1.0; }
    if (mode == FS_PARILU) {
        wave_sync();
        double anorm2 = 0.0;
        if (stop_tol > 0.0) {
            for (int e = lane; e < nnz; e += WAVE) anorm2 = fma(F[e], F[e], anorm2);
            anorm2 = wave_sum(anorm2);
        }
        if (Yp) {
            // the warm start from the network's raw outputs y (FP16):
            // guess = SIGMA sinh(YS (base + y)) for the entries with work, the
            // same operations as bilu_mlp_infer (guess_from); the others keep
            // a_ij.  The diagonal of A goes to Dg first (FP32, as the base guess
            // rounds it), so the positions are visited in one coalesced pass
            // with WU outputs and position codes per lane in flight: a lower
            // entry (i, j) (code = j) reads its pivot a_jj from Dg, an upper
            // entry with work has code 0xFFFE, one without work 0xFFFF.
            constexpr int WU = 24;
            for (int j = lane; j < n; j += WAVE) Dg[j] = (float)F[fp.diag[j]];
            wave_sync();
            // the loads are unconditional (positions past nnz re-read the last
            // one), so all 2 WU of a lane are in flight together
            for (int e0 = lane; e0 < nnz; e0 += WU * WAVE) {
                _Float16 yh[WU];
                uint32_t cd[WU];
#pragma unroll
                for (int j = 0; j < WU; j++) {
                    const int e = min(e0 + j * WAVE, nnz - 1);
                    yh[j] = Yp[e];
                    cd[j] = fp.wcode[e];
                }
#pragma unroll
                for (int j = 0; j < WU; j++) {
                    const int e = e0 + j * WAVE;
                    if (e < nnz && cd[j] != 0xFFFFu) {
                        const bool lower = cd[j] != 0xFFFEu;
                        F[e] = bilu_mlp::guess_from((float)F[e], lower ? Dg[cd[j]] : 1.0f, lower, (float)yh[j]);
                    }
                }
            }
        } else if (Ip) {
            // entries with work start from the guess; the others are exact (= a)
            // unconditional loads (entries past ne re-read the last one)
            for (int i0 = lane; i0 < fp.ne; i0 += 8 * WAVE) {
                int pos[8];
                double g[8];
#pragma unroll
                for (int j = 0; j < 8; j++) pos[j] = fp.etab[min(i0 + j * WAVE, fp.ne - 1)].x;
#pragma unroll
                for (int j = 0; j < 8; j++) g[j] = Ip[pos[j]];
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (i0 + j * WAVE < fp.ne) F[pos[j]] = g[j];
            }
        }
        wave_sync();
        const double lim2 = stop_tol * stop_tol * anorm2;
        int used = 0;
        bool met = stop_tol <= 0.0;
        if (inplace) {
            // in-place sweeps over F; the stop test uses the residuals the
            // rounds saw, and the iterate after the sweep is kept
            while (used < sweeps) {
                const double rr = wave_sum(fp.par_rounds > 0
                    ? (par_lds ? parilu_inplace<true>(par_l, par_l + fp.par_psteps * WAVE, fp.par_rounds,
                                                      fp.par_psteps, Ap, F, lane)
                               : parilu_inplace<true>(fp.parp, fp.parf, fp.par_rounds, fp.par_psteps, Ap, F,
                                                      lane))
                    : 0.0);
                wave_sync();
                used++;
                if (stop_tol > 0.0 && rr <= lim2) { met = true; break; }
            }
        } else {
            // Jacobi sweeps between F and W (entries without work and the zero cell in both)
            wave_copy(W, F, nnz, lane);
            if (lane == 0) W[nnz] = 0.0;
            wave_sync();
            double *src = F, *dst = W;
            while (used < sweeps) {
                const double rr = wave_sum(fp.par_rounds > 0
                    ? (par_lds ? parilu_rounds<true>(par_l, par_l + fp.par_psteps * WAVE, fp.par_rounds,
                                                     fp.par_psteps, Ap, src, dst, lane)
                               : parilu_rounds<true>(fp.parp, fp.parf, fp.par_rounds, fp.par_psteps, Ap, src,
                                                     dst, lane))
                    : 0.0);
                wave_sync();
                used++;
                if (stop_tol > 0.0 && rr <= lim2) {
                    // the iterate this sweep started from (src) meets the tolerance: keep it
                    used--;
                    met = true;
                    break;
                }
                double *tmp = src; src = dst; dst = tmp;
            }
            if (src != F) {
                wave_copy(F, src, nnz, lane);
                wave_sync();
            }
        }
        if (!met) {
            // fallback: exact ILU(0)
            wave_copy(F, Ap, nnz, lane);
            wave_sync();
            ilu_stream(fp.ilu0, fp.ilu1, fp.ilu_steps, F, F + nnz + 2, fp.diag, n, lane);
            status = -1;
        } else {
            status = used;
        }
    } else {
        wave_sync();
        ilu_stream(fp.ilu0, fp.ilu1, fp.ilu_steps, F, F + nnz + 2, fp.diag, n, lane);
    }
    int bad = 0;
    if (isai) {
        if constexpr (sizeof(TO) == 4) {
            wave_sync();
            round_in_place(F, nnz + 4, lane);
            if (lane == 0) { M[nnz] = 0.0f; M[nnz + 1] = 1.0f; M[nnz + 2] = -1.0f; }
            wave_sync();
            bad = isai_gen<false>(fp.isL0, fp.isL1, fp.isL_steps, F32, M, M + nnz + 3, lane);
            bad |= isai_gen<true>(fp.isU0, fp.isU1, fp.isU_steps, F32, M, M + nnz + 3, lane);
        } else {
            double *M64 = W;
            wave_sync();
            if (lane == 0) { M64[nnz] = 0.0; M64[nnz + 1] = 1.0; M64[nnz + 2] = -1.0; }
            wave_sync();
            bad = isai_gen<false, double>(fp.isL0, fp.isL1, fp.isL_steps, F, M64, M64 + nnz + 3, lane);
            bad |= isai_gen<true, double>(fp.isU0, fp.isU1, fp.isU_steps, F, M64, M64 + nnz + 3, lane);
        }
        bad = __any(bad);
    }
    return bad;
}

template <typename TO>
__global__ __launch_bounds__(KF_MAX_WAVES * WAVE) void k_factor_fast(
    FastParams fp, int n, int nnz, int batch, const double *__restrict__ gA,
    const double *__restrict__ gFinit, TO *__restrict__ gF, int mode, int sweeps, double stop_tol,
    int *__restrict__ g_status, int isai, void *__restrict__ gM, int *__restrict__ gZ,
    const int *__restrict__ sysidx, const _Float16 *__restrict__ gY, int P3, int inplace)
{
    // job j factors system sysidx[j] (identity when sysidx is null); outputs are per job
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    // per wave: F[nnz + 4] FP64 (F[nnz] = 0, F[nnz + 1] = 1, F[nnz + 2] scratch),
    // then the Jacobi ParILU buffer W[nnz + 2] FP64 (not for in-place sweeps).
    // ISAI (TO = float): the factors are rounded to FP32 in place when the
    // factorization is done (the first half of F), the ISAI values M[nnz + 4]
    // FP32 (M[nnz] = 0, M[nnz + 1] = 1, M[nnz + 2] = -1, M[nnz + 3] scratch)
    // take the second half.  ISAI (TO = double): the factors stay FP64 and
    // the ISAI values, FP64 with the same cells, take the second region (the
    // Jacobi buffer, free once the factorization is done).
    // ParILU with room for it: the sweep records (products, then finishes)
    // in a shared region ahead of the wave regions, read by LDS instructions
    // instead of global loads (fp.shared_bytes = 0 otherwise)
    uint32_t *par_l = reinterpret_cast<uint32_t *>(smem);
    const bool par_lds = fp.shared_bytes > 0;
    if (par_lds) {
        const int np = fp.par_psteps * WAVE, nf = fp.par_rounds * WAVE;
        for (int i = threadIdx.x; i < np; i += blockDim.x) par_l[i] = fp.parp[i];
        for (int i = threadIdx.x; i < nf; i += blockDim.x) par_l[np + i] = fp.parf[i];
        __syncthreads();
    }
    double *F = reinterpret_cast<double *>(smem + fp.shared_bytes + (size_t)wave * fp.wave_bytes);
    double *W = F + (nnz + 4);
    float *M = reinterpret_cast<float *>(F) + (nnz + 4);
    // the fused warm start's diagonal of A (FP32), after F (and W)
    float *Dg = reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(F) + fp.wave_bytes) - fp.dg_floats;
    for (;;) {
        int sys = 0;
        if (lane == 0) sys = atomicAdd(fp.counter, 1);
        sys = __builtin_amdgcn_readfirstlane(sys);
        if (sys >= batch) break;
        const int si = sysidx ? sysidx[sys] : sys;
        int status;
        const int bad = factor_system<TO>(fp, n, nnz, gA + (size_t)si * nnz,
                                          gFinit ? gFinit + (size_t)si * nnz : nullptr,
                                          gY ? gY + (size_t)sys * P3 : nullptr, mode, sweeps, stop_tol, isai,
                                          inplace, par_l, par_lds, F, W, Dg, lane, status);
        if (isai) {
            wave_sync();
            if constexpr (sizeof(TO) == 4) wave_copy(static_cast<float *>(gM) + (size_t)sys * nnz, M, nnz, lane);
            else wave_copy(static_cast<double *>(gM) + (size_t)sys * nnz, W, nnz, lane);
            if (lane == 0) gZ[sys] = bad;
        } else {
            TO *Fo = gF + (size_t)sys * nnz;
            wave_copy(Fo, F, nnz, lane);
        }
        if (g_status && lane == 0) g_status[sys] = status;
        wave_sync();
    }
}

// the factorization of the merged mixed kernel (FM = 1): the arguments of
// factor_system, and where the ParILU status goes (by result index; null:
// nowhere)
struct FactorArgs {
    int mode, sweeps, inplace;
    double stop_tol;
    const double *Finit;    // ParILU guess [system][nnz] or null
    const _Float16 *Y;      // network outputs [job][P3] or null
    int P3;
    int *status;
};

struct MixedParams {
    double tol;          // outer tolerance (relative or absolute, tol_mode)
    int tol_mode;
    float inner_tol;     // inner relative tolerance (FP32 BiCGSTAB)
    int max_iter;        // inner iterations per refinement step
    int max_refine;      // refinement steps
    double inner_tol_d;  // inner_tol as given (FP64)
    double inner_adapt;  // stop the inner solve at max(inner_tol, inner_adapt * tol / |r|) (0: inner_tol)
};

// w = A x in FP64 by CSR rows: lane l computes the rows l + 64 k, each an
// fma chain left to right over the row's entries (the same chain as the
// slot SpMVs, whose lanes hold whole rows in CSR order), with A and the
// column indices read from global memory in batches of 8 entries (the A
// loads do not wait on each other or on the index loads) and x gathered from
// LDS (FP64 cells).
template <int NR>
__device__ __forceinline__ void residual_rows(const int *__restrict__ rp, const int *__restrict__ ci,
                                              const double *__restrict__ Ap, const double *__restrict__ x,
                                              Vec<NR, double> &w, int n, int lane)
{
    constexpr int K = 8;
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int i = lane + WAVE * k;
        const int pb = (i < n) ? rp[i] : 0, pe = (i < n) ? rp[i + 1] : 0;
        int len = pe - pb;
        // wave-uniform trip count: the longest row of this row slot
        int lmax = len;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) lmax = max(lmax, __shfl_xor(lmax, o));
        double acc = 0.0;
        for (int q0 = 0; q0 < lmax; q0 += K) {
            int c[K];
            double av[K];
#pragma unroll
            for (int j = 0; j < K; j++) {
                const bool in = q0 + j < len;
                c[j] = in ? ci[pb + q0 + j] : n;
                av[j] = in ? Ap[pb + q0 + j] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < K; j++) {
                const double xv = x[c[j]];
                if (q0 + j < len) acc = fma(av[j], xv, acc);
            }
        }
        w.e[k] = acc;
    }
}

// y (LDS, by row) = C x over A's slot layout with the coefficients read
// from the wave's LDS, slot-major (cl[s * 64 + lane]) and one packed record
// per slot: bits 0-15 the byte offset of x (FP32 cells); bits 16-31 the byte
// offset of the output cell (the row's cell when the slot ends a row, else
// the output vector's scratch cell n + 1, the largest offset in use).  A
// record below thr = (4 (n + 1)) << 16 ends a row: the running sum restarts.
// One LDS word per slot instead of a separate x and output record (the
// kernel is bound by instruction issue), each half taken by one add.  The
// same CSR-order fma chain as spmv_slots.
template <int SM>
__device__ __forceinline__ void spmv32(const float *__restrict__ cl, const uint32_t *__restrict__ meta,
                                       const float *__restrict__ x, float *__restrict__ y, uint32_t thr, int lane)
{
    const char *xb = reinterpret_cast<const char *>(x);
    char *yb = reinterpret_cast<char *>(y);
    float acc = 0.0f;
    constexpr int G = SPMV_G;
    uint32_t rw[SM];
    float xv[SM], cw[SM];
    auto fetch = [&](int s) {
        rw[s] = meta[s * WAVE + lane];
        cw[s] = cl[s * WAVE + lane];
        xv[s] = *reinterpret_cast<const float *>(xb + (rw[s] & 0xFFFFu));
    };
#pragma unroll
    for (int s = 0; s < G && s < SM; s++) fetch(s);
#pragma unroll
    for (int s = 0; s < SM; s++) {
        if (s % G == 0) {
#pragma unroll
            for (int t = s + G; t < s + 2 * G && t < SM; t++) fetch(t);
        }
        const uint32_t r = rw[s];
        acc = __builtin_fmaf(cw[s], xv[s], acc);
        *reinterpret_cast<float *>(yb + (r >> 16)) = acc;
        acc = r < thr ? 0.0f : acc;
    }
}

// y (LDS, by row) = C x over one of the triangle slot layouts: the
// coefficient of slot s is c[s] (registers), the record one packed word
// (x offset | output offset << 16, as for spmv32); only the first `ns`
// slots are visited (a wave-uniform bound, checked per gather group).
// Each row sum is the CSR-order fma chain over the triangle's entries.
template <int SMX>
__device__ __forceinline__ void spmv_tri(const float (&c)[SMX], int ns, const uint32_t *__restrict__ rec,
                                         const float *__restrict__ x, float *__restrict__ y, uint32_t thr,
                                         int lane)
{
    const char *xb = reinterpret_cast<const char *>(x);
    char *yb = reinterpret_cast<char *>(y);
    float acc = 0.0f;
    constexpr int G = SPMV_G;
    uint32_t rw[SMX];
    float xv[SMX];
    auto fetch = [&](int s) {
        rw[s] = rec[s * WAVE + lane];
        xv[s] = *reinterpret_cast<const float *>(xb + (rw[s] & 0xFFFFu));
    };
#pragma unroll
    for (int s = 0; s < G && s < SMX; s++) fetch(s);
#pragma unroll
    for (int s = 0; s < SMX; s++) {
        if (s % G == 0) {
            if (s >= ns) break;
#pragma unroll
            for (int t = s + G; t < s + 2 * G && t < SMX; t++) fetch(t);
        }
        const uint32_t r = rw[s];
        acc = __builtin_fmaf(c[s], xv[s], acc);
        *reinterpret_cast<float *>(yb + (r >> 16)) = acc;
        acc = r < thr ? 0.0f : acc;
    }
}

#ifndef KM_MAX_WAVES
#define KM_MAX_WAVES 12      // LU variant
#endif
#ifndef KMI_MAX_WAVES
#define KMI_MAX_WAVES 12     // ISAI variant (two slot coefficient sets in VGPRs; 12 waves fill the LDS)
#endif

constexpr int PCK_LU = 0, PCK_ISAI = 1, PCK_GEN = 2;

// PCK_LU: gF = FP32 in-place factors (triangular solves); PCK_ISAI: gF =
// FP32 ISAI(L) / ISAI(U) values on the pattern, gZ = zero-pivot flags of
// their generation (two SpMVs per apply); PCK_GEN: gF = FP32 general ISAI of
// A on the pattern, gZ likewise (one SpMV per apply, over A's slot records;
// segregated only)
// waves per workgroup (launch bound) of a k_mixed instantiation: 72 slots
// (patterns up to ~4600 nnz) need the register budget of 8 waves (LU: A in
// VGPRs; ISAI: 2 x 40 triangle coefficients)
constexpr int km_waves(int SM, int PCK, int FM = 0)
{
    // the merged kernel holds the factorization's and the solve's registers:
    // the budget of 8 waves beyond 24 slots (its LDS region seldom fits more;
    // small patterns keep 12, measured faster at n = 32)
    return (SM > 48 || (FM && SM > 24)) ? 8 : (PCK == PCK_LU ? KM_MAX_WAVES : KMI_MAX_WAVES);
}

// FM = 0 (segregated): the FP32 factors / ISAI values come from gF (a
// factor launch before).  FM = 1 (merged): each system is factored in the
// wave's own region first (factor_system, FP64 in LDS, then FP32), and the
// solve reads the factors / ISAI values there; the region is the larger of
// the two layouts, the solve's overlaying the factorization's once its
// values are in registers.  Same arithmetic, same results bit for bit.
// RG = 1: the slot records are read from global memory (fp.rec32, the same
// layout as the shared LDS copy) instead, freeing that LDS for more waves.
template <int NR, int SM, int PCK, int SLU = 1, int FM = 0, int RG = 0>
__global__ __launch_bounds__(km_waves(SM, PCK, FM) * WAVE) void k_mixed(
    DevPattern P, FastParams fp, int batch, const double *__restrict__ gA,
    const double *__restrict__ gb, double *__restrict__ gx, const float *__restrict__ gF,
    const int *__restrict__ gZ, MixedParams mp, float *__restrict__ g_iter, int *__restrict__ g_conv,
    double *__restrict__ g_ires, double *__restrict__ g_tres, int *__restrict__ g_refine,
    const int *__restrict__ sysidx, const int *__restrict__ outidx, FactorArgs fa)
{
    // job j solves system sysidx[j] (A, b, x) and writes its results at
    // outidx[j] (identity when null); factors / ISAI values are per job
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int n = P.n, nnz = P.nnz, S = fp.S;
    uint32_t *meta_l = reinterpret_cast<uint32_t *>(smem);
    const uint32_t *meta = RG ? fp.rec32 : meta_l;
    const uint32_t *trL = fp.trL;
    const uint32_t *trU = fp.trU;
    // LU: the records of spmv_slots.  ISAI: the packed records of spmv32
    // for the A matvec, from those (8 col | 8 row << 16, 0xFFFF: no row end);
    // padding slots read x[0] with coefficient 0.  ISAI: then the records of
    // the ISAI(L) and ISAI(U) slot layouts (their not-a-row-end bit 31
    // dropped: the output offset tells), padded to SLU slots each.
    const uint32_t *lrec = meta + SM * WAVE, *urec = lrec + SLU * WAVE;
    uint32_t *lrec_l = meta_l + SM * WAVE, *urec_l = lrec_l + SLU * WAVE;
    for (int i = threadIdx.x; i < (RG ? 0 : SM * WAVE); i += blockDim.x) {
        const uint32_t m = (i < S * WAVE) ? fp.slot_meta[i] : 0xFFFF0000u;
        if constexpr (PCK == PCK_LU) {
            meta_l[i] = m;
        } else {
            const bool end = (m >> 16) != 0xFFFFu;
            const uint32_t yo = end ? (m >> 16) >> 1 : (uint32_t)(n + 1) * 4u;
            meta_l[i] = ((m & 0xFFFFu) >> 1) | (yo << 16);
        }
    }
    // FP32 records: an output offset below that of the scratch cell n + 1 ends a row
    const uint32_t thr = (uint32_t)(n + 1) * 4u << 16;
    if constexpr (PCK == PCK_ISAI && !RG) {
        const uint32_t pad = (uint32_t)n * 4u | thr;
        for (int i = threadIdx.x; i < SLU * WAVE; i += blockDim.x) {
            lrec_l[i] = (i < fp.SL * WAVE) ? fp.isl_meta[i] & 0x7FFFFFFFu : pad;
            urec_l[i] = (i < fp.SU * WAVE) ? fp.isu_meta[i] & 0x7FFFFFFFu : pad;
        }
    }
    // per wave: F[nnz + 2] FP32 (F[nnz] = 0), then two vector regions of
    // (n + 2) doubles each, used as FP32 vectors by the inner solve (zero
    // cell at [n], scratch at [n + 1]) and as FP64 vectors by the outer
    // residual, then the FP64 row accumulators of the solves acc[64 NR]
    // (PCK_ISAI: instead of the factors, A (FP32) slot-major [SM][64]; no accumulators)
    unsigned char *wb = smem + fp.shared_bytes + (size_t)wave * fp.wave_bytes;
    float *F = reinterpret_cast<float *>(wb);
    float *Asl = F;
    double *d0 = reinterpret_cast<double *>(wb + (PCK == PCK_LU ? (((size_t)(nnz + 2) * 4 + 15) / 16) * 16
                                                                : (size_t)SM * WAVE * 4));
    double *d1 = d0 + (n + 2);
    float *e0 = reinterpret_cast<float *>(d0), *e1 = reinterpret_cast<float *>(d1);
    float *scratch = e1 + (n + 1);
    double *accb = d1 + (n + 2);
    if (PCK == PCK_LU && lane == 0) F[nnz] = 0.0f;
    __syncthreads();

    int rlev[NR];
#pragma unroll
    for (int k = 0; k < NR; k++)
        rlev[k] = (int)fp.rlevL[k * WAVE + lane] | ((int)fp.rlevU[k * WAVE + lane] << 16);
    const int vcntL = fp.cntL[lane], vcntU = fp.cntU[lane];

    // LU: A (FP32) in VGPRs.  ISAI: the ISAI(L) and ISAI(U) coefficients of
    // the triangle slot layouts in VGPRs (1 on ISAI(L)'s unit diagonal), A in LDS.
    // GEN: the general ISAI's coefficients in A's slot layout in VGPRs, A in LDS.
    static_assert(PCK != PCK_GEN || FM == 0, "the general ISAI is generated by its own launch");
    constexpr int SA = PCK == PCK_LU ? SM : 1, SC = PCK == PCK_ISAI ? SLU : 1, SG = PCK == PCK_GEN ? SM : 1;
    float a[SA];
    float cL[SC], cU[SC], cG[SG];

    for (;;) {
        int sys = 0;
        if (lane == 0) sys = atomicAdd(fp.counter, 1);
        sys = __builtin_amdgcn_readfirstlane(sys);
        if (sys >= batch) break;
        const int si = sysidx ? sysidx[sys] : sys;
        const int oi = outidx ? outidx[sys] : sys;

        const double *Ap = gA + (size_t)si * nnz;
        const double *bp = gb + (size_t)si * n;
        const float *Fp = gF + (size_t)sys * nnz;
        int fbad = 0;
        if constexpr (FM == 1) {
            // merged: the factors (LU: FP32 in place at F) or the ISAI values
            // (FP32 at M, the second half of the FP64 factor region) in LDS
            double *F64 = reinterpret_cast<double *>(wb);
            float *Dg = reinterpret_cast<float *>(wb + fp.wave_bytes) - fp.dg_floats;
            int status;
            fbad = factor_system<float>(fp, n, nnz, Ap, fa.Finit ? fa.Finit + (size_t)si * nnz : nullptr,
                                        fa.Y ? fa.Y + (size_t)sys * fa.P3 : nullptr, fa.mode, fa.sweeps,
                                        fa.stop_tol, PCK == PCK_ISAI ? 1 : 0, fa.inplace, nullptr, false, F64,
                                        F64 + (nnz + 4), Dg, lane, status);
            if (fa.status && lane == 0) fa.status[oi] = status;
            wave_sync();
            if constexpr (PCK == PCK_LU) {
                // the factors as the solve reads them: FP32 in place (F[nnz] = 0)
                round_in_place(F64, nnz + 4, lane);
                wave_sync();
            }
            Fp = reinterpret_cast<const float *>(wb) + (PCK == PCK_ISAI ? nnz + 4 : 0);
        } else if (PCK == PCK_LU) {
            wave_copy(F, Fp, nnz, lane);
        }
        if constexpr (PCK == PCK_ISAI) {
            // ISAI(L)'s unit diagonal (position -2) reads as 1; gathered before
            // A's slot values, which overlay them in the merged kernel
            slot_gather<SLU>(fp.isl_pos, fp.SL, Fp, lane, [&](int q, float v) { cL[q] = v; });
            slot_gather<SLU>(fp.isu_pos, fp.SU, Fp, lane, [&](int q, float v) { cU[q] = v; });
            if constexpr (FM == 1) wave_sync();
        }
        if constexpr (PCK == PCK_GEN) slot_gather<SM>(fp.slot_pos, S, Fp, lane, [&](int q, float v) { cG[q] = v; });
        slot_gather<SM>(fp.slot_pos, S, Ap, lane, [&](int q, double v) {
            if constexpr (PCK == PCK_LU) a[q] = (float)v;
            else Asl[q * WAVE + lane] = (float)v;
        });
        wave_sync();
        // zero cells of the FP32 views (the FP64 residual overwrites them)
        auto zero_cells = [&]() {
            if (lane == 0) { e0[n] = 0.0f; e1[n] = 0.0f; }
            wave_sync();
        };
        zero_cells();
        int bad = 0;
        Vec<NR, float> rinv;
        if constexpr (PCK == PCK_LU) {
#pragma unroll
            for (int k = 0; k < NR; k++) {
                const int i = lane + WAVE * k;
                const float u = (i < n) ? F[fp.diag[i]] : 1.0f;
                bad |= (u == 0.0f);
                rinv.e[k] = 1.0f / u;
            }
            bad = __any(bad);
        } else {
            bad = FM == 1 ? fbad : gZ[sys];
        }

        Vec<NR, double> x, r;
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const int i = lane + WAVE * k;
            x.e[k] = 0.0;
            r.e[k] = (i < n) ? bp[i] : 0.0;
        }
        const double bnorm = sqrt(vdot(r, r, lane, n));
        const double tol_abs = (mp.tol_mode == BILU_TOL_RELATIVE) ? mp.tol * bnorm : mp.tol;
        double rn = bnorm;
        float iter = 0.0f;
        int steps = 0;
        int conv = BILU_NOT_CONVERGED;

        // out = M^{-1} in, also left in e0 for the following SpMV
        auto apply = [&](const Vec<NR, float> &in, Vec<NR, float> &out) {
            if constexpr (PCK == PCK_LU) {
                Vec<NR, float> y;
                trsv_avail<NR, true, float>(trL, vcntL, fp.nlev_l, rlev, F, in, rinv, e1, accb, scratch, y, lane);
                trsv_avail<NR, false, float>(trU, vcntU, fp.nlev_u, rlev, F, y, rinv, e0, accb, scratch, out, lane);
                wave_sync();
            } else if constexpr (PCK == PCK_ISAI) {
                to_lds(in, e0, lane, n);
                wave_sync();
                spmv_tri<SC>(cL, fp.SL, lrec, e0, e1, thr, lane);
                wave_sync();
                spmv_tri<SC>(cU, fp.SU, urec, e1, e0, thr, lane);
                wave_sync();
                from_lds(out, e0, lane, n);
            } else {
                // rows of M v with A's records (row ends, scratch cell n + 1), e1 -> e0
                to_lds(in, e1, lane, n);
                wave_sync();
                spmv_tri<SG>(cG, S, meta, e1, e0, thr, lane);
                wave_sync();
                from_lds(out, e0, lane, n);
            }
        };
        auto matvec = [&](Vec<NR, float> &out) {   // out = A32 e0
            if constexpr (PCK == PCK_LU) spmv_slots<SM, float>(a, meta, e0, e1, scratch, lane);
            else spmv32<SM>(Asl, meta, e0, e1, thr, lane);
            wave_sync();
            from_lds(out, e1, lane, n);
            wave_sync();
        };

        if (bad) {
            conv = BILU_ZERO_PIVOT;
        } else if (!isfinite(bnorm)) {
            conv = BILU_NONFINITE;
        } else if (bnorm == 0.0) {
            conv = BILU_CONVERGED;
            rn = 0.0;
        } else if (rn < tol_abs) {
            conv = BILU_CONVERGED;
        } else {
            while (steps < mp.max_refine) {
                // inner FP32 BiCGSTAB on A d = r / |r|, d0 = 0
                const double sc = 1.0 / rn;
                Vec<NR, float> d, ri, rh, p, v, s, t, ph, sh;
#pragma unroll
                for (int k = 0; k < NR; k++) {
                    ri.e[k] = (float)(r.e[k] * sc);
                    rh.e[k] = ri.e[k];
                    d.e[k] = 0.0f; p.e[k] = 0.0f; v.e[k] = 0.0f;
                }
                const float r0 = sqrtf(vdot(ri, ri, lane, n));
                // the inner solve stops at the relative residual inner_tol, or looser
                // when that already takes |r| below the outer tolerance with the
                // margin inner_adapt (the last refinement step of most systems)
                double want = mp.inner_tol_d;
                if (mp.inner_adapt > 0.0 && mp.inner_adapt * tol_abs / rn > want) want = mp.inner_adapt * tol_abs / rn;
                const float itol = (float)want * r0;
                float ires = r0;
                float it = 0.0f;
                float rho_old = 1.0f, omega_old = 1.0f, alpha = 1.0f;
                bool inner_bad = false;
                while (it < (float)mp.max_iter && ires >= itol) {
                    const float rho_new = vdot(ri, rh, lane, n);
                    if (rho_new == 0.0f) break;
                    const float beta = (rho_new / rho_old) * (alpha / omega_old);
#pragma unroll
                    for (int k = 0; k < NR; k++) p.e[k] = fmaT(beta, fmaT(-omega_old, v.e[k], p.e[k]), ri.e[k]);
                    apply(p, ph);
                    matvec(v);
                    alpha = rho_new / vdot(rh, v, lane, n);
#pragma unroll
                    for (int k = 0; k < NR; k++) d.e[k] = fmaT(alpha, ph.e[k], d.e[k]);
#pragma unroll
                    for (int k = 0; k < NR; k++) s.e[k] = fmaT(-alpha, v.e[k], ri.e[k]);
                    ires = sqrtf(vdot(s, s, lane, n));
                    it += 0.5f;
                    if (!isfinite(ires)) { inner_bad = true; break; }
                    if (ires < itol) break;
                    apply(s, sh);
                    matvec(t);
                    const float ts = vdot(t, s, lane, n);
                    const float tt = vdot(t, t, lane, n);
                    const float omega_new = ts / tt;
#pragma unroll
                    for (int k = 0; k < NR; k++) d.e[k] = fmaT(omega_new, sh.e[k], d.e[k]);
                    it += 0.5f;
#pragma unroll
                    for (int k = 0; k < NR; k++) ri.e[k] = fmaT(-omega_new, t.e[k], s.e[k]);
                    ires = sqrtf(vdot(ri, ri, lane, n));
                    rho_old = rho_new;
                    omega_old = omega_new;
                    if (!isfinite(ires)) { inner_bad = true; break; }
                }
                iter += it;
                steps++;
                if (inner_bad) { conv = BILU_NONFINITE; break; }
                // x += |r| d; r = b - A x in FP64
#pragma unroll
                for (int k = 0; k < NR; k++) x.e[k] = fma((double)d.e[k], rn, x.e[k]);
                to_lds(x, d0, lane, n);
                wave_sync();
                Vec<NR, double> w;
                residual_rows<NR>(P.rp, P.ci, Ap, d0, w, n, lane);
#pragma unroll
                for (int k = 0; k < NR; k++) {
                    const int i = lane + WAVE * k;
                    r.e[k] = ((i < n) ? bp[i] : 0.0) - w.e[k];
                }
                wave_sync();
                zero_cells();
                rn = sqrt(vdot(r, r, lane, n));
                if (!isfinite(rn)) { conv = BILU_NONFINITE; break; }
                if (rn < tol_abs) { conv = BILU_CONVERGED; break; }
                if (it == 0.0f) break;   // the inner solve could not move (rho = 0)
            }
        }
        if (conv == BILU_ZERO_PIVOT) {
#pragma unroll
            for (int k = 0; k < NR; k++) x.e[k] = 0.0;
            rn = bnorm;
        }
        double *xp = gx + (size_t)si * n;
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const int i = lane + WAVE * k;
            if (i < n) xp[i] = x.e[k];
        }
        if (lane == 0) {
            g_iter[oi] = iter;
            g_conv[oi] = conv;
            if (g_ires) g_ires[oi] = (conv == BILU_ZERO_PIVOT) ? 0.0 : rn;
            if (g_tres) g_tres[oi] = rn;
            if (g_refine) g_refine[oi] = steps;
        }
        wave_sync();
    }
}
