// Persistent multi-wave solver kernel for the LU-preconditioned paths
// (exact ILU(0), ParILU(0), precomputed factors).  Included by solver.hip
// inside its anonymous namespace (uses DevPattern, SolveParams, Vec, vdot,
// to_lds, from_lds, wave_sum).
//
// Layout of one workgroup (W <= KF_MAX_WAVES waves, one system per wave at a time):
//   LDS  [ wave 0: F[nnz] fp64 | (ParILU: W[nnz]) | e0[n] | e1[n] ] [ wave 1 ... ]
//   global (cache-resident, shared by every wave): solve and factorization
//        schedules, SpMV slot tables
//   VGPR A values in the balanced slot layout (lane l, slot s <-> one CSR
//        position; whole rows per lane, so an SpMV row sum is a register-
//        only left-to-right fma chain over x gathered from LDS), Krylov
//        vectors in the row-cyclic layout, slot metadata.
// Waves take systems from a global counter (dynamic balance: systems need
// 5 to 14 iterations), so the grid is sized to the machine, not the batch.
// Waves of one workgroup never synchronise with each other: intra-wave
// ordering of LDS traffic is enough (a wave's LDS instructions execute in
// issue order), enforced against compiler reordering by wavefront-scope
// fences.
//
// The kernel is bound by LDS and VALU instruction issue, so every phase is
// laid out to make each LDS instruction carry useful lanes: triangular
// solves pull entries as their columns become available (no per-level
// row scans), Krylov vectors never round-trip through LDS except as the
// gather source of an SpMV or solve.

struct FastParams {
    const uint32_t *trL, *trU;            // solve chunk records [chunk][CH][64], tail-padded
    int trL_words, trU_words;
    const uint16_t *cntL, *cntU;          // per level: chunks | row-slot mask << 8, [64]
    const uint16_t *rlevL, *rlevU;        // [NR][64] level of the lane's rows
    const ushort4 *etab;                  // factorization entry records
    const uint16_t *wcode;                // per position: column j of a strictly lower entry, 0xFFFE an
                                          // upper entry with work, 0xFFFF no work
    int dg_floats;                        // k_factor_fast: floats at the end of a wave region for the diagonal
    const uint16_t *eptr;                 // per entry-DAG level
    const ushort2 *dep;                   // (pos(i,k), pos(k,j)) pairs
    const int *diag;                      // pattern diagonal positions
    int nlev_l, nlev_u, nlev_e, ne;
    int S;
    const uint32_t *slot_meta;  // [S][64]
    const int32_t *slot_pos;    // [S][64]
    int shared_bytes;           // shared LDS: SpMV meta [SM][64], trL, trU (16-byte aligned)
    int o_trL, o_trU;
    int wave_bytes;
    int f_cells;    // k_fast: FP64 cells of a wave region ahead of its vectors
    int par_w;      // 1: the wave region holds a second factor buffer (ParILU Jacobi)
    int *counter;
    // ISAI generation streams [step][64] and the slot triangles (ISAI applies)
    const uint32_t *isL0, *isL1, *isU0, *isU1;
    int isL_steps, isU_steps;
    // triangle slot layouts of the FP32 ISAI applies [SL or SU][64]
    const uint32_t *isl_meta, *isu_meta;
    const int32_t *isl_pos, *isu_pos;
    int SL, SU;
    // exact ILU(0) lane streams [step][64]
    const uint32_t *ilu0, *ilu1;
    int ilu_steps;
    // Jacobi ParILU(0) sweep rounds: products [step][64], finishes [round][64]
    const uint32_t *parp, *parf;
    int par_rounds, par_psteps;
    // the FP32 mixed ISAI kernel's slot records in the layout of its LDS copy
    // (null unless the pattern's template has 48 or 72 slots)
    const uint32_t *rec32;
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void k_zero_counter(int *c) { *c = 0; }

// A zero the compiler cannot see through.  Table addresses formed from it
// inside the persistent loop are recomputed there: otherwise the loop-
// invariant per-slot addresses of the slot tables are hoisted out of the
// loop, outnumber the registers, and every load first waits for its
// address to come back from scratch memory.
__device__ __forceinline__ int opaque_zero()
{
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}

// A system's values in a slot layout: for slots q < SMX, put(q, v) with
// v = V[pos[q][lane]] for q < ns (position -1: 0, position -2: 1, the unit
// diagonal of ISAI(L)) and v = 0 for ns <= q.  Every load is unconditional
// -- slots past ns re-read record ns - 1, negative positions read V[0] --
// and a batch of KB slots issues its position loads, then its value loads,
// together: a load that sits inside a branch on its own value waits alone,
// one memory round trip per slot.  pos holds at least one slot when ns > 0.
template <int SMX, typename TV, typename PUT>
__device__ __forceinline__ void slot_gather(const int32_t *__restrict__ pos, int ns, const TV *V, int lane,
                                            PUT &&put)
{
    constexpr int KB = 12;
    if (ns <= 0) {
#pragma unroll
        for (int q = 0; q < SMX; q++) put(q, TV(0));
        return;
    }
    // the last record's index is opaque, so the (uniform) record addresses are
    // formed per use
    const int last = ns - 1 + opaque_zero();
#pragma unroll
    for (int q0 = 0; q0 < SMX; q0 += KB) {
        int ps[KB];
        TV v[KB];
#pragma unroll
        for (int j = 0; j < KB; j++)
            if (q0 + j < SMX) {
                const int32_t *rec = pos + min(q0 + j, last) * WAVE;
                ps[j] = rec[lane];
            }
#pragma unroll
        for (int j = 0; j < KB; j++)
            if (q0 + j < SMX) v[j] = V[max(ps[j], 0)];
#pragma unroll
        for (int j = 0; j < KB; j++)
            if (q0 + j < SMX) {
                const bool in = q0 + j < ns;
                put(q0 + j, (in && ps[j] >= 0) ? v[j] : (in && ps[j] == -2) ? TV(1) : TV(0));
            }
    }
}

// waves per workgroup of the persistent kernel (its launch bound; the
// register budget per lane is 512 / (waves / 4) rounded to the allocation
// granule)
#ifndef KF_MAX_WAVES
#define KF_MAX_WAVES 8
#endif

// y (LDS, by row) = A x over the slot layout; x in LDS.  Slot metadata
// (8*col | 8*row << 16, high half 0xFFFF when the slot does not end a row)
// is read from the workgroup's shared LDS copy, padded to SM slots.  The
// body is straight-line: every slot stores its running sum, to the row's
// cell when the slot ends a row and to a scratch cell otherwise.
constexpr int SPMV_G = 3;   // slots per gather group (best of 1..8 measured at n = 144)
template <int SM, typename T = double>
__device__ __forceinline__ void spmv_slots(const T (&a)[SM], const uint32_t *__restrict__ meta,
                                           const T *__restrict__ x, T *__restrict__ y,
                                           T *__restrict__ scratch, int lane)
{
    // offsets in the tables are for 8-byte elements
    constexpr int SH = sizeof(T) == 4 ? 1 : 0;
    const char *xb = reinterpret_cast<const char *>(x);
    char *yb = reinterpret_cast<char *>(y);
    T acc = T(0);
    // the meta records and x values of the next SPMV_G slots are read before
    // this group's row stores (x and y never overlap), so the LDS latency of
    // the gathers overlaps the fma chain instead of following each store
    constexpr int G = SPMV_G;
    uint32_t m[SM];
    T xv[SM];
    auto fetch = [&](int s) {
        m[s] = meta[s * WAVE + lane];
        xv[s] = *reinterpret_cast<const T *>(xb + ((m[s] & 0xFFFFu) >> SH));
    };
#pragma unroll
    for (int s = 0; s < G && s < SM; s++) fetch(s);
#pragma unroll
    for (int s = 0; s < SM; s++) {
        if (s % G == 0) {
#pragma unroll
            for (int t = s + G; t < s + 2 * G && t < SM; t++) fetch(t);
        }
        acc = fmaT(a[s], xv[s], acc);
        const bool end = (m[s] >> 16) != 0xFFFFu;
        T *dst = end ? reinterpret_cast<T *>(yb + ((m[s] >> 16) >> SH)) : scratch;
        *dst = acc;
        acc = end ? T(0) : acc;
    }
}

// y = C x over a slot layout whose FP64 coefficients C sit in LDS
// slot-major (C[s 64 + lane]), ns slots (wave-uniform); lanes hold whole
// rows, each row an fma chain left to right from 0.  TRI: the triangle
// records of the ISAI applies (FP32 byte offsets, doubled here: x in bits
// 0-15, the output cell in bits 16-30 -- the scratch cell n + 1 of y while
// the row continues -- and bit 31 set while it continues); otherwise A's slot
// records (FP64 byte offsets: x in bits 0-15, bits 16-31 the row's cell or
// 0xFFFF while the row continues, its running sum then going to `scratch`).
// Records, coefficients and x values of G slots are read ahead of their fmas.
template <bool TRI>
__device__ __forceinline__ void spmv_coef(const double *__restrict__ C, int ns, const uint32_t *__restrict__ rec,
                                          const double *__restrict__ x, double *__restrict__ y,
                                          double *__restrict__ scratch, int lane)
{
    constexpr int G = 4;
    const char *xb = reinterpret_cast<const char *>(x);
    char *yb = reinterpret_cast<char *>(y);
    double acc = 0.0;
    for (int s0 = 0; s0 < ns; s0 += G) {
        uint32_t r[G];
        double xv[G], cv[G];
#pragma unroll
        for (int j = 0; j < G; j++) {
            const int s = min(s0 + j, ns - 1);   // past the end: the last slot again, not used
            r[j] = rec[s * WAVE + lane];
            cv[j] = C[s * WAVE + lane];
            xv[j] = *reinterpret_cast<const double *>(xb + (TRI ? (r[j] & 0xFFFFu) << 1 : (r[j] & 0xFFFFu)));
        }
#pragma unroll
        for (int j = 0; j < G; j++) {
            if (s0 + j < ns) {
                acc = fma(cv[j], xv[j], acc);
                bool end;
                double *dst;
                if constexpr (TRI) {
                    end = !(r[j] >> 31);
                    dst = reinterpret_cast<double *>(yb + (((r[j] >> 16) & 0x7FFFu) << 1));
                } else {
                    end = (r[j] >> 16) != 0xFFFFu;
                    dst = end ? reinterpret_cast<double *>(yb + (r[j] >> 16)) : scratch;
                }
                *dst = acc;
                acc = end ? 0.0 : acc;
            }
        }
    }
}

// Triangular solve by availability.  Lane l owns rows l + 64 k; their
// partial sums live in FP64 LDS cells acc[row] (per wave; for FP32 solves
// too: the FP32 products are added into FP64 cells, rounded to FP32 once
// when the row is final -- the LDS FP32 atomic add runs about 20x slower
// than the FP64 one on gfx950).  Level ℓ: (1) every
// row of level ℓ is final: L: y_r = vin_r - acc_r, U: x_r = (vin_r - acc_r)
// * (1 / u_rr), kept in `res` and stored to `out` (LDS) for the gathers of
// later levels; (2) each lane adds the products of its entries whose column
// lies in level ℓ into its rows' cells, acc_r += F_p * out_c, with the
// LDS floating-point atomic (fire and forget; a lane only touches its own
// rows, so the additions of one row land in program order).  Records (LDS
// byte offsets + row slot k) come in chunks of CH per level; padding records
// point at zero cells, so every chunk body is straight-line code.  Lane ℓ of
// `vcnt` holds level ℓ's chunk count (low byte) and the mask of the row slots
// k that hold a row of level ℓ (bits 8..); it is read with readlane (no
// memory access on the loop bound), and row slots outside the mask skip the
// finalisation.  The next chunk's records are loaded one chunk ahead.
template <int NR, bool LOWER, typename T = double>
__device__ __forceinline__ void trsv_avail(const uint32_t *__restrict__ rec, int vcnt, int nlev,
                                           const int (&rlev)[NR], const T *__restrict__ F,
                                           const Vec<NR, T> &vin, const Vec<NR, T> &rinv, T *out,
                                           double *acc, T *scratch, Vec<NR, T> &res, int lane)
{
    constexpr int CH = BILU_SOLVE_CH;
    constexpr int SH = sizeof(T) == 4 ? 1 : 0;   // record offsets are for 8-byte elements
    const char *Fb = reinterpret_cast<const char *>(F);
    const char *ob = reinterpret_cast<const char *>(out);
    char *ab = reinterpret_cast<char *>(acc) + lane * 8;   // row lane + 64 k at k * 512
#pragma unroll
    for (int k = 0; k < NR; k++) acc[lane + WAVE * k] = 0.0;
    uint32_t r[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) r[j] = rec[j * WAVE + lane];
    int chunk = 0;
    wave_sync();
    for (int l = 0; l < nlev; l++) {
        const int cm = __builtin_amdgcn_readlane(vcnt, l);
        const int nch = cm & 0xFF;
#pragma unroll
        for (int k = 0; k < NR; k++) {
            if (!(cm & (0x100 << k))) continue;   // no row of this level in slot k
            const bool mine = (LOWER ? (rlev[k] & 0xFFFF) : (rlev[k] >> 16)) == l;
            const T a = (T)acc[lane + WAVE * k];
            const T v = LOWER ? vin.e[k] - a : (vin.e[k] - a) * rinv.e[k];
            res.e[k] = mine ? v : res.e[k];
            T *dst = mine ? out + (lane + WAVE * k) : scratch;
            *dst = v;
        }
        wave_sync();
        for (int c = 0; c < nch; c++, chunk++) {
            uint32_t rn[CH];
#pragma unroll
            for (int j = 0; j < CH; j++) rn[j] = rec[((chunk + 1) * CH + j) * WAVE + lane];
            T f[CH], yv[CH];
#pragma unroll
            for (int j = 0; j < CH; j++) {
                f[j] = *reinterpret_cast<const T *>(Fb + ((r[j] & 0xFFFFu) >> SH));
                yv[j] = *reinterpret_cast<const T *>(ob + (((r[j] >> 16) & 0xFFF8u) >> SH));
            }
#pragma unroll
            for (int j = 0; j < CH; j++) {
                double *cell = reinterpret_cast<double *>(ab + (((r[j] >> 16) & 7u) << 9));
                __hip_atomic_fetch_add(cell, (double)(f[j] * yv[j]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
#pragma unroll
            for (int j = 0; j < CH; j++) r[j] = rn[j];
        }
        wave_sync();
    }
}

// Exact ILU(0) in place by the lane streams of the pattern plan: each step
// folds one dependency pair acc = fma(-l_ik, u_kj, acc) of the lane's
// current entry (acc starts at a_ij on the entry's first step; the k order
// is the reference KIJ order, merged.cu:168-219), and an entry's last step
// writes it (lower entries divided by u_jj).  Every value is read at least
// ILU_GAP steps after its write, so a step's LDS operands are loaded one
// step ahead, before the previous step's store; records arrive in batches of
// ILU_BATCH steps, the next batch in flight during the current one.
// Returns nonzero on a zero pivot.
__device__ __forceinline__ int ilu_stream(const uint32_t *__restrict__ r0p, const uint32_t *__restrict__ r1p,
                                          int steps, double *F, double *scratch, const int *__restrict__ diag,
                                          int n, int lane)
{
    constexpr int K = bilu::ILU_BATCH;
    char *Fb = reinterpret_cast<char *>(F);
    auto ld = [&](uint32_t off) { return *reinterpret_cast<const double *>(Fb + off); };
    uint32_t ra[K], rb[K];
#pragma unroll
    for (int j = 0; j < K; j++) { ra[j] = r0p[j * WAVE + lane]; rb[j] = r1p[j * WAVE + lane]; }
    double acc = 0.0;
    int bad = 0;
    for (int s0 = 0; s0 < steps; s0 += K) {
        const int nx = (s0 + K < steps) ? s0 + K : s0;
        uint32_t na[K], nb[K];
#pragma unroll
        for (int j = 0; j < K; j++) { na[j] = r0p[(nx + j) * WAVE + lane]; nb[j] = r1p[(nx + j) * WAVE + lane]; }
        double l = ld(ra[0] & 0xFFFFu), u = ld(ra[0] >> 16), in = ld(rb[0] & 0xFFF8u);
        uint32_t pv = rb[0] >> 16;
        double dv = ld(pv == 0xFFFFu ? 0u : pv);
#pragma unroll
        for (int j = 0; j < K; j++) {
            double l1 = 0.0, u1 = 0.0, in1 = 0.0, dv1 = 0.0;
            uint32_t pv1 = 0xFFFFu;
            if (j + 1 < K) {
                l1 = ld(ra[j + 1] & 0xFFFFu);
                u1 = ld(ra[j + 1] >> 16);
                in1 = ld(rb[j + 1] & 0xFFF8u);
                pv1 = rb[j + 1] >> 16;
                dv1 = ld(pv1 == 0xFFFFu ? 0u : pv1);
            }
            const uint32_t b = rb[j];
            acc = (b & 1u) ? in : acc;
            acc = fma(-l, u, acc);
            const bool div = pv != 0xFFFFu, end = (b & 2u) != 0;
            const double val = div ? acc / dv : acc;
            bad |= (end && div && dv == 0.0);
            double *dst = end ? reinterpret_cast<double *>(Fb + (b & 0xFFF8u)) : scratch;
            *dst = val;
            l = l1; u = u1; in = in1; pv = pv1; dv = dv1;
        }
#pragma unroll
        for (int j = 0; j < K; j++) { ra[j] = na[j]; rb[j] = nb[j]; }
    }
    wave_sync();
    // every U diagonal must be nonzero for the solves
    for (int k = lane; k < n; k += WAVE) bad |= (F[diag[k]] == 0.0);
    return __any(bad);
}

// One Jacobi ParILU(0) sweep (parILU_0_gpu.cu:18-57 formula) in
// lane-synchronous rounds: src = the current iterate, dst = the next one
// (entries without work hold a_ij in both).  In round r lane l owns entry
// 64 r + l of the sweep order; the round's product steps accumulate
// sum = fma(l_ik, u_kj, sum) (zero pairs pad the shorter entries) and its
// last step finishes every lane's entry at once: t = a_ij - sum, dst = t /
// src[pivot] (lower) or t (upper), residual t - src*pivot (or t - src)
// squared into the lane's sum.  Every operand comes from src.  Product
// records arrive in batches, one batch ahead; finish records and a_ij (global
// memory, L2 hits after the first sweep) are queued PAR_AHEAD rounds ahead.
// Returns the lane's sum of squared residuals (RES) or 0.
#ifndef PAR_AHEAD_ROUNDS
#define PAR_AHEAD_ROUNDS 2
#endif
constexpr int PAR_AHEAD = PAR_AHEAD_ROUNDS;
template <bool RES>
__device__ __forceinline__ double parilu_rounds(const uint32_t *__restrict__ pp, const uint32_t *__restrict__ pf,
                                                int R, int PS, const double *__restrict__ Ap, const double *src,
                                                double *dst, int lane)
{
    constexpr int K = bilu::PAR_BATCH, D = PAR_AHEAD;
    const char *Sb = reinterpret_cast<const char *>(src);
    char *Db = reinterpret_cast<char *>(dst);
    auto ld = [&](uint32_t off) { return *reinterpret_cast<const double *>(Sb + off); };
    auto frec = [&](int r) { return pf[(r < R ? r : R - 1) * WAVE + lane]; };
    uint32_t fq[D + 1];
    double aq[D];
#pragma unroll
    for (int k = 0; k <= D; k++) fq[k] = frec(k);
#pragma unroll
    for (int k = 0; k < D; k++) aq[k] = Ap[(fq[k] & 0xFFF8u) >> 3];
    int r = 0;
    uint32_t ra[K];
#pragma unroll
    for (int j = 0; j < K; j++) ra[j] = pp[j * WAVE + lane];
    double acc = 0.0, rr = 0.0;
    for (int s0 = 0; s0 < PS; s0 += K) {
        const int nx = (s0 + K < PS) ? s0 + K : s0;
        uint32_t na[K];
#pragma unroll
        for (int j = 0; j < K; j++) na[j] = pp[(nx + j) * WAVE + lane];
        // the batch's operands first: src is never written during the sweep
        double lv[K], uv[K];
#pragma unroll
        for (int j = 0; j < K; j++) { lv[j] = ld(ra[j] & 0xFFF8u); uv[j] = ld(ra[j] >> 16); }
#pragma unroll
        for (int j = 0; j < K; j++) {
            const uint32_t rec = ra[j];
            acc = fma(lv[j], uv[j], acc);
            if (__builtin_amdgcn_readfirstlane(rec) & 1u) {   // the round's last step (all lanes)
                const uint32_t f = fq[0], pos = f & 0xFFF8u, pv = f >> 16;
                const bool valid = (f & 1u) != 0, div = pv != 0xFFFFu;
                const double old = ld(pos), dv = ld(div ? pv : pos);
                const double t = aq[0] - acc;
                const double val = div ? t / dv : t;
                if (RES) {
                    const double res = div ? t - old * dv : t - old;
                    rr = valid ? fma(res, res, rr) : rr;
                }
                if (valid) *reinterpret_cast<double *>(Db + pos) = val;
                acc = 0.0;
                r++;
#pragma unroll
                for (int k = 0; k < D; k++) fq[k] = fq[k + 1];
                fq[D] = frec(r + D);
#pragma unroll
                for (int k = 0; k + 1 < D; k++) aq[k] = aq[k + 1];
                aq[D - 1] = Ap[(fq[D - 1] & 0xFFF8u) >> 3];
            }
        }
#pragma unroll
        for (int j = 0; j < K; j++) ra[j] = na[j];
    }
    return rr;
}

// One in-place ParILU(0) sweep (parILU_0_gpu.cu:18-57 formula; the
// reference also sweeps in place, in an unspecified order) over the same
// lane-synchronous rounds: round r reads F as the earlier rounds of the sweep
// left it and writes its entries at its last step, so its lanes read the
// round's starting values.  A wave's LDS instructions execute in program
// order, so the loads of the steps after a round end are issued after that
// round's writes: at each segment start (the batch start, or the step after
// a round end) the operands of every step up to the next round end in the
// batch are loaded together.  The round ends of a batch are wave-uniform
// (bit 0 of every lane's record).  Returns the lane's sum of squared
// residuals, each as its round saw it (RES), or 0.
template <bool RES>
__device__ __forceinline__ double parilu_inplace(const uint32_t *__restrict__ pp, const uint32_t *__restrict__ pf,
                                                 int R, int PS, const double *__restrict__ Ap, double *F, int lane)
{
    constexpr int K = bilu::PAR_BATCH, D = PAR_AHEAD;
    char *Fb = reinterpret_cast<char *>(F);
    auto ld = [&](uint32_t off) { return *reinterpret_cast<const double *>(Fb + off); };
    auto frec = [&](int r) { return pf[(r < R ? r : R - 1) * WAVE + lane]; };
    uint32_t fq[D + 1];
    double aq[D];
#pragma unroll
    for (int k = 0; k <= D; k++) fq[k] = frec(k);
#pragma unroll
    for (int k = 0; k < D; k++) aq[k] = Ap[(fq[k] & 0xFFF8u) >> 3];
    int r = 0;
    uint32_t ra[K];
#pragma unroll
    for (int j = 0; j < K; j++) ra[j] = pp[j * WAVE + lane];
    double acc = 0.0, rr = 0.0;
    for (int s0 = 0; s0 < PS; s0 += K) {
        const int nx = (s0 + K < PS) ? s0 + K : s0;
        uint32_t na[K];
#pragma unroll
        for (int j = 0; j < K; j++) na[j] = pp[(nx + j) * WAVE + lane];
        uint32_t fin = 0;
#pragma unroll
        for (int j = 0; j < K; j++) fin |= (__builtin_amdgcn_readfirstlane(ra[j]) & 1u) << j;
        double lv[K], uv[K];
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j == 0 || ((fin >> (j - 1)) & 1u)) {   // segment start: load up to the next round end
#pragma unroll
                for (int t = j; t < K; t++) {
                    lv[t] = ld(ra[t] & 0xFFF8u);
                    uv[t] = ld(ra[t] >> 16);
                    if ((fin >> t) & 1u) break;
                }
            }
            acc = fma(lv[j], uv[j], acc);
            if ((fin >> j) & 1u) {   // the round's last step (all lanes)
                const uint32_t f = fq[0], pos = f & 0xFFF8u, pv = f >> 16;
                const bool valid = (f & 1u) != 0, div = pv != 0xFFFFu;
                const double old = ld(pos), dv = ld(div ? pv : pos);
                const double t = aq[0] - acc;
                const double val = div ? t / dv : t;
                if (RES) {
                    const double res = div ? t - old * dv : t - old;
                    rr = valid ? fma(res, res, rr) : rr;
                }
                if (valid) *reinterpret_cast<double *>(Fb + pos) = val;
                acc = 0.0;
                r++;
#pragma unroll
                for (int k = 0; k < D; k++) fq[k] = fq[k + 1];
                fq[D] = frec(r + D);
#pragma unroll
                for (int k = 0; k + 1 < D; k++) aq[k] = aq[k + 1];
                aq[D - 1] = Ap[(fq[D - 1] & 0xFFF8u) >> 3];
            }
        }
#pragma unroll
        for (int j = 0; j < K; j++) ra[j] = na[j];
    }
    return rr;
}

// Jacobi ParILU(0) sweep (parILU_0_gpu.cu:18-57 formula) entry by entry
// (entry idx on lane idx % 64), W = the old iterate; entries without work
// keep a_ij.  The register-lean form for the merged FP64 kernel, whose slot
// values stay live in VGPRs across the factorization.
__device__ __forceinline__ void parilu_sweep(const ushort4 *__restrict__ etab, int ne,
                                             const ushort2 *__restrict__ dep, const double *__restrict__ Ap,
                                             double *F, double *W, int nnz, int lane)
{
    for (int e = lane; e < nnz; e += WAVE) W[e] = F[e];
    wave_sync();
    for (int idx = lane; idx < ne; idx += WAVE) {
        const ushort4 r = etab[idx];
        const int pos = r.x, db = r.y, de = r.z, pv = r.w;
        const double a = Ap[pos];
        const double dv = W[pv == 0xFFFF ? pos : pv];
        double sum = 0.0;
        for (int d = db; d < de; d++) {
            const ushort2 pq = dep[d];
            sum = fma(W[pq.x], W[pq.y], sum);
        }
        F[pos] = (pv != 0xFFFF) ? (a - sum) / dv : a - sum;
    }
    wave_sync();
}

// Preconditioners of k_fast: FPC_LU = the in-place LU factors (built in the
// kernel or loaded), applied by triangular solves; FPC_ISAI = ISAI(L) and
// ISAI(U) values on the pattern (strictly lower = ISAI(L), its unit diagonal
// implicit; diagonal and upper = ISAI(U)) with zero-pivot flags, applied as
// two SpMVs over the triangle slot layouts; FPC_GEN = the general ISAI of A
// on its pattern with flags, applied as an SpMV over A's slot layout.  The
// ISAI values come from a separate generation launch.
constexpr int FPC_LU = 0, FPC_ISAI = 1, FPC_GEN = 2;

template <int NR, int SM, int PC = FPC_LU>
__global__ __launch_bounds__(KF_MAX_WAVES * WAVE) void k_fast(DevPattern P, FastParams fp, int batch,
                                              const double *__restrict__ gA, const double *__restrict__ gb,
                                              double *__restrict__ gx, const double *__restrict__ gF,
                                              SolveParams sp, float *__restrict__ g_iter,
                                              int *__restrict__ g_conv, double *__restrict__ g_ires,
                                              double *__restrict__ g_tres, const int *__restrict__ gZ)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int n = P.n, nnz = P.nnz, S = fp.S;
    // shared: SpMV slot metadata [SM][64] (padded), FPC_ISAI: then the
    // records of the two triangle layouts [SL + SU][64]; LU solve chunk records
    // are read from global memory (cache-resident tables shared by every wave)
    uint32_t *meta = reinterpret_cast<uint32_t *>(smem);
    uint32_t *trec = meta + SM * WAVE;
    const uint32_t *trL = fp.trL;
    const uint32_t *trU = fp.trU;
    for (int i = threadIdx.x; i < SM * WAVE; i += blockDim.x) meta[i] = (i < S * WAVE) ? fp.slot_meta[i] : 0xFFFF0000u;
    if constexpr (PC == FPC_ISAI) {
        const int nl = fp.SL * WAVE, nu = fp.SU * WAVE;
        for (int i = threadIdx.x; i < nl; i += blockDim.x) trec[i] = fp.isl_meta[i];
        for (int i = threadIdx.x; i < nu; i += blockDim.x) trec[nl + i] = fp.isu_meta[i];
    }
    // per wave: f_cells FP64 cells -- FPC_LU: F[nnz + 2] (F[nnz] = 0), (ParILU:
    // W[nnz + 2]); FPC_ISAI / FPC_GEN: the coefficients slot-major -- then
    // e0[n + 2], e1[n + 2], (FPC_LU: acc[64 NR]) (e[n] = 0 is the zero cell of
    // padding records, e[n + 1] a scratch cell)
    double *F = reinterpret_cast<double *>(smem + fp.shared_bytes + (size_t)wave * fp.wave_bytes);
    double *Wb = F + (nnz + 2);                       // ParILU only
    double *e0 = F + fp.f_cells, *e1 = e0 + (n + 2);
    double *scratch = e1 + (n + 1);
    double *accb = e1 + (n + 2);
    if (lane == 0) {
        if constexpr (PC == FPC_LU) F[nnz] = 0.0;
        e0[n] = 0.0;
        e1[n] = 0.0;
    }
    __syncthreads();
    // level of the lane's rows in the lower (low half) and upper (high half) solve
    int rlev[NR];
#pragma unroll
    for (int k = 0; k < NR; k++)
        rlev[k] = (int)fp.rlevL[k * WAVE + lane] | ((int)fp.rlevU[k * WAVE + lane] << 16);
    // chunks per level: lane l holds the count of level l
    const int vcntL = fp.cntL[lane], vcntU = fp.cntU[lane];

    double a[SM];

    for (;;) {
        int sys = 0;
        if (lane == 0) sys = atomicAdd(fp.counter, 1);
        sys = __builtin_amdgcn_readfirstlane(sys);
        if (sys >= batch) break;

        const double *Ap = gA + (size_t)sys * nnz;
        const double *bp = gb + (size_t)sys * n;
        Vec<NR> x, r, rh, p, v, s, t, ph, sh;
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const int i = lane + WAVE * k;
            r.e[k] = (i < n) ? bp[i] : 0.0;
        }
        int bad = 0;
        if constexpr (PC != FPC_LU) {
            // A's slot values, and the ISAI values gathered slot-major into the
            // wave's region (ISAI(L)'s unit diagonal = 1, padding = 0)
            const double *Mp = gF + (size_t)sys * nnz;
            slot_gather<SM>(fp.slot_pos, S, Ap, lane, [&](int q, double v) { a[q] = v; });
            // runtime slot counts: batches of U slots, loads unconditional (clamped
            // record and position) so that a batch's loads are in flight together
            auto gather = [&](const int32_t *__restrict__ pos, int cnt, double *dst) {
                constexpr int U = 8;
                for (int s0 = 0; s0 < cnt; s0 += U) {
                    int ps[U];
                    double v[U];
#pragma unroll
                    for (int j = 0; j < U; j++) ps[j] = pos[min(s0 + j, cnt - 1) * WAVE + lane];
#pragma unroll
                    for (int j = 0; j < U; j++) v[j] = Mp[max(ps[j], 0)];
#pragma unroll
                    for (int j = 0; j < U; j++) v[j] = ps[j] >= 0 ? v[j] : (ps[j] == -2 ? 1.0 : 0.0);
#pragma unroll
                    for (int j = 0; j < U; j++)
                        if (s0 + j < cnt) dst[(s0 + j) * WAVE + lane] = v[j];
                }
            };
            if constexpr (PC == FPC_ISAI) {
                gather(fp.isl_pos, fp.SL, F);
                gather(fp.isu_pos, fp.SU, F + fp.SL * WAVE);
            } else {
                gather(fp.slot_pos, S, F);
            }
            bad = gZ[sys] != 0;
            wave_sync();
        } else if (sp.factor == FK_LOAD) {
            const double *Fp = gF + (size_t)sys * nnz;
            wave_copy(F, Fp, nnz, lane);
            slot_gather<SM>(fp.slot_pos, S, Ap, lane, [&](int q, double v) { a[q] = v; });
            wave_sync();
        } else {
            wave_copy(F, Ap, nnz, lane);
            wave_sync();
            slot_gather<SM>(fp.slot_pos, S, (const double *)F, lane, [&](int q, double v) { a[q] = v; });
            if (sp.factor == FK_ILU) {
                bad = ilu_stream(fp.ilu0, fp.ilu1, fp.ilu_steps, F, F + nnz + 1, fp.diag, n, lane);
            } else if (sp.factor == FK_PARILU) {
                for (int sw = 0; sw < sp.sweeps; sw++) parilu_sweep(fp.etab, fp.ne, fp.dep, Ap, F, Wb, nnz, lane);
            }
        }
        if (PC == FPC_LU && sp.factor != FK_ILU) {
            // a zero on U's diagonal is a zero pivot, whatever produced the factors
            // (exact ILU(0) checks its pivots as it divides)
            wave_sync();
            for (int k = lane; k < n; k += WAVE) bad |= (F[fp.diag[k]] == 0.0);
            bad = __any(bad);
        }

        float iter = 0.0f;
        int conv = BILU_NOT_CONVERGED;
#pragma unroll
        for (int k = 0; k < NR; k++) {
            x.e[k] = 0.0; v.e[k] = 0.0; p.e[k] = 0.0; rh.e[k] = r.e[k];
            ph.e[k] = 0.0; sh.e[k] = 0.0; s.e[k] = 0.0; t.e[k] = 0.0;
        }
        const double bnorm = sqrt(vdot(r, r, lane, n));   // r0 = b
        const double res0 = bnorm;
        const double tol_abs = (sp.tol_mode == BILU_TOL_RELATIVE) ? sp.tol * bnorm : sp.tol;
        double iter_res = res0;

        // reciprocal U diagonal of the lane's rows (multipliers of the upper solve)
        Vec<NR> rinv;
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const int i = lane + WAVE * k;
            rinv.e[k] = (PC == FPC_LU && i < n) ? 1.0 / F[fp.diag[i]] : 1.0;
        }
        // out = M in (LU: U^{-1} L^{-1} in), in registers; also left in e0 for
        // the following SpMV
        auto apply = [&](const Vec<NR> &in, Vec<NR> &out) {
            if constexpr (PC == FPC_LU) {
                Vec<NR> y;
                trsv_avail<NR, true>(trL, vcntL, fp.nlev_l, rlev, F, in, rinv, e1, accb, scratch, y, lane);
                trsv_avail<NR, false>(trU, vcntU, fp.nlev_u, rlev, F, y, rinv, e0, accb, scratch, out, lane);
                wave_sync();
            } else if constexpr (PC == FPC_ISAI) {
                // ilu_isai_solve.cu:25-45: t = ISAI(L) in, out = ISAI(U) t
                to_lds(in, e0, lane, n);
                wave_sync();
                spmv_coef<true>(F, fp.SL, trec, e0, e1, nullptr, lane);
                wave_sync();
                spmv_coef<true>(F + fp.SL * WAVE, fp.SU, trec + fp.SL * WAVE, e1, e0, nullptr, lane);
                wave_sync();
                from_lds(out, e0, lane, n);
                wave_sync();
            } else {
                // gisai_solve.cu:23-30: out = G in
                to_lds(in, e1, lane, n);
                wave_sync();
                spmv_coef<false>(F, S, meta, e1, e0, e1 + (n + 1), lane);
                wave_sync();
                from_lds(out, e0, lane, n);
                wave_sync();
            }
        };
        auto matvec = [&](Vec<NR> &out) {  // out = A e0
            spmv_slots<SM>(a, meta, e0, e1, scratch, lane);
            wave_sync();
            from_lds(out, e1, lane, n);
            wave_sync();
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
                apply(p, ph);
                matvec(v);
                alpha = rho_new / vdot(rh, v, lane, n);
#pragma unroll
                for (int k = 0; k < NR; k++) s.e[k] = fma(-alpha, v.e[k], r.e[k]);
                iter_res = sqrt(vdot(s, s, lane, n));
                iter += 0.5f;
                if (!isfinite(iter_res)) { conv = BILU_NONFINITE; break; }
                // x + alpha*ph now (x keeps the last finite iterate on a break
                // above): the same fma the combined update performs first
#pragma unroll
                for (int k = 0; k < NR; k++) x.e[k] = fma(alpha, ph.e[k], x.e[k]);
                const bool half_conv = iter_res < tol_abs;
                if (!half_conv) {
                    apply(s, sh);
                    matvec(t);
                    const double ts = vdot(t, s, lane, n);
                    const double tt = vdot(t, t, lane, n);
                    const double omega_new = ts / tt;
#pragma unroll
                    for (int k = 0; k < NR; k++) x.e[k] = fma(omega_new, sh.e[k], x.e[k]);
                    iter += 0.5f;
#pragma unroll
                    for (int k = 0; k < NR; k++) r.e[k] = fma(-omega_new, t.e[k], s.e[k]);
                    iter_res = sqrt(vdot(r, r, lane, n));
                    rho_old = rho_new;
                    omega_old = omega_new;
                    if (!isfinite(iter_res)) { conv = BILU_NONFINITE; break; }
                }
                if (half_conv || iter_res < tol_abs) {
                    conv = BILU_CONVERGED;
                    if (!sp.restart || restarts >= sp.max_restarts) break;
                    to_lds(x, e0, lane, n);
                    wave_sync();
                    Vec<NR> w;
                    matvec(w);
#pragma unroll
                    for (int k = 0; k < NR; k++) {
                        const int i = lane + WAVE * k;
                        w.e[k] = ((i < n) ? bp[i] : 0.0) - w.e[k];
                    }
                    const double tr = sqrt(vdot(w, w, lane, n));
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
        // fused epilogue: true residual from the register-resident A
        wave_sync();
        to_lds(x, e0, lane, n);
        wave_sync();
        Vec<NR> w;
        matvec(w);
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const int i = lane + WAVE * k;
            w.e[k] = ((i < n) ? bp[i] : 0.0) - w.e[k];
        }
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
        wave_sync();
    }
}
