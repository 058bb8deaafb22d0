// Persistent multi-wave solver kernel for the LU-preconditioned paths
// (exact ILU(0), ParILU(0), precomputed factors).  Included by solver.hip
// inside its anonymous namespace (uses DevPattern, SolveParams, Vec, vdot,
// to_lds, from_lds, wave_sum).
//
// Layout of one workgroup (W waves, W <= 8, one system per wave at a time):
//   LDS  [ schedule blob (u16 tables, shared by all waves) ]
//        [ wave 0: F[nnz] fp64 | e0[n] | e1[n] | e2[n] ] [ wave 1 ... ]
//   VGPR A values in the balanced slot layout (lane l, slot s <-> one CSR
//        position; whole rows per lane, so an SpMV row sum is a register-
//        only left-to-right fma chain over x gathered from LDS), Krylov
//        vectors in the row-cyclic layout, slot metadata.
// Waves take systems from a global counter (dynamic balance: systems need
// 5 to 14 iterations), so the grid is sized to the machine, not the batch.
// Waves of one workgroup never synchronise with each other after the blob
// copy: intra-wave ordering of LDS traffic is enough (a wave's LDS
// instructions execute in issue order), enforced against compiler
// reordering by wavefront-scope fences.

struct FastParams {
    const uint8_t *blob;
    int blob_bytes;
    int o_ci, o_diag, o_ldesc, o_udesc, o_lptr, o_uptr, o_upd_ptr, o_upd, o_dep_ptr, o_dep, o_pdiv;
    int nlev_l, nlev_u;
    int S;
    const uint32_t *slot_meta;  // [S][64]
    const int32_t *slot_pos;    // [S][64]
    int wave_bytes;
    int par_w;      // 1: the wave region holds a second factor buffer (ParILU Jacobi)
    int *counter;
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void k_zero_counter(int *c) { *c = 0; }

// y (LDS, by row) = A x over the slot layout; x in LDS
template <int SM>
__device__ __forceinline__ void spmv_slots(const double (&a)[SM], const uint32_t (&meta)[SM], int S,
                                           const double *x, double *y)
{
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < SM; s++) {
        if (s < S) {
            uint32_t m = meta[s];
            // opaque copy: keeps the index extraction here instead of letting
            // it be hoisted into long-lived registers across the whole solve
            asm volatile("" : "+v"(m));
            acc = fma(a[s], x[m & 0xFFFFu], acc);
            if ((m >> 16) != 0xFFFFu) {
                y[m >> 16] = acc;
                acc = 0.0;
            }
        }
    }
}

// y = L^{-1} vin (unit diagonal), rows level by level
__device__ __forceinline__ void lsolve_fast(const ushort4 *desc, const uint16_t *lptr, int nlev,
                                            const uint16_t *ci, const double *F, const double *vin,
                                            double *y, int lane)
{
    for (int l = 0; l < nlev; l++) {
        const int b = lptr[l], e = lptr[l + 1];
        for (int idx = b + lane; idx < e; idx += WAVE) {
            const ushort4 d = desc[idx];
            const int i = d.x, p0 = d.y, pd = d.z;
            double acc = 0.0;
            for (int p = p0; p < pd; p++) acc = fma(F[p], y[ci[p]], acc);
            y[i] = vin[i] - acc;
        }
        wave_sync();
    }
}

// out = U^{-1} y, rows level by level from the bottom; sums right to left
__device__ __forceinline__ void usolve_fast(const ushort4 *desc, const uint16_t *uptr, int nlev,
                                            const uint16_t *ci, const double *F, const double *y,
                                            double *out, int lane)
{
    for (int l = 0; l < nlev; l++) {
        const int b = uptr[l], e = uptr[l + 1];
        for (int idx = b + lane; idx < e; idx += WAVE) {
            const ushort4 d = desc[idx];
            const int i = d.x, pd = d.z, pe = d.w;
            double acc = 0.0;
            for (int p = pe - 1; p > pd; p--) acc = fma(F[p], out[ci[p]], acc);
            out[i] = (y[i] - acc) / F[pd];
        }
        wave_sync();
    }
}

// exact ILU(0) in place, IKJ per row, rows level by level; nonzero on a zero pivot
__device__ __forceinline__ int ilu_fast(const ushort4 *desc, const uint16_t *lptr, int nlev,
                                        const uint16_t *ci, const uint16_t *diag,
                                        const uint16_t *upd_ptr, const ushort2 *upd, int n,
                                        double *F, int lane)
{
    int bad = 0;
    for (int l = 0; l < nlev; l++) {
        const int b = lptr[l], e = lptr[l + 1];
        for (int idx = b + lane; idx < e; idx += WAVE) {
            const ushort4 d = desc[idx];
            const int p0 = d.y, pd = d.z;
            for (int p = p0; p < pd; p++) {
                const double ukk = F[diag[ci[p]]];
                bad |= (ukk == 0.0);
                const double lik = F[p] / ukk;
                F[p] = lik;
                const int ue = upd_ptr[p + 1];
                for (int u = upd_ptr[p]; u < ue; u++) {
                    const ushort2 du = upd[u];
                    F[du.x] = fma(-lik, F[du.y], F[du.x]);
                }
            }
        }
        wave_sync();
    }
    for (int i = lane; i < n; i += WAVE) bad |= (F[diag[i]] == 0.0);
    return __any(bad);
}

template <int NR, int SM>
__global__ __launch_bounds__(512) void k_fast(DevPattern P, FastParams fp, int batch,
                                              const double *__restrict__ gA, const double *__restrict__ gb,
                                              double *__restrict__ gx, const double *__restrict__ gF,
                                              SolveParams sp, float *__restrict__ g_iter,
                                              int *__restrict__ g_conv, double *__restrict__ g_ires,
                                              double *__restrict__ g_tres)
{
    extern __shared__ __align__(16) unsigned char smem[];
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(fp.blob);
        uint4 *dst = reinterpret_cast<uint4 *>(smem);
        for (int i = threadIdx.x; i < fp.blob_bytes / 16; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int n = P.n, nnz = P.nnz, S = fp.S;
    double *F = reinterpret_cast<double *>(smem + fp.blob_bytes + (size_t)wave * fp.wave_bytes);
    double *Wb = F + nnz;                       // ParILU only
    double *e0 = Wb + (fp.par_w ? nnz : 0), *e1 = e0 + n, *e2 = e1 + n;
    const uint16_t *ci = reinterpret_cast<const uint16_t *>(smem + fp.o_ci);
    const uint16_t *diag = reinterpret_cast<const uint16_t *>(smem + fp.o_diag);
    const ushort4 *ldesc = reinterpret_cast<const ushort4 *>(smem + fp.o_ldesc);
    const ushort4 *udesc = reinterpret_cast<const ushort4 *>(smem + fp.o_udesc);
    const uint16_t *lptr = reinterpret_cast<const uint16_t *>(smem + fp.o_lptr);
    const uint16_t *uptr = reinterpret_cast<const uint16_t *>(smem + fp.o_uptr);

    uint32_t meta[SM];
#pragma unroll
    for (int s = 0; s < SM; s++) meta[s] = (s < S) ? fp.slot_meta[s * WAVE + lane] : 0xFFFF0000u;
    double a[SM];

    for (;;) {
        int sys = 0;
        if (lane == 0) sys = atomicAdd(fp.counter, 1);
        sys = __builtin_amdgcn_readfirstlane(sys);
        if (sys >= batch) break;

        const double *Ap = gA + (size_t)sys * nnz;
        const double *bp = gb + (size_t)sys * n;
        Vec<NR> b, x, r, rh, p, v, s, t, ph, sh;
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const int i = lane + WAVE * k;
            b.e[k] = (i < n) ? bp[i] : 0.0;
        }
        int bad = 0;
        if (sp.factor == FK_LOAD) {
            const double *Fp = gF + (size_t)sys * nnz;
            for (int e = lane; e < nnz; e += WAVE) F[e] = Fp[e];
#pragma unroll
            for (int q = 0; q < SM; q++) {
                a[q] = 0.0;
                const int e = (q < S) ? fp.slot_pos[q * WAVE + lane] : -1;
                if (e >= 0) a[q] = Ap[e];
            }
            wave_sync();
        } else {
            for (int e = lane; e < nnz; e += WAVE) F[e] = Ap[e];
            wave_sync();
#pragma unroll
            for (int q = 0; q < SM; q++) {
                a[q] = 0.0;
                const int e = (q < S) ? fp.slot_pos[q * WAVE + lane] : -1;
                if (e >= 0) a[q] = F[e];
            }
            if (sp.factor == FK_ILU) {
                bad = ilu_fast(ldesc, lptr, fp.nlev_l, ci, diag,
                               reinterpret_cast<const uint16_t *>(smem + fp.o_upd_ptr),
                               reinterpret_cast<const ushort2 *>(smem + fp.o_upd), n, F, lane);
            } else if (sp.factor == FK_PARILU) {
                // Jacobi sweeps (parILU_0_gpu.cu:18-57 formula): Wb holds the
                // previous sweep, A comes from the slot registers
                const uint16_t *dptr = reinterpret_cast<const uint16_t *>(smem + fp.o_dep_ptr);
                const ushort2 *dep = reinterpret_cast<const ushort2 *>(smem + fp.o_dep);
                const uint16_t *pdiv = reinterpret_cast<const uint16_t *>(smem + fp.o_pdiv);
                for (int sw = 0; sw < sp.sweeps; sw++) {
                    for (int e = lane; e < nnz; e += WAVE) Wb[e] = F[e];
                    wave_sync();
#pragma unroll
                    for (int q = 0; q < SM; q++) {
                        const int e = (q < S) ? fp.slot_pos[q * WAVE + lane] : -1;
                        if (e >= 0) {
                            double sum = 0.0;
                            const int de = dptr[e + 1];
                            for (int dd = dptr[e]; dd < de; dd++) {
                                const ushort2 pq = dep[dd];
                                sum = fma(Wb[pq.x], Wb[pq.y], sum);
                            }
                            const int pv = pdiv[e];
                            if (pv != 0xFFFF) F[e] = (a[q] - sum) / Wb[pv];
                            else F[e] = a[q] - sum;
                        }
                    }
                    wave_sync();
                }
            }
        }

        float iter = 0.0f;
        int conv = BILU_NOT_CONVERGED;
#pragma unroll
        for (int k = 0; k < NR; k++) {
            x.e[k] = 0.0; v.e[k] = 0.0; p.e[k] = 0.0; r.e[k] = b.e[k]; rh.e[k] = b.e[k];
            ph.e[k] = 0.0; sh.e[k] = 0.0; s.e[k] = 0.0; t.e[k] = 0.0;
        }
        const double bnorm = sqrt(vdot(b, b, lane, n));
        const double res0 = bnorm;
        const double tol_abs = (sp.tol_mode == BILU_TOL_RELATIVE) ? sp.tol * bnorm : sp.tol;
        double iter_res = res0;

        auto apply = [&](const Vec<NR> &in, Vec<NR> &out) {
            to_lds(in, e0, lane, n);
            wave_sync();
            lsolve_fast(ldesc, lptr, fp.nlev_l, ci, F, e0, e1, lane);
            usolve_fast(udesc, uptr, fp.nlev_u, ci, F, e1, e2, lane);
            from_lds(out, e2, lane, n);
        };
        auto matvec = [&](Vec<NR> &out) {  // out = A e2
            spmv_slots<SM>(a, meta, S, e2, e1);
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
                const bool half_conv = iter_res < tol_abs;
                if (half_conv) {
#pragma unroll
                    for (int k = 0; k < NR; k++) x.e[k] = fma(alpha, ph.e[k], x.e[k]);
                } else {
                    apply(s, sh);
                    matvec(t);
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
                if (half_conv || iter_res < tol_abs) {
                    conv = BILU_CONVERGED;
                    if (!sp.restart || restarts >= sp.max_restarts) break;
                    to_lds(x, e2, lane, n);
                    wave_sync();
                    Vec<NR> w;
                    matvec(w);
#pragma unroll
                    for (int k = 0; k < NR; k++) w.e[k] = b.e[k] - w.e[k];
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
        to_lds(x, e2, lane, n);
        wave_sync();
        Vec<NR> w;
        matvec(w);
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
        wave_sync();
    }
}
