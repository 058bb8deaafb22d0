// mlp.hip -- the learned warm start of ParILU(0): an MLP maps the scaled
// values of A (one system = one row of features on the pattern) to the
// values of its in-place ILU(0) factor, with BF16 MFMA GEMMs (FP32
// accumulation) for training and inference on gfx950.
//
//   x_e = asinh(a_e / SIGMA) / YS           (features, BF16)
//   h1 = gelu(x W1^T + b1), h2 = gelu(h1 W2^T + b2)   (hidden 512, BF16)
//   y = h2 W3^T + b3,  F_guess_e = SIGMA sinh(YS (base_e + y_e))  (FP64 out)
// with the base guess base_e = asinh(s_e / SIGMA) / YS, s_e = a_ij for
// (i, j) on or above the diagonal and a_ij / a_jj below it (the first
// Jacobi guess of the factor): the network learns the correction, which
// carries far less information than the factor itself.  Trained on exact
// ILU(0) factors F with the mean squared error of y against t_e - base_e,
// t_e = asinh(F_e / SIGMA) / YS (sign-preserving log scale: errors in y are
// relative errors in F).  Weights are kept in FP32 (master, Adam moments)
// and copied to BF16 for the GEMMs.
//
// GEMMs: the forward products (both operands k-contiguous) run k_gemm_nt,
// the backward ones k_gemm.  k_gemm: C[m][n] = sum_k A(m, k) B(k, n),
// A(m, k) = TA ? A[k lda + m] : A[m lda + k], B(k, n) = TB ? B[n ldb + k] :
// B[k ldb + n], all dimensions
// multiples of the tile (buffers are padded).  256 threads = 4 waves, tile
// 128 x 128 x 64: each wave owns a 64 x 64 quarter as 4 x 4 MFMA
// 16x16x32 blocks; the next k tile's global loads are issued before the
// MFMAs of the current one.  Operands are staged in LDS as As[m][k], Bs[n][k] (k
// contiguous, so each lane's MFMA fragment -- 8 consecutive k -- is one
// 16-byte read); a transposed global layout is transposed on the way in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "../../include/bilu.h"
#include "mlp_scaling.hpp"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int TM = 128, TN = 128, TK = 64, KP = TK + 8;   // LDS row pitch (bf16), padded
using bilu_mlp::SIGMA;
using bilu_mlp::YS;
using bilu_mlp::INV_SIGMA_F;
using bilu_mlp::INV_YS;

enum Epi { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_GELU_BWD = 3, EPI_BIAS_F16 = 4, EPI_BIAS_GELU_H = 5 };

__device__ __forceinline__ float gelu(float z)
{
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    return 0.5f * z * (1.0f + tanhf(u));
}

__device__ __forceinline__ float gelu_grad(float z)
{
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    const float t = tanhf(u);
    const float du = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * z * z);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
}

// EPI_F32: C = acc; EPI_BIAS: C = acc + bias[n]; EPI_BIAS_GELU: Z = acc +
// bias[n] (FP32, kept for the backward pass), H = gelu(Z) (BF16);
// EPI_GELU_BWD: H = acc * gelu'(Z[m][n]) (BF16); EPI_BIAS_F16: C (as FP16,
// pitch ldc) = acc + bias[n] rounded to FP16; EPI_BIAS_GELU_H: H only
// (inference keeps no Z)
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const __bf16 *__restrict__ A, int lda, const __bf16 *__restrict__ B,
                                              int ldb, int K, float *__restrict__ C, int ldc,
                                              const float *__restrict__ bias, float *__restrict__ Z,
                                              __bf16 *__restrict__ H)
{
    __shared__ __align__(16) __bf16 As[TM * KP];
    __shared__ __align__(16) __bf16 Bs[TN * KP];
    const int t = threadIdx.x, lane = t % 64, wave = t / 64;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
    floatx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // a tile of one operand is 128 x TK bf16 = TK * 16 vectors of 8, VPT per thread
    constexpr int VPT = TK * 16 / 256;
    constexpr int KV = TK / 8;   // vectors along k in a row
    auto gload = [&](const __bf16 *P, int ld, bool trans, int r0, int k0, bf16x8 (&g)[VPT]) {
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const int id = t + 256 * v;
            if (!trans) {   // rows r (128), KV vectors of 8 k each
                const int r = id / KV, c = (id % KV) * 8;
                g[v] = *reinterpret_cast<const bf16x8 *>(&P[(size_t)(r0 + r) * ld + k0 + c]);
            } else {        // rows k (TK), 16 vectors of 8 r each
                const int kk = id / 16, c = (id % 16) * 8;
                g[v] = *reinterpret_cast<const bf16x8 *>(&P[(size_t)(k0 + kk) * ld + r0 + c]);
            }
        }
    };
    auto sstore = [&](__bf16 *S, bool trans, const bf16x8 (&g)[VPT]) {
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const int id = t + 256 * v;
            if (!trans) {
                const int r = id / KV, c = (id % KV) * 8;
                *reinterpret_cast<bf16x8 *>(&S[r * KP + c]) = g[v];
            } else {        // transpose into S[r][k]
                const int kk = id / 16, c = (id % 16) * 8;
#pragma unroll
                for (int j = 0; j < 8; j++) S[(c + j) * KP + kk] = g[v][j];
            }
        }
    };
    bf16x8 ga[VPT], gb[VPT];
    gload(A, lda, TA, m0, 0, ga);
    gload(B, ldb, !TB, n0, 0, gb);
    for (int k0 = 0; k0 < K; k0 += TK) {
        sstore(As, TA, ga);
        sstore(Bs, !TB, gb);
        __syncthreads();
        // the next tile's global loads are in flight while this one is multiplied
        if (k0 + TK < K) {
            gload(A, lda, TA, m0, k0 + TK, ga);
            gload(B, ldb, !TB, n0, k0 + TK, gb);
        }
#pragma unroll
        for (int kq = 0; kq < TK; kq += 32) {
            bf16x8 af[4], bfr[4];
#pragma unroll
            for (int i = 0; i < 4; i++)
                af[i] = *reinterpret_cast<const bf16x8 *>(&As[(wm + 16 * i + lane % 16) * KP + kq + 8 * (lane / 16)]);
#pragma unroll
            for (int j = 0; j < 4; j++)
                bfr[j] = *reinterpret_cast<const bf16x8 *>(&Bs[(wn + 16 * j + lane % 16) * KP + kq + 8 * (lane / 16)]);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // lane holds D[m = 4 (lane / 16) + r][n = lane % 16] of each 16 x 16 block
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + 4 * (lane / 16) + r;
                const int n = n0 + wn + 16 * j + lane % 16;
                const float v = acc[i][j][r];
                const size_t o = (size_t)m * ldc + n;
                if (EPI == EPI_F32) {
                    C[o] = v;
                } else if (EPI == EPI_BIAS) {
                    C[o] = v + bias[n];
                } else if (EPI == EPI_BIAS_F16) {
                    reinterpret_cast<_Float16 *>(C)[o] = (_Float16)(v + bias[n]);
                } else if (EPI == EPI_BIAS_GELU) {
                    const float z = v + bias[n];
                    Z[o] = z;
                    H[o] = (__bf16)gelu(z);
                } else {
                    H[o] = (__bf16)(v * gelu_grad(Z[o]));
                }
            }
}

// Forward GEMM C = A B^T (+ epilogue) with both operands k-contiguous: A
// [M][K] (lda), B [N][K] (ldb); M, N, K multiples of FM, FN, FK.  256
// threads = 4 waves, block tile FM x FN = 256 x 128: wave w owns rows
// 64 w .. 64 w + 63 and all 128 columns as 4 x 8 MFMA 16x16x32 blocks, so
// each LDS fragment of B feeds 4 MFMAs and each register fragment of A 8.
// A's fragments belong to one wave only: they are loaded straight from
// global memory into registers in the MFMA layout (one 16-byte load per
// fragment), two 32-wide k steps ahead.  B's 128 x 32 step tile is shared
// by the four waves and staged in LDS, double buffered (one barrier per k
// step), its global loads also issued two steps ahead.  Measured at 65536
// rows the kernel spends most of its time waiting on those loads (about 27%
// of the BF16 MFMA peak; a 256 x 256 block tile with 8 waves was no faster).
// Every accumulator receives the same MFMAs in the same k order as in
// k_gemm, so both kernels give the same bits.  Blocks are renumbered so that
// consecutive tiles (the column tiles of one row tile) run on one XCD and
// share its L2 (dispatch places block b on XCD b mod 8).
constexpr int FM = 256, FN = 128, FK = 32, FKP = FK + 8;
template <int EPI>
__global__ __launch_bounds__(256, 2) void k_gemm_nt(const __bf16 *__restrict__ A, int lda,
                                                    const __bf16 *__restrict__ B, int ldb, int K, int ntn,
                                                    float *__restrict__ C, int ldc, const float *__restrict__ bias,
                                                    float *__restrict__ Z, __bf16 *__restrict__ H)
{
    __shared__ __align__(16) __bf16 Bs[2][FN * FKP];
    const int t = threadIdx.x, lane = t % 64, wave = t / 64;
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int lg = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    const int m0 = (lg / ntn) * FM + wave * 64, n0 = (lg % ntn) * FN;
    floatx4 acc[4][8];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    // lane's A fragment i of k step s: row m0 + 16 i + lane % 16, k = 32 s + 8 (lane / 16) .. + 7
    const __bf16 *Aw = A + (size_t)(m0 + lane % 16) * lda + 8 * (lane / 16);
    constexpr int KV = FK / 8, VB = FN * KV / 256;   // B step tile: FN rows x KV vectors, VB per thread
    bf16x8 fa[3][4], gb[2][VB];
    auto aload = [&](int s, bf16x8 (&f)[4]) {
#pragma unroll
        for (int i = 0; i < 4; i++) f[i] = *reinterpret_cast<const bf16x8 *>(Aw + (size_t)16 * i * lda + FK * s);
    };
    auto bload = [&](int s, bf16x8 (&g)[VB]) {
#pragma unroll
        for (int v = 0; v < VB; v++) {
            const int id = t + 256 * v, r = id / KV, c = (id % KV) * 8;
            g[v] = *reinterpret_cast<const bf16x8 *>(&B[(size_t)(n0 + r) * ldb + FK * s + c]);
        }
    };
    auto bstore = [&](const bf16x8 (&g)[VB], __bf16 *S) {
#pragma unroll
        for (int v = 0; v < VB; v++) {
            const int id = t + 256 * v, r = id / KV, c = (id % KV) * 8;
            *reinterpret_cast<bf16x8 *>(&S[r * FKP + c]) = g[v];
        }
    };
    // k step s: A(s + 2) and B(s + 2) loads issued (two steps of MFMAs hide
    // their latency), the MFMAs of step s, B(s + 1) (loaded in step s - 1)
    // stored to the LDS buffer that step s - 1 read, one barrier.  The loads
    // and stores are unconditional (steps past the end reload step S - 1 and
    // store into the idle buffer), so every step issues the same memory
    // operations and the compiler's wait counts stay exact across steps.
    const int S = K / FK;
    auto step = [&](auto R, int s) {
        constexpr int r = decltype(R)::value;
        const int sn = min(s + 2, S - 1);
        aload(sn, fa[(r + 2) % 3]);
        bload(sn, gb[r % 2]);
        __builtin_amdgcn_sched_barrier(0);   // the loads stay ahead of the MFMAs
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const bf16x8 bf =
                *reinterpret_cast<const bf16x8 *>(&Bs[r % 2][(16 * j + lane % 16) * FKP + 8 * (lane / 16)]);
#pragma unroll
            for (int i = 0; i < 4; i++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[r % 3][i], bf, acc[i][j], 0, 0, 0);
        }
        bstore(gb[(r + 1) % 2], Bs[(r + 1) % 2]);
        __syncthreads();
    };
    aload(0, fa[0]);
    bload(0, gb[0]);
    aload(min(1, S - 1), fa[1]);
    bload(min(1, S - 1), gb[1]);
    bstore(gb[0], Bs[0]);
    __syncthreads();
    int s0 = 0;
    for (; s0 + 6 <= S; s0 += 6) {
        step(std::integral_constant<int, 0>(), s0);
        step(std::integral_constant<int, 1>(), s0 + 1);
        step(std::integral_constant<int, 2>(), s0 + 2);
        step(std::integral_constant<int, 3>(), s0 + 3);
        step(std::integral_constant<int, 4>(), s0 + 4);
        step(std::integral_constant<int, 5>(), s0 + 5);
    }
    if (s0 < S) step(std::integral_constant<int, 0>(), s0);
    if (s0 + 1 < S) step(std::integral_constant<int, 1>(), s0 + 1);
    if (s0 + 2 < S) step(std::integral_constant<int, 2>(), s0 + 2);
    if (s0 + 3 < S) step(std::integral_constant<int, 3>(), s0 + 3);
    if (s0 + 4 < S) step(std::integral_constant<int, 4>(), s0 + 4);
    if constexpr (EPI == EPI_BIAS_F16 || EPI == EPI_BIAS_GELU_H) {
        // 16-bit outputs leave through the LDS: in the MFMA layout a lane
        // holds four rows of one column, so direct stores write 32-byte
        // pieces of rows.  Each wave converts its 64 x 128 tile in two
        // 64 x 64 halves into its own LDS region, then writes whole 128-byte
        // row pieces, 16 bytes per lane.  Values and rounding as below.
        using T16 = std::conditional_t<EPI == EPI_BIAS_F16, _Float16, __bf16>;
        constexpr int CP = 64 + 8;   // row pitch (elements): 144 bytes
        __shared__ __align__(16) T16 Cs[4][64 * CP];
        T16 *cw = Cs[wave];
        T16 *out = EPI == EPI_BIAS_F16 ? reinterpret_cast<T16 *>(C) : reinterpret_cast<T16 *>(H);
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    const float bn = bias[n0 + 64 * h + 16 * jj + lane % 16];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const float z = acc[i][4 * h + jj][r] + bn;
                        T16 y;
                        if constexpr (EPI == EPI_BIAS_F16) y = (T16)z;
                        else y = (T16)gelu(z);
                        cw[(16 * i + 4 * (lane / 16) + r) * CP + 16 * jj + lane % 16] = y;
                    }
                }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int row = 8 * q + lane / 8, c = 8 * (lane % 8);
                const bf16x8 v = *reinterpret_cast<const bf16x8 *>(&cw[row * CP + c]);
                *reinterpret_cast<bf16x8 *>(&out[(size_t)(m0 + row) * ldc + n0 + 64 * h + c]) = v;
            }
            __syncthreads();
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 8; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + 16 * i + 4 * (lane / 16) + r;
                const int n = n0 + 16 * j + lane % 16;
                const float v = acc[i][j][r];
                const size_t o = (size_t)m * ldc + n;
                if (EPI == EPI_BIAS) {
                    C[o] = v + bias[n];
                } else if (EPI == EPI_BIAS_F16) {
                    reinterpret_cast<_Float16 *>(C)[o] = (_Float16)(v + bias[n]);
                } else if (EPI == EPI_BIAS_GELU) {
                    const float z = v + bias[n];
                    Z[o] = z;
                    H[o] = (__bf16)gelu(z);
                } else if (EPI == EPI_BIAS_GELU_H) {
                    H[o] = (__bf16)gelu(v + bias[n]);
                } else {
                    C[o] = v;
                }
            }
}

// features: X[s][e] = asinh(a / SIGMA) / YS (BF16) in FP32 arithmetic
// (bilu_mlp::feature); padding
// columns and rows s >= batch are 0.  One block per system row s (its reads
// and writes stay in one XCD's L2).
__global__ __launch_bounds__(256) void k_features(const double *__restrict__ A, int nnz, int P0, int batch,
                                                 __bf16 *__restrict__ X)
{
    constexpr int U = 8;   // loads in flight per thread
    const int s = blockIdx.x;
    const double *As = A + (size_t)s * nnz;
    __bf16 *Xs = X + (size_t)s * P0;
    const int lim = s < batch ? nnz : 0;
    for (int e0 = threadIdx.x; e0 < P0; e0 += U * 256) {
        double a[U];
#pragma unroll
        for (int j = 0; j < U; j++) a[j] = (e0 + j * 256 < lim) ? As[e0 + j * 256] : 0.0;
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int e = e0 + j * 256;
            if (e < P0) Xs[e] = (__bf16)(e < lim ? bilu_mlp::feature((float)a[j]) : 0.0f);
        }
    }
}

// base guess of entry e of system s (piv[e] = position of a_jj for a lower
// entry (i, j), -1 otherwise), in the scaled space (FP32: the warm start is
// a guess; scaled values carry relative accuracy ~1e-7)
__device__ __forceinline__ float base_guess(const double *__restrict__ As, const int *__restrict__ piv, int e)
{
    const int pv = piv[e];
    return bilu_mlp::base_scaled((float)As[e], pv >= 0 ? (float)As[pv] : 1.0f, pv >= 0);
}

// loss gradient dY = 2 (y - (t - base)) / (batch nnz) (BF16), t from the exact
// factors; per-block partial sums of the squared error into part[]
__global__ void k_loss_grad(const float *__restrict__ Y, const double *__restrict__ F, const double *__restrict__ A,
                            const int *__restrict__ piv, int nnz, int P3,
                            int batch, __bf16 *__restrict__ dY, float *__restrict__ part)
{
    __shared__ float red[256];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float sq = 0.0f;
    if (i < (size_t)batch * P3) {
        const int s = int(i / P3), e = int(i % P3);
        float g = 0.0f;
        if (e < nnz) {
            const float tv = asinhf((float)(F[(size_t)s * nnz + e] / SIGMA)) / YS -
                             base_guess(A + (size_t)s * nnz, piv, e);
            const float d = Y[i] - tv;
            sq = d * d;
            g = 2.0f * d / ((float)batch * (float)nnz);
        }
        dY[i] = (__bf16)g;
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int o = 128; o > 0; o /= 2) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_loss_final(const float *__restrict__ part, int nparts, float scale, float *__restrict__ loss)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < nparts; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o /= 2) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] * scale;
}

// bias gradient g[n] = sum_m D[m][n] over BF16 rows, in two passes: block
// (x, y) sums rows y, y + CS, ... of 256 columns into part[y][n], then the CS
// partial sums of each column are added in order (deterministic)
constexpr int CS = 32;
__global__ void k_colsum_part(const __bf16 *__restrict__ D, int M, int N, float *__restrict__ part)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int m = blockIdx.y; m < M; m += CS) s += (float)D[(size_t)m * N + n];
    part[(size_t)blockIdx.y * N + n] = s;
}

__global__ void k_colsum_final(const float *__restrict__ part, int N, float *__restrict__ g)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int y = 0; y < CS; y++) s += part[(size_t)y * N + n];
    g[n] = s;
}

// output: F_guess = SIGMA sinh(YS (base + y)) (FP64, unpadded).  One block
// per system: its A row (FP32, as base_guess rounds it) is staged in LDS, so
// the pivot reads of the base guess stay on chip.
__global__ __launch_bounds__(256) void k_outputs(const float *__restrict__ Y, const double *__restrict__ A,
                                                const int *__restrict__ piv, int nnz, int P3, int batch,
                                                double *__restrict__ Fo)
{
    constexpr int U = 8;   // loads in flight per thread
    extern __shared__ float sa[];
    const int s = blockIdx.x;
    if (s >= batch) return;
    const double *As = A + (size_t)s * nnz;
    for (int e0 = threadIdx.x; e0 < nnz; e0 += U * 256) {
        double a[U];
#pragma unroll
        for (int j = 0; j < U; j++) a[j] = (e0 + j * 256 < nnz) ? As[e0 + j * 256] : 0.0;
#pragma unroll
        for (int j = 0; j < U; j++)
            if (e0 + j * 256 < nnz) sa[e0 + j * 256] = (float)a[j];
    }
    __syncthreads();
    const float *Ys = Y + (size_t)s * P3;   // y is used rounded to FP16, as by the fused warm start
    double *Fs = Fo + (size_t)s * nnz;
    for (int e0 = threadIdx.x; e0 < nnz; e0 += U * 256) {
        float y[U];
        int pv[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const bool in = e0 + j * 256 < nnz;
            y[j] = in ? Ys[e0 + j * 256] : 0.0f;
            pv[j] = in ? piv[e0 + j * 256] : -1;
        }
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int e = e0 + j * 256;
            if (e < nnz) {
                const bool lower = pv[j] >= 0;
                Fs[e] = bilu_mlp::guess_from(sa[e], lower ? sa[pv[j]] : 1.0f, lower, (float)(_Float16)y[j]);
            }
        }
    }
}

// Adam step on FP32 master weights w (moments m, v), BF16 copy wb
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, __bf16 *__restrict__ wb, size_t n, float lr, float b1t, float b2t)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - b1t), vh = vi / (1.0f - b2t);
    const float wi = w[i] - lr * mh / (sqrtf(vh) + eps);
    w[i] = wi;
    if (wb) wb[i] = (__bf16)wi;
}

__global__ void k_to_bf16(const float *__restrict__ w, __bf16 *__restrict__ wb, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = (__bf16)w[i];
}

// [rows][cols] FP32 block copy between a packed and a padded (pitch) layout
__global__ void k_copy2d(const float *__restrict__ src, int sp, float *__restrict__ dst, int dp, int rows, int cols)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    const int r = int(i / cols), c = int(i % cols);
    dst[(size_t)r * dp + c] = src[(size_t)r * sp + c];
}

int pad_to(int x, int q) { return (x + q - 1) / q * q; }

unsigned blocks(size_t n, unsigned b = 256) { return unsigned((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------- handle
struct bilu_mlp_s {
    int nnz = 0, P0 = 0, H = 0, P3 = 0;
    int device = 0;
    long step = 0;
    // parameters, FP32, one allocation: W1[H][P0], b1[H], W2[H][H], b2[H], W3[P3][H], b3[P3]
    size_t off[7] = {};
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    __bf16 *wb = nullptr;
    // activations for a minibatch of up to cap rows
    int cap = 0;
    __bf16 *X = nullptr, *H1 = nullptr, *H2 = nullptr, *dY = nullptr, *dZ2 = nullptr, *dZ1 = nullptr;
    float *Z1 = nullptr, *Z2 = nullptr, *Y = nullptr, *part = nullptr, *cpart = nullptr;
    int *piv = nullptr;   // [nnz]: position of a_jj for lower entries (i, j), else -1
    const void *pattern = nullptr;   // the pattern handle the network was created on
};

namespace {

#define MLP_TRY(x)                                  \
    do {                                            \
        if ((x) != hipSuccess) return BILU_ERR_HIP; \
    } while (0)

size_t nparams(const bilu_mlp_s &q) { return q.off[6]; }

void free_acts(bilu_mlp_s &q)
{
    for (void *p : {(void *)q.X, (void *)q.H1, (void *)q.H2, (void *)q.dY, (void *)q.dZ2, (void *)q.dZ1,
                    (void *)q.Z1, (void *)q.Z2, (void *)q.Y, (void *)q.part, (void *)q.cpart})
        if (p) (void)hipFree(p);
    q.X = q.H1 = q.H2 = q.dY = q.dZ2 = q.dZ1 = nullptr;
    q.Z1 = q.Z2 = q.Y = q.part = q.cpart = nullptr;
    q.cap = 0;
}

// activations for `rows` (a multiple of the tile) rows
int reserve(bilu_mlp_s &q, int rows)
{
    if (rows <= q.cap) return BILU_OK;
    MLP_TRY(hipDeviceSynchronize());
    free_acts(q);
    const size_t r = (size_t)rows;
    MLP_TRY(hipMalloc((void **)&q.X, r * q.P0 * 2));
    MLP_TRY(hipMalloc((void **)&q.H1, r * q.H * 2));
    MLP_TRY(hipMalloc((void **)&q.H2, r * q.H * 2));
    MLP_TRY(hipMalloc((void **)&q.dY, r * q.P3 * 2));
    MLP_TRY(hipMalloc((void **)&q.dZ2, r * q.H * 2));
    MLP_TRY(hipMalloc((void **)&q.dZ1, r * q.H * 2));
    MLP_TRY(hipMalloc((void **)&q.Z1, r * q.H * 4));
    MLP_TRY(hipMalloc((void **)&q.Z2, r * q.H * 4));
    MLP_TRY(hipMalloc((void **)&q.Y, r * q.P3 * 4));
    MLP_TRY(hipMalloc((void **)&q.part, (r * q.P3 / 256 + 1) * 4));
    MLP_TRY(hipMalloc((void **)&q.cpart, (size_t)CS * std::max(q.P3, q.H) * 4));
    q.cap = rows;
    return BILU_OK;
}

template <bool TA, bool TB, int EPI>
void gemm(hipStream_t st, int M, int N, int K, const __bf16 *A, int lda, const __bf16 *B, int ldb, float *C,
          int ldc, const float *bias = nullptr, float *Z = nullptr, __bf16 *H = nullptr)
{
    if constexpr (!TA && TB) {
        // the 16-bit epilogues store 8 outputs per lane: rows 16-byte aligned
        constexpr bool w16 = EPI == EPI_BIAS_F16 || EPI == EPI_BIAS_GELU_H;
        if (M % FM == 0 && N % FN == 0 && K % FK == 0 && (!w16 || ldc % 8 == 0)) {
            hipLaunchKernelGGL((k_gemm_nt<EPI>), dim3((M / FM) * (N / FN)), dim3(256), 0, st, A, lda, B, ldb, K,
                               N / FN, C, ldc, bias, Z, H);
            return;
        }
    }
    static_assert((EPI != EPI_BIAS_GELU_H && EPI != EPI_BIAS_F16) || (!TA && TB), "forward epilogues");
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), dim3(N / TN, M / TM), dim3(256), 0, st, A, lda, B, ldb, K, C, ldc,
                       bias, Z, H);
}

void colsum(hipStream_t st, const __bf16 *D, int M, int N, float *part, float *g)
{
    hipLaunchKernelGGL(k_colsum_part, dim3(blocks(N), CS), dim3(256), 0, st, D, M, N, part);
    hipLaunchKernelGGL(k_colsum_final, dim3(blocks(N)), dim3(256), 0, st, part, N, g);
}

// forward of `rows` padded rows (features already in X): Y (FP32); `keep`:
// the pre-activations Z1, Z2 for the backward pass
void forward(bilu_mlp_s &q, int rows, hipStream_t st, bool keep = true)
{
    const __bf16 *W1 = q.wb + q.off[0], *W2 = q.wb + q.off[2], *W3 = q.wb + q.off[4];
    const float *b1 = q.w + q.off[1], *b2 = q.w + q.off[3], *b3 = q.w + q.off[5];
    if (keep) {
        gemm<false, true, EPI_BIAS_GELU>(st, rows, q.H, q.P0, q.X, q.P0, W1, q.P0, nullptr, q.H, b1, q.Z1, q.H1);
        gemm<false, true, EPI_BIAS_GELU>(st, rows, q.H, q.H, q.H1, q.H, W2, q.H, nullptr, q.H, b2, q.Z2, q.H2);
    } else {
        gemm<false, true, EPI_BIAS_GELU_H>(st, rows, q.H, q.P0, q.X, q.P0, W1, q.P0, nullptr, q.H, b1, nullptr, q.H1);
        gemm<false, true, EPI_BIAS_GELU_H>(st, rows, q.H, q.H, q.H1, q.H, W2, q.H, nullptr, q.H, b2, nullptr, q.H2);
    }
    gemm<false, true, EPI_BIAS>(st, rows, q.P3, q.H, q.H2, q.H, W3, q.H, q.Y, q.P3, b3);
}

// loss and gradient into q.g for `batch` systems (rows padded with zero features
// and zero loss gradient)
int loss_and_grad(bilu_mlp_s &q, int batch, const double *A, const double *F, float *d_loss, hipStream_t st)
{
    const int rows = pad_to(batch, FM);
    int rc = reserve(q, rows);
    if (rc != BILU_OK) return rc;
    hipLaunchKernelGGL(k_features, dim3(rows), dim3(256), 0, st, A, q.nnz, q.P0, batch, q.X);
    forward(q, rows, st);
    MLP_TRY(hipMemsetAsync(q.dY, 0, (size_t)rows * q.P3 * 2, st));
    const unsigned nb = blocks((size_t)batch * q.P3);
    hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, st, q.Y, F, A, q.piv, q.nnz, q.P3, batch, q.dY, q.part);
    if (d_loss)
        hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, st, q.part, (int)nb,
                           1.0f / ((float)batch * (float)q.nnz), d_loss);
    float *gW1 = q.g + q.off[0], *gb1 = q.g + q.off[1], *gW2 = q.g + q.off[2], *gb2 = q.g + q.off[3];
    float *gW3 = q.g + q.off[4], *gb3 = q.g + q.off[5];
    const __bf16 *W2 = q.wb + q.off[2], *W3 = q.wb + q.off[4];
    // layer 3: gW3[n][k] = sum_m dY[m][n] H2[m][k]; gb3 = colsum dY; dZ2 = (dY W3) * gelu'(Z2)
    gemm<true, false, EPI_F32>(st, q.P3, q.H, rows, q.dY, q.P3, q.H2, q.H, gW3, q.H);
    colsum(st, q.dY, rows, q.P3, q.cpart, gb3);
    gemm<false, false, EPI_GELU_BWD>(st, rows, q.H, q.P3, q.dY, q.P3, W3, q.H, nullptr, q.H, nullptr, q.Z2, q.dZ2);
    // layer 2
    gemm<true, false, EPI_F32>(st, q.H, q.H, rows, q.dZ2, q.H, q.H1, q.H, gW2, q.H);
    colsum(st, q.dZ2, rows, q.H, q.cpart, gb2);
    gemm<false, false, EPI_GELU_BWD>(st, rows, q.H, q.H, q.dZ2, q.H, W2, q.H, nullptr, q.H, nullptr, q.Z1, q.dZ1);
    // layer 1
    gemm<true, false, EPI_F32>(st, q.H, q.P0, rows, q.dZ1, q.H, q.X, q.P0, gW1, q.P0);
    colsum(st, q.dZ1, rows, q.H, q.cpart, gb1);
    MLP_TRY(hipGetLastError());
    return BILU_OK;
}

// packed (exported) layout: W1[H][nnz], b1[H], W2[H][H], b2[H], W3[nnz][H], b3[nnz]
size_t packed_count(const bilu_mlp_s &q)
{
    return (size_t)q.H * q.nnz + q.H + (size_t)q.H * q.H + q.H + (size_t)q.nnz * q.H + q.nnz;
}

// copy between the packed layout and the padded internal one
void copy_params(const bilu_mlp_s &q, const float *src, float *dst, bool to_packed, hipStream_t st)
{
    struct Part { size_t ioff, poff; int rows, cols, ipitch; };
    const size_t p1 = 0, p2 = p1 + (size_t)q.H * q.nnz, p3 = p2 + q.H, p4 = p3 + (size_t)q.H * q.H, p5 = p4 + q.H,
                 p6 = p5 + (size_t)q.nnz * q.H;
    const Part parts[6] = {{q.off[0], p1, q.H, q.nnz, q.P0}, {q.off[1], p2, 1, q.H, q.H},
                           {q.off[2], p3, q.H, q.H, q.H},    {q.off[3], p4, 1, q.H, q.H},
                           {q.off[4], p5, q.nnz, q.H, q.H},  {q.off[5], p6, 1, q.nnz, q.P3}};
    for (const auto &pt : parts) {
        const size_t n = (size_t)pt.rows * pt.cols;
        if (to_packed)
            hipLaunchKernelGGL(k_copy2d, dim3(blocks(n)), dim3(256), 0, st, src + pt.ioff, pt.ipitch,
                               dst + pt.poff, pt.cols, pt.rows, pt.cols);
        else
            hipLaunchKernelGGL(k_copy2d, dim3(blocks(n)), dim3(256), 0, st, src + pt.poff, pt.cols,
                               dst + pt.ioff, pt.ipitch, pt.rows, pt.cols);
    }
}

}  // namespace

int bilu_mlp_forward_rows() { return 65536; }

int bilu_mlp_forward_raw(bilu_mlp_s *q, const void *pattern, int batch, const double *d_A, hipStream_t st,
                         const _Float16 **Y, int *P3)
{
    if (!q || q->pattern != pattern || batch <= 0 || batch > bilu_mlp_forward_rows() || !d_A)
        return BILU_ERR_INVALID_ARG;
    const int rows = pad_to(batch, FM);
    const int rc = reserve(*q, rows);
    if (rc != BILU_OK) return rc;
    hipLaunchKernelGGL(k_features, dim3(rows), dim3(256), 0, st, d_A, q->nnz, q->P0, batch, q->X);
    // the three layers, the last one writing the outputs in FP16
    const __bf16 *W1 = q->wb + q->off[0], *W2 = q->wb + q->off[2], *W3 = q->wb + q->off[4];
    const float *b1 = q->w + q->off[1], *b2 = q->w + q->off[3], *b3 = q->w + q->off[5];
    gemm<false, true, EPI_BIAS_GELU_H>(st, rows, q->H, q->P0, q->X, q->P0, W1, q->P0, nullptr, q->H, b1, nullptr, q->H1);
    gemm<false, true, EPI_BIAS_GELU_H>(st, rows, q->H, q->H, q->H1, q->H, W2, q->H, nullptr, q->H, b2, nullptr, q->H2);
    gemm<false, true, EPI_BIAS_F16>(st, rows, q->P3, q->H, q->H2, q->H, W3, q->H, q->Y, q->P3, b3);
    MLP_TRY(hipGetLastError());
    *Y = reinterpret_cast<const _Float16 *>(q->Y);
    *P3 = q->P3;
    return BILU_OK;
}

extern "C" {

int bilu_mlp_create(bilu_pattern_t p, int hidden, unsigned long long seed, bilu_mlp_t *out)
{
    if (!p || !out || hidden <= 0 || hidden % TN) return BILU_ERR_INVALID_ARG;
    bilu_pattern_info info;
    int rc = bilu_pattern_get_info(p, &info);
    if (rc != BILU_OK) return rc;
    auto *q = new bilu_mlp_s();
    (void)hipGetDevice(&q->device);
    // pivot positions of the lower entries (base guess), from the reference L/U split
    std::vector<int> lrp(info.n + 1), lci(info.l_nnz), urp(info.n + 1), uci(info.u_nnz);
    rc = bilu_pattern_lu_split(p, lrp.data(), lci.data(), urp.data(), uci.data());
    if (rc != BILU_OK) { delete q; return rc; }
    std::vector<int> piv(info.nnz, -1), dpos(info.n, -1);
    {
        // pattern position of (i, j): row i's entries are L's strict part then U's
        int e = 0;
        std::vector<int> rowof(info.nnz), colof(info.nnz);
        for (int i = 0; i < info.n; i++) {
            for (int k = lrp[i]; k < lrp[i + 1]; k++)
                if (lci[k] != i) { rowof[e] = i; colof[e] = lci[k]; e++; }
            for (int k = urp[i]; k < urp[i + 1]; k++) {
                if (uci[k] == i) dpos[i] = e;
                rowof[e] = i; colof[e] = uci[k]; e++;
            }
        }
        for (int x = 0; x < info.nnz; x++)
            if (colof[x] < rowof[x]) piv[x] = dpos[colof[x]];
    }
    q->nnz = info.nnz;
    q->pattern = p;
    q->P0 = pad_to(info.nnz, TN);
    q->P3 = pad_to(info.nnz, TN);
    q->H = hidden;
    size_t o = 0;
    const size_t sizes[6] = {(size_t)q->H * q->P0, (size_t)q->H, (size_t)q->H * q->H, (size_t)q->H,
                             (size_t)q->P3 * q->H, (size_t)q->P3};
    for (int i = 0; i < 6; i++) { q->off[i] = o; o += (sizes[i] + 63) / 64 * 64; }
    q->off[6] = o;
    // Glorot-uniform hidden weights, zero biases and a zero output layer (the
    // untrained network returns the base guess); splitmix64 stream from the seed
    std::vector<float> hw(o, 0.0f);
    uint64_t z = seed;
    auto next = [&]() {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t r = z;
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
        r ^= r >> 31;
        return (double)(r >> 11) * (1.0 / 9007199254740992.0);
    };
    const int fan[3][2] = {{q->nnz, q->H}, {q->H, q->H}, {q->H, q->nnz}};
    const int pitch[3] = {q->P0, q->H, q->H}, rowsn[3] = {q->H, q->H, q->nnz};
    for (int l = 0; l < 2; l++) {
        const double lim = std::sqrt(6.0 / (fan[l][0] + fan[l][1]));
        for (int r = 0; r < rowsn[l]; r++)
            for (int c = 0; c < fan[l][0]; c++) hw[q->off[2 * l] + (size_t)r * pitch[l] + c] = (float)((2.0 * next() - 1.0) * lim);
    }
    if (hipMalloc((void **)&q->w, o * 4) != hipSuccess || hipMalloc((void **)&q->g, o * 4) != hipSuccess ||
        hipMalloc((void **)&q->m, o * 4) != hipSuccess || hipMalloc((void **)&q->v, o * 4) != hipSuccess ||
        hipMalloc((void **)&q->wb, o * 2) != hipSuccess) {
        bilu_mlp_destroy(q);
        return BILU_ERR_HIP;
    }
    if (hipMemcpy(q->w, hw.data(), o * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(q->g, 0, o * 4) != hipSuccess || hipMemset(q->m, 0, o * 4) != hipSuccess ||
        hipMemset(q->v, 0, o * 4) != hipSuccess) {
        bilu_mlp_destroy(q);
        return BILU_ERR_HIP;
    }
    if (hipMalloc((void **)&q->piv, (size_t)info.nnz * sizeof(int)) != hipSuccess ||
        hipMemcpy(q->piv, piv.data(), (size_t)info.nnz * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        bilu_mlp_destroy(q);
        return BILU_ERR_HIP;
    }
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(o)), dim3(256), 0, 0, q->w, q->wb, o);
    if (hipDeviceSynchronize() != hipSuccess) {
        bilu_mlp_destroy(q);
        return BILU_ERR_HIP;
    }
    *out = q;
    return BILU_OK;
}

int bilu_mlp_destroy(bilu_mlp_t q)
{
    if (!q) return BILU_ERR_INVALID_ARG;
    free_acts(*q);
    for (void *p : {(void *)q->w, (void *)q->g, (void *)q->m, (void *)q->v, (void *)q->wb, (void *)q->piv})
        if (p) (void)hipFree(p);
    delete q;
    return BILU_OK;
}

int bilu_mlp_num_params(bilu_mlp_t q, long long *count)
{
    if (!q || !count) return BILU_ERR_INVALID_ARG;
    *count = (long long)packed_count(*q);
    return BILU_OK;
}

int bilu_mlp_get_weights(bilu_mlp_t q, float *d_w, void *stream)
{
    if (!q || !d_w) return BILU_ERR_INVALID_ARG;
    copy_params(*q, q->w, d_w, true, (hipStream_t)stream);
    MLP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_mlp_set_weights(bilu_mlp_t q, const float *d_w, void *stream)
{
    if (!q || !d_w) return BILU_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    MLP_TRY(hipMemsetAsync(q->w, 0, nparams(*q) * 4, st));
    copy_params(*q, d_w, q->w, false, st);
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(nparams(*q))), dim3(256), 0, st, q->w, q->wb, nparams(*q));
    MLP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_mlp_gradient(bilu_mlp_t q, int batch, const double *d_A, const double *d_F, float *d_loss,
                      float *d_grad, void *stream)
{
    if (!q || batch <= 0 || !d_A || !d_F) return BILU_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    int rc = loss_and_grad(*q, batch, d_A, d_F, d_loss, st);
    if (rc != BILU_OK) return rc;
    if (d_grad) copy_params(*q, q->g, d_grad, true, st);
    MLP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_mlp_train_step(bilu_mlp_t q, int batch, const double *d_A, const double *d_F, float lr,
                        float *d_loss, void *stream)
{
    if (!q || batch <= 0 || !d_A || !d_F || !(lr > 0.0f)) return BILU_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    int rc = loss_and_grad(*q, batch, d_A, d_F, d_loss, st);
    if (rc != BILU_OK) return rc;
    q->step++;
    const float b1t = std::pow(0.9f, (float)q->step), b2t = std::pow(0.999f, (float)q->step);
    hipLaunchKernelGGL(k_adam, dim3(blocks(nparams(*q))), dim3(256), 0, st, q->w, q->g, q->m, q->v, q->wb,
                       nparams(*q), lr, b1t, b2t);
    MLP_TRY(hipGetLastError());
    return BILU_OK;
}

int bilu_mlp_infer(bilu_mlp_t q, int batch, const double *d_A, double *d_F_init, void *stream)
{
    if (!q || batch < 0 || (batch > 0 && (!d_A || !d_F_init))) return BILU_ERR_INVALID_ARG;
    if (batch == 0) return BILU_OK;
    hipStream_t st = (hipStream_t)stream;
    // in chunks, so the activations stay bounded
    const int chunk = bilu_mlp_forward_rows();
    int rc = reserve(*q, std::min(pad_to(batch, FM), chunk));
    if (rc != BILU_OK) return rc;
    for (int s0 = 0; s0 < batch; s0 += chunk) {
        const int nb = std::min(chunk, batch - s0), rows = pad_to(nb, FM);
        hipLaunchKernelGGL(k_features, dim3(rows), dim3(256), 0, st,
                           d_A + (size_t)s0 * q->nnz, q->nnz, q->P0, nb, q->X);
        forward(*q, rows, st, false);
        hipLaunchKernelGGL(k_outputs, dim3(nb), dim3(256), (size_t)q->nnz * sizeof(float), st, q->Y,
                           d_A + (size_t)s0 * q->nnz, q->piv, q->nnz, q->P3, nb, d_F_init + (size_t)s0 * q->nnz);
    }
    MLP_TRY(hipGetLastError());
    return BILU_OK;
}

}  // extern "C"
