// Scaling maps of the learned ParILU(0) warm start, shared by the network
// kernels (mlp.hip) and the factor kernel's fused warm start, so both compute
// the guess with the same operations:
//   features / targets: x = asinh(v / SIGMA) / YS
//   base guess: s = a_ij on / above the diagonal, a_ij / a_jj below it
//   guess: F = SIGMA sinh(YS (base + y)), y the network output
#pragma once

#include <hip/hip_runtime.h>

namespace bilu_mlp {

constexpr double SIGMA = 1e-8;
constexpr float YS = 24.0f;
constexpr float INV_SIGMA_F = 1e8f;          // 1 / SIGMA, exact in FP32
constexpr float INV_YS = 1.0f / YS;

// base guess in the scaled space from the FP32-rounded a_ij and (lower) a_jj
__device__ __forceinline__ float base_scaled(float a, float ajj, bool lower)
{
    const float sv = lower ? a / ajj : a;
    return asinhf(sv * INV_SIGMA_F) * INV_YS;
}

__device__ __forceinline__ double guess(float base, float y)
{
    return SIGMA * (double)sinhf(YS * (base + y));
}

}  // namespace bilu_mlp

struct bilu_mlp_s;

// Network outputs for the fused learned solve (defined in mlp.hip): the
// feature map and the three GEMMs for batch <= bilu_mlp_forward_rows()
// systems of d_A on `st`; *Y = the raw outputs [batch][*P3] (FP32, valid
// until the network's next call), *piv = [nnz] the base guess's pivot
// positions (-1 on / above the diagonal).  Rejects (BILU_ERR_INVALID_ARG)
// a network created on another pattern handle than `pattern`, before any
// launch reads d_A.
int bilu_mlp_forward_raw(bilu_mlp_s *q, const void *pattern, int batch, const double *d_A, hipStream_t st,
                         const float **Y, int *P3, const int **piv);
int bilu_mlp_forward_rows();
