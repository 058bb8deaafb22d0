// Scaling maps of the learned ParILU(0) warm start, shared by the network
// kernels (mlp.hip) and the factor kernel's fused warm start, so both compute
// the guess with the same operations:
//   features / targets: x = asinh(v / SIGMA) / YS
//   base guess: s = a_ij on / above the diagonal, a_ij / a_jj below it
//   guess: F = SIGMA sinh(YS (base + y)), y the network output (guess_from,
//   by an identity that needs no asinh / sinh)
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

// network input: asinh(u) / YS for u = a / SIGMA, as sign(u) ln(t) with
// t = |u| + sqrt(1 + u^2) (2|u| once u^2 would overflow) and the hardware
// log and sqrt: the value is rounded to BF16 next, far coarser than their
// error
__device__ __forceinline__ float feature(float a)
{
    const float u = a * INV_SIGMA_F;
    const float au = fabsf(u);
    const float t = au < 1e18f ? au + __builtin_amdgcn_sqrtf(fmaf(au, au, 1.0f)) : 2.0f * au;
    return copysignf(__logf(t) * INV_YS, u);
}

// the same guess from a_ij, a_jj and y without the asinh / sinh pair:
// SIGMA sinh(asinh(u) + b) with u = s / SIGMA, b = YS y equals
// SIGMA (q - 1/q) / 2 for q = t e^b (u >= 0) or e^b / t (u < 0), where
// t = |u| + sqrt(1 + u^2) = e^asinh|u| is formed without cancellation
// (2|u| once u^2 would overflow); the hardware exp, sqrt and reciprocal
// approximations (also for a_ij / a_jj) suffice for a warm start.
// Overflow and NaN behave as in the asinh / sinh form.
__device__ __forceinline__ double guess_from(float a, float ajj, bool lower, float y)
{
    const float sv = lower ? a * __builtin_amdgcn_rcpf(ajj) : a;
    const float u = sv * INV_SIGMA_F;
    const float au = fabsf(u);
    const float t = au < 1e18f ? au + __builtin_amdgcn_sqrtf(fmaf(au, au, 1.0f)) : 2.0f * au;
    const float eb = __expf(YS * y);
    const float q = u >= 0.0f ? t * eb : eb * __builtin_amdgcn_rcpf(t);
    return SIGMA * (double)(0.5f * (q - __builtin_amdgcn_rcpf(q)));
}

}  // namespace bilu_mlp

struct bilu_mlp_s;

// Network outputs for the fused learned solve (defined in mlp.hip): the
// feature map and the three GEMMs for batch <= bilu_mlp_forward_rows()
// systems of d_A on `st`; *Y = the raw outputs [batch][*P3] rounded to FP16
// (valid until the network's next call; the warm start uses y only through
// e^(YS y), and FP16 keeps it to ~1e-3 relative, far inside the network's
// error).  bilu_mlp_infer rounds y the same way.  Rejects (BILU_ERR_INVALID_ARG) a network
// created on another pattern handle than `pattern`, before any launch reads
// d_A.
int bilu_mlp_forward_raw(bilu_mlp_s *q, const void *pattern, int batch, const double *d_A, hipStream_t st,
                         const _Float16 **Y, int *P3);
int bilu_mlp_forward_rows();
