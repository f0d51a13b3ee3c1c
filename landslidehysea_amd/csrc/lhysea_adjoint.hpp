// Discrete adjoint device functions (forward recomputation + reverse sweep),
// generated from the step's arithmetic by csrc/gen/gen_adjoint.py.
#pragma once
#include "lhysea_common.hpp"

namespace lhysea {
namespace adj {

// FP32: hardware reciprocal / square root / exp / log2, as in the forward step
__device__ __forceinline__ float ad_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float ad_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float ad_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float ad_exp(float x) { return __expf(x); }
__device__ __forceinline__ float ad_pow(float x, float c)
{
    return __builtin_amdgcn_exp2f(c * __builtin_amdgcn_logf(x));   // x >= 0, as the forward step
}
__device__ __forceinline__ float ad_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// FP64
__device__ __forceinline__ double ad_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ double ad_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double ad_rsq(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ double ad_exp(double x) { return exp(x); }
__device__ __forceinline__ double ad_pow(double x, double c) { return pow(x, c); }
__device__ __forceinline__ double ad_div(double a, double b) { return a / b; }

#include "lhysea_adjoint_gen.inc"

}  // namespace adj
}  // namespace lhysea
