// Minimal host stand-in for <hip/hip_runtime.h>: enough to compile the
// header-only device functions of landslidehysea_amd/csrc/lhysea_fp32.hpp and
// lhysea_common.hpp with g++ for CPU checks.  Hardware reciprocal, square
// root and exp become correctly rounded host functions.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
using std::fabs;
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
inline float __builtin_amdgcn_sqrtf(float x) { return std::sqrt(x); }
inline float __builtin_amdgcn_rsqf(float x) { return 1.0f / std::sqrt(x); }
inline float __expf(float x) { return std::exp(x); }
inline float __powf(float a, float b) { return std::pow(a, b); }
inline float __builtin_amdgcn_logf(float x) { return std::log2(x); }
inline float __builtin_amdgcn_exp2f(float x) { return std::exp2(x); }
// single-threaded host checks: plain read-modify-write
inline unsigned atomicOr(unsigned* p, unsigned v) { unsigned o = *p; *p = o | v; return o; }
inline int atomicMin(int* p, int v) { int o = *p; if (v < o) *p = v; return o; }
