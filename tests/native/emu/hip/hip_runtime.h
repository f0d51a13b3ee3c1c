// CPU emulation of the subset of HIP that the library uses (test
// infrastructure).  Kernels run on the host: the workgroups of a launch one
// after another in index order, the lanes of a workgroup as cooperative fibers
// on one thread that switch only at barriers and shuffles.  Device memory is
// host memory, streams and events are synchronous, a graph replays its
// recorded launches in order, atomics are plain read-modify-writes, casts to
// the global address space are ignored, and __all() answers for the calling
// lane alone (a wavefront's lanes may then take different branches of a
// wave-uniform choice, which the kernels allow only between forms that give
// the same result).  Workgroups in index order (or, with EMU_ORDER=reverse,
// from the last down) is the most adverse schedule for protocols between the
// workgroups of one launch: a workgroup starts only after every other one
// before it in that order has finished.
#pragma once
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <functional>
#include <mutex>
#include <vector>
#include <chrono>

using std::min;
using std::max;
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__
#define HIP_SYMBOL(x) (&x)
#define __align__(n) __attribute__((aligned(n)))

struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
struct EmuIdx { unsigned x, y, z; };
extern EmuIdx threadIdx, blockIdx;
extern dim3 blockDim, gridDim;
void emu_block_sync();
void emu_wave_sync();
inline void __syncthreads() { emu_block_sync(); }
extern std::atomic<int> emu_or_flag;
inline int __syncthreads_or(int p) {
    __syncthreads();
    if (threadIdx.x == 0) emu_or_flag = 0;
    __syncthreads();
    if (p) emu_or_flag = 1;
    __syncthreads();
    const int r = emu_or_flag;
    __syncthreads();
    return r;
}
inline void __builtin_amdgcn_wave_barrier() { emu_wave_sync(); }
inline void __builtin_amdgcn_fence(int, const char*) { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }

// shuffles within a wave (every lane of the wave calls them together)
extern std::vector<double> emu_shfl_d;
inline double __shfl_xor(double v, int off, int w = 64) {
    unsigned t = threadIdx.x;
    emu_shfl_d[t] = v; emu_wave_sync();
    unsigned lane = t & 63, base = t - lane;
    double r = emu_shfl_d[base + ((lane ^ off) & 63)];
    emu_wave_sync();
    return r;
}
inline float __shfl_xor(float v, int off, int w = 64) { return (float)__shfl_xor((double)v, off, w); }
inline int __shfl_xor(int v, int off, int w = 64) { return (int)__shfl_xor((double)v, off, w); }

extern std::mutex emu_atomic_mu;
inline int atomicOr(int* p, int v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); int o = *p; *p = o | v; return o; }
inline unsigned atomicOr(unsigned* p, unsigned v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); unsigned o = *p; *p = o | v; return o; }
inline int atomicMin(int* p, int v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); int o = *p; if (v < o) *p = v; return o; }
inline int atomicMax(int* p, int v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); int o = *p; if (v > o) *p = v; return o; }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); unsigned long long o = *p; if (v > o) *p = v; return o; }
inline unsigned int atomicAdd(unsigned int* p, unsigned int v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); unsigned int o = *p; *p = o + v; return o; }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); unsigned long long o = *p; *p = o + v; return o; }
inline unsigned long long atomicExch(unsigned long long* p, unsigned long long v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); unsigned long long o = *p; *p = v; return o; }
inline unsigned int atomicExch(unsigned int* p, unsigned int v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); unsigned int o = *p; *p = v; return o; }
inline double atomicAdd(double* p, double v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); double o = *p; *p = o + v; return o; }
inline float atomicAdd(float* p, float v) {
    std::lock_guard<std::mutex> g(emu_atomic_mu); float o = *p; *p = o + v; return o; }
inline long long __double_as_longlong(double d) { long long l; std::memcpy(&l, &d, 8); return l; }
inline double __longlong_as_double(long long l) { double d; std::memcpy(&d, &l, 8); return d; }
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
inline float __builtin_amdgcn_sqrtf(float x) { return std::sqrt(x); }
inline float __builtin_amdgcn_rsqf(float x) { return 1.0f / std::sqrt(x); }
inline float __expf(float x) { return std::exp(x); }
inline float __powf(float a, float b) { return std::pow(a, b); }
inline float __builtin_amdgcn_logf(float x) { return std::log2(x); }
inline float __builtin_amdgcn_exp2f(float x) { return std::exp2(x); }
// raw buffer resources: base pointer and size in bytes; an access outside
// [0, size) aborts (the kernels never rely on the hardware's range check)
struct __amdgpu_buffer_rsrc_t { char* base; unsigned size; };
inline __amdgpu_buffer_rsrc_t __builtin_amdgcn_make_buffer_rsrc(void* p, short, int n, int) {
    return __amdgpu_buffer_rsrc_t{(char*)p, (unsigned)n}; }
inline char* emu_buf_at(const __amdgpu_buffer_rsrc_t& r, unsigned voff, unsigned soff, unsigned bytes) {
    const unsigned long long o = (unsigned long long)voff + soff;
    if (o + bytes > r.size) { std::fprintf(stderr, "buffer access out of range: %llu + %u > %u\n", o, bytes, r.size); std::abort(); }
    return r.base + o;
}
inline unsigned __builtin_amdgcn_raw_buffer_load_b32(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, int) {
    unsigned v; std::memcpy(&v, emu_buf_at(r, voff, soff, 4), 4); return v; }
inline unsigned long long __builtin_amdgcn_raw_buffer_load_b64(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, int) {
    unsigned long long v; std::memcpy(&v, emu_buf_at(r, voff, soff, 8), 8); return v; }
inline void __builtin_amdgcn_raw_buffer_store_b32(unsigned v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, int) {
    std::memcpy(emu_buf_at(r, voff, soff, 4), &v, 4); }
inline void __builtin_amdgcn_raw_buffer_store_b64(unsigned long long v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, int) {
    std::memcpy(emu_buf_at(r, voff, soff, 8), &v, 8); }
extern "C" float fmaf(float, float, float);
inline float emu_fmaf(float a, float b, float c) { return ::fmaf(a, b, c); }
#define __builtin_fmaf(a, b, c) emu_fmaf(a, b, c)

typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
#define hipSuccess 0
#define hipErrorOutOfMemory 2
#define hipErrorNotSupported 801
#define hipErrorUnknown 999
#define hipErrorInvalidValue 1
#define hipStreamNonBlocking 1
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
struct hipDeviceProp_t { int major; };
inline const char* hipGetErrorString(hipError_t) { return "emu error"; }
inline hipError_t hipGetLastError() { return 0; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return 0; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) { p->major = 9; return 0; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return 0; }
// device allocation: calloc, refused (out of memory) for a single request above
// EMU_MEM_CAP bytes when that variable is set
void* emu_device_alloc(size_t n);
template <class T> inline hipError_t hipMalloc(T** p, size_t n) { *p = (T*)emu_device_alloc(n); return *p ? 0 : 2; }
inline hipError_t hipFree(void* p) { free(p); return 0; }
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return 0; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return 0; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) { std::memmove(d, s, n); return 0; }
inline hipError_t hipMemset(void* d, int v, size_t n) { std::memset(d, v, n); return 0; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { std::memset(d, v, n); return 0; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)1; return 0; }
inline hipError_t hipStreamDestroy(hipStream_t) { return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipDeviceSynchronize() { return 0; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = new std::chrono::steady_clock::time_point; return 0; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete (std::chrono::steady_clock::time_point*)e; return 0; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { *(std::chrono::steady_clock::time_point*)e = std::chrono::steady_clock::now(); return 0; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
#define hipEventDisableTiming 2
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return 0; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(*(std::chrono::steady_clock::time_point*)b - *(std::chrono::steady_clock::time_point*)a).count(); return 0; }

void emu_launch(dim3 grid, dim3 block, const std::function<void()>& body);
inline int __all(int p) { return p; }
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode);
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*);
hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, char*, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t);
hipError_t hipGraphDestroy(hipGraph_t);
hipError_t hipGraphExecDestroy(hipGraphExec_t);
#define hipLaunchKernelGGL(K, G, B, SH, S, ...) emu_launch(dim3(G), dim3(B), [=]() { K(__VA_ARGS__); })
