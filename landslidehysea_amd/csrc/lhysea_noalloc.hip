// lhysea_step_device: steps of a state held in caller-owned device memory,
// with no device allocation.
//
// For callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).  The caller owns the
// depth, the state and a second state buffer of the same layout (the step
// reads one and writes the other); the library's own device scratch -- the
// step's control block (dt, time, wave-speed max, step and last-workgroup
// counters) and the tile class tables of the split step -- lives in the
// file-scope array `scratch` below, of fixed size: grids of at most
// kMaxClassTiles class tiles (FP64: 30 x 14 cells per tile, FP32C: 62 x 14),
// e.g. every grid up to 4096 x 4096 in either precision.  Each GPU holds its
// own copy of the array until the process ends; its device address is looked
// up once per device (hipGetSymbolAddress) and used wherever host code hands
// the scratch to a kernel.  Every call resets the control block and rebuilds
// the class tables with kernels (no memset), so no state carries over from
// the previous call; calls that overlap on one GPU share the array and race.
#include <cmath>
#include <mutex>
#include "lhysea_internal.hpp"
#include "lhysea_tiles.hpp"
#include "../../include/lhysea.h"

using namespace lhysea;

namespace {

constexpr int kMaxClassTiles = 65536;
constexpr size_t kTablesOffset = 256;   // class tables after the control block
constexpr size_t kScratchBytes = kTablesOffset + 2 * (2 * (size_t)kMaxClassTiles) * sizeof(int);
static_assert(sizeof(Ctl) <= kTablesOffset, "control block fits before the tables");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// device address of this device's scratch array (looked up on its first use)
int scratch_on_current_device(unsigned char** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_last_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
        return LHYSEA_ERR_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        set_last_error("device index beyond the scratch table");
        return LHYSEA_ERR_HIP;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return LHYSEA_ERR_HIP;
        }
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return LHYSEA_OK;
}

// Control block at step 0 with time t and the dt of the first step (0: to be
// computed); no class table yet, no health bit.
__global__ void k_ctl_reset(Ctl* ctl, double t, double dt)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->maxA_next = 0ull;
        ctl->blocks_done = 0u;
        ctl->step = 0;
        ctl->dt_cur = dt;
        ctl->t = t;
        ctl->maxA_last = 0.0;
        ctl->gcount[0] = ctl->gcount[1] = 0u;
        ctl->cls_from = 0x7fffffff;
        ctl->health = 0u;
        ctl->health_step = 0x7fffffff;
    }
}

int hip_err(hipError_t e, const char* where)
{
    set_last_error(std::string(where) + ": " + hipGetErrorString(e));
    return LHYSEA_ERR_HIP;
}

}  // namespace

extern "C" int lhysea_step_device(const lhysea_params* p, const double* H, void* W, void* W_work, int32_t nsteps,
                                  double* t, double* dt, int32_t* result_in_work, void* stream)
{
    if (!p || !H || !W || !W_work || !t || !dt || !result_in_work) {
        set_last_error("null argument");
        return LHYSEA_ERR_ARGUMENT;
    }
    if (nsteps < 0 || W == W_work) {
        set_last_error("nsteps must be >= 0 and W, W_work distinct buffers");
        return LHYSEA_ERR_ARGUMENT;
    }
    if (const int rc = check_params(*p)) return rc;
    const Consts k = make_consts(*p);
    const bool f64 = p->precision == LHYSEA_FP64;
    const int ntiles = f64 ? f64_tiles(k) : f32c_tiles(k);
    if (ntiles > kMaxClassTiles) {
        set_last_error("grid beyond the no-allocation scratch: at most 65536 class tiles (FP64 30 x 14, "
                       "FP32C 62 x 14 cells each)");
        return LHYSEA_ERR_ARGUMENT;
    }
    if (lhysea_device_probe() != 0) {
        set_last_error("no usable HIP device");
        return LHYSEA_ERR_NO_DEVICE;
    }
    unsigned char* base = nullptr;
    if (const int rc = scratch_on_current_device(&base)) return rc;
    Ctl* ctl = (Ctl*)base;
    int* tcls = (int*)(base + kTablesOffset);
    int* tlist = tcls + 2 * kMaxClassTiles;
    hipStream_t s = (hipStream_t)stream;

    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(64), 0, s, ctl, *t, (*dt > 0.0) ? *dt : 0.0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_err(e, "control block reset");
    if (!(*dt > 0.0)) {
        e = f64 ? f64_initial_dt(k, (const double*)W, ctl, s) : f32c_initial_dt(k, (const float*)W, ctl, s);
        if (e != hipSuccess) return hip_err(e, "initial dt");
    }
    StepArgs a{};
    a.ctl = ctl;
    a.t_end = INFINITY;
    a.tcls = tcls;
    a.tlist = tlist;
    a.wn_prev = 1;   // ping-pong: the first two steps after the classification write every plane
    e = f64 ? f64_classify(k, (const double*)W, a, s) : f32c_classify(k, (const float*)W, a, s);
    if (e != hipSuccess) return hip_err(e, "tile classification");
    // the split launch pays when enough tiles are sediment-free
    Ctl c;
    e = hipMemcpyAsync(&c, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(e, "control block read");
    a.split = (ntiles - (int)c.gcount[0]) >= f32c_split_min_bare();

    void* buf[2] = {W, W_work};
    for (int n = 0; n < nsteps; n++) {
        a.step = n;
        const void* src = buf[n & 1];
        void* dst = buf[(n + 1) & 1];
        e = f64 ? f64_step_fused(k, (const double*)src, (double*)dst, H, a, s)
                : f32c_step_fused(k, (const float*)src, (float*)dst, H, a, s);
        if (e != hipSuccess) return hip_err(e, "step launch");
    }
    e = hipMemcpyAsync(&c, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(e, "control block read");
    *t = c.t;
    *dt = c.dt_cur;
    *result_in_work = nsteps & 1;
    return c.health ? numeric_fail(c) : LHYSEA_OK;
}
