// Internal host-side definitions of the LandSlide-HySEA MI355X library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <string>
#include "lhysea_common.hpp"

namespace lhysea {

enum Precision { PREC_FP64 = 0, PREC_FP32C = 1 };
enum KernelMode { MODE_FOURPASS = 0, MODE_FUSED = 1 };

// Per-step options shared by every launcher.
struct StepArgs {
    Ctl* ctl;
    const double* dt_seq;   // non-null: frozen dt sequence indexed by ctl->step
    double* emax;           // 2N (value, time) or null
    const int* gauges;      // gauge cells
    int ng;
    double* gseries;        // [gcap][ng]
    int gcap;
    double* dt_hist;        // [hcap]
    int hcap;
    hipEvent_t ev0, ev1;    // when set: recorded around the main step kernel only
};

// Device buffers of one simulation in FP64.
struct Buf64 {
    double* W[2];           // ping-pong state, 6 planes of N each
    double* H;              // N
    double* acc;            // 7N, four-pass path only
};

// Device buffers of one simulation in FP32 with compensation.
struct Buf32 {
    float* W[2];            // ping-pong: 12 planes of N (6 values, then 6 compensations)
    double* H;              // N
    float* acc;             // 7N, four-pass path only
};

// FP64 launchers (lhysea_fp64.hip)
hipError_t f64_initial_dt(const Consts& k, const double* W, Ctl* ctl, hipStream_t s);
hipError_t f64_step_fourpass(const Consts& k, const double* W, double* Wn, const double* H, double* acc,
                             const StepArgs& a, hipStream_t s);
hipError_t f64_step_fused(const Consts& k, const double* W, double* Wn, const double* H, const StepArgs& a,
                          hipStream_t s);

// FP32-compensated launchers (lhysea_fp32c.hip)
hipError_t f32c_initial_dt(const Consts& k, const float* W, Ctl* ctl, hipStream_t s);
hipError_t f32c_step_fused(const Consts& k, const float* W, float* Wn, const double* H, const StepArgs& a,
                           hipStream_t s);

}  // namespace lhysea
