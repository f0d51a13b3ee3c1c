// Internal host-side definitions of the LandSlide-HySEA MI355X library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <string>
#include <utility>
#include <vector>
#include "lhysea_common.hpp"

struct lhysea_params;

namespace lhysea {

// Shared by the C-ABI files (lhysea_capi.hip)
Consts make_consts(const lhysea_params& p);
// message of the failure being returned (also clears HIP's per-thread last error)
void set_last_error(const std::string& msg);
// LHYSEA_ERR_NUMERIC, the last error naming the health bits of c and their first step
int numeric_fail(const Ctl& c);
// argument checks of lhysea_create (0, or an LHYSEA_ERR_ code with the message set)
int check_params(const lhysea_params& p);

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
    double t_end;           // fused kernels: a step launched at t >= t_end does nothing
    // FP32C fused step by tile class: tiles whose tile + halo box holds
    // sediment at step s (tcls[s & 1][t] == s, listed in tlist[s & 1][0 ..
    // ctl->gcount[s & 1])) take the general tile function, the others the
    // sediment-free one
    int* tcls;              // [2][ntiles]
    int* tlist;             // [2][ntiles]
    int split;              // 1: sediment-free and general launches, 0: every tile general
    int step;               // the step this launch performs (ctl->step when it starts)
    // 1: Wn holds the state two steps back (ping-pong buffers), whose layer-2
    // planes are zero on every tile that step s - 2 took as sediment-free, so
    // the sediment-free launch need not write them again; 0: Wn is a fresh
    // buffer (every state kept), all planes are written
    int wn_prev;
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
// eta max of a fresh FP64-layout state W (6 planes of N): (h1 + h2 - H, 0) per cell
hipError_t init_eta_max(const double* W, const double* H, double* emax, size_t N, hipStream_t s);
hipError_t f64_initial_dt(const Consts& k, const double* W, Ctl* ctl, hipStream_t s);
hipError_t f64_step_fourpass(const Consts& k, const double* W, double* Wn, const double* H, double* acc,
                             const StepArgs& a, hipStream_t s);
hipError_t f64_step_fused(const Consts& k, const double* W, double* Wn, const double* H, const StepArgs& a,
                          hipStream_t s);
// tiles of the FP64 split step and its class table (tcls, tlist: 2 * f64_tiles ints each)
int f64_tiles(const Consts& k);
hipError_t f64_classify(const Consts& k, const double* W, const StepArgs& a, hipStream_t s);

// FP32-compensated launchers (lhysea_fp32c.hip)
// FP64 planes (6 of N) to the compensated layout and back
hipError_t f32c_split_state(const double* src, float* W, size_t N, hipStream_t s);
hipError_t f32c_merge_state(const float* W, double* dst, size_t N, hipStream_t s);
hipError_t f32c_initial_dt(const Consts& k, const float* W, Ctl* ctl, hipStream_t s);
hipError_t f32c_step_fused(const Consts& k, const float* W, float* Wn, const double* H, const StepArgs& a,
                           hipStream_t s);
// tiles of the split step and their class table (tcls, tlist: 2 * f32c_tiles ints each)
int f32c_tiles(const Consts& k);
// fewest sediment-free tiles for which the split step pays
int f32c_split_min_bare();
// class table of state W at step ctl->step (after an upload or a restore)
hipError_t f32c_classify(const Consts& k, const float* W, const StepArgs& a, hipStream_t s);
// partial h sums read by classes 2..4 (P, [3][2][N]) and flux sums (ACC, [6][N])
// of the step from W with dt = dt_seq[step], by the fused tile kernel
// FP64 forward recomputation for the adjoint (partial sums P [3][2][N], flux sums ACC [6][N])
hipError_t f64_recompute(const Consts& k, const double* W, const double* H, const double* dt_seq, int step, double* P,
                         double* ACC, hipStream_t s);
hipError_t f32c_recompute(const Consts& k, const float* W, const double* H, const double* dt_seq, int step, float* P,
                          float* ACC, hipStream_t s);

// One scenario of an ensemble batch as the batched kernels read it (device copy).
struct BatchScen {
    Consts k;
    const double* H;        // N
    float* W[2];            // ping-pong, 12 planes of N; the current one is W[ctl->step & 1]
    Ctl* ctl;
    double* emax;           // 2N or null
    const int* gauges;
    double* gseries;        // row r holds step r (base already offset)
    double* dt_hist;
    double t_end;           // a scenario steps while t < t_end and step < max_steps
    int ng, gcap, hcap, max_steps;
    int tiles_x, tiles_y;
    // class tiles (62 x 14) of the split shape: table and list per parity
    int* cls;
    int* list;
    int ctiles_x, ctiles_y;
};

// Tile shapes of the batched kernel (TX x TY cells, NT threads)
struct BatchTile {
    int tx, ty, nt;
};
constexpr int kNumBatchTiles = 4;
extern const BatchTile kBatchTiles[kNumBatchTiles];
// buf: every scenario's state passes state_buf_fits (buffer-addressed kernels)
hipError_t f32c_batch_step(int tile, const BatchScen* sd, int nscen, const unsigned* tab, int nblocks, bool buf,
                           hipStream_t s);
// scenarios of a batch at most
constexpr int kMaxBatchScen = 4096;
// shape kBatchSplit steps sediment-free class tiles and general ones apart
constexpr int kBatchSplit = 0;
// class table of scenario sid of the batch at its current step (device descriptors in sd)
hipError_t f32c_batch_classify(const BatchScen& d, const BatchScen* sd_dev, int sid, hipStream_t s);

// Adjoint (lhysea_adjoint.hip).  gmu: the friction-coefficient adjoint slots,
// one per workgroup: d/dmu (Coulomb) at gmu[slot], or d/dtan(angle i) of the
// Pouliquen law (friction_law 1) at gmu[i * gstride + slot], i = 0..3.
hipError_t adj_reverse_step_f32(const Consts& k, const float* W, const double* H, float* P, float* ACC, float* AB,
                                double* LAM, double* gmu, size_t gstride, const double* dt_hist, int step,
                                hipStream_t s);
// FP32 reverse step as two launches: the recomputation and one tiled kernel
// for the cell and edge adjoints (lambda^{n+1} in LAMin, lambda^n to LAMout),
// which also adds the gauge seeds Gk (ng values at the gauge cells; null: none)
hipError_t adj_reverse_step_f32_fused(const Consts& k, const float* W, const double* H, float* P, float* ACC,
                                      const double* LAMin, double* LAMout, double* gmu, size_t gstride,
                                      const double* dt_hist, int step, const int* gauges, int ng, const double* Gk,
                                      hipStream_t s);
hipError_t adj_reverse_step_f64(const Consts& k, const double* W, const double* H, double* P, double* ACC,
                                double* AB, double* LAM, double* gmu, size_t gstride, const double* dt_hist, int step,
                                hipStream_t s);
hipError_t adj_inject(double* LAM, size_t N, const int* gauges, int ng, const double* G, hipStream_t s);
// friction-coefficient adjoint slots per component (one per workgroup of the
// largest reverse-step launch)
size_t adj_gmu_slots(const Consts& k);
hipError_t adj_set_step(Ctl* ctl, int step, hipStream_t s);

struct AdjointHooks {
    hipError_t (*advance)(void* ctx, int k0, int n);
    hipError_t (*store)(void* ctx, int slot);
    hipError_t (*restore)(void* ctx, int slot);
    hipError_t (*reverse)(void* ctx, int k);
    void* ctx;
    long advances, reverses;
};
long revolve_advances(long l, int c);
long revolve_split(long l, int c);
void revolve_chain(long N, int S, std::vector<std::pair<long, int>>& out);
hipError_t run_treeverse(AdjointHooks& h, int S, long N);

}  // namespace lhysea
