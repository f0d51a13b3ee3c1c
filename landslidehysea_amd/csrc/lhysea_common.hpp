// Shared device/host definitions of the LandSlide-HySEA MI355X solver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

// EPSILON = FLT_EPSILON is a model constant of the reference
// (src/GPU/Constantes.hxx:13); it stays FLT_EPSILON in every precision.
#define LH_EPS 1.1920928955078125e-07
#define LH_EPSF 1.1920928955078125e-07f
#define LH_SQRT2 1.41421356237309504880
#define LH_BIG 1e30

namespace lhysea {

// FP32 forms of the per-run constants read by the FP32-compensated kernels:
// each is (float) of the FP64 constant or of the FP64 expression named, set
// once on the host so that no kernel wave spends instructions deriving them.
struct ConstsF32 {
    float eps, eps4;          // eps_h, eps^4 (FP32 products)
    float g, gp, r, mu, peso; // gp = g (1 - r)
    float area, len_x, len_y; // dx dy; edge lengths dy (x-edges), dx (y-edges)
    float igp, ten_eps, ieps; // 1 / (g (1 - r)), 10 / eps_h, 1 / eps_h
    float mfc, mf0, mfs, vmax1, vmax2;
    float tan1, tan2, tan3, tan4, LoverH;
};

// Per-run constants, derived on the host from the reference's parameter set
// (src/GPU/main.cxx:157-164, src/GPU/Problema.cxx:226-296).
struct Consts {
    int nx, ny;
    int friction_law;   // 0 Coulomb, 1 Pouliquen
    int pad_;
    double dx, dy, area;
    double borde_sup, borde_inf, borde_izq, borde_der;
    double cfl, r, g, eps_h;
    double mu;          // Coulomb coefficient |tan(angle1)| L/H (src/GPU/Arista_kernel.cu:11-16)
    double tan1, tan2, tan3, tan4, LoverH;   // Pouliquen law
    double mfc, mf0, mfs, vmax1, vmax2;
    double peso, beta;
    ConstsF32 f;
};

// Raw buffer resources on gfx9 targets (gfx950 among them): word 3 selects
// the 32-bit data format, no swizzle, no index stride.  The sediment-free
// step kernels address a state of N cells through one resource with a
// 32-bit byte offset per cell and the plane offset in a scalar register
// while its 48 bytes per cell (12 FP32 or 6 FP64 planes) stay below 4 GiB
// (N < 89,478,486, about 9459 x 9459); larger grids take 64-bit addresses.
constexpr int BUF_DWORD3 = 0x00020000;
__host__ __device__ inline bool state_buf_fits(size_t N) { return N < ((size_t)1 << 32) / 48; }

// Device-side step control block (one per simulation).
// maxA[] holds max_i A_i bit patterns (A_i >= 0, so the IEEE bit order is the
// numeric order and an integer atomicMax is an exact float max).
struct Ctl {
    unsigned long long maxA_next;   // wave-speed max accumulated by the running step
    unsigned int blocks_done;       // last-block-done counter
    int step;                       // steps completed
    double dt_cur;                  // dt of the running step
    double t;                       // simulated time at the start of the running step
    double maxA_last;               // max A of the last completed step (diagnostic)
    unsigned int gcount[2];         // general tiles listed for the steps of parity 0 / 1 (FP32C)
    // step at which the class tables were last rebuilt from a state (upload,
    // restore); the tables describe every step from it on
    int cls_from = 0x7fffffff;
    // health of the state: HEALTH_* bits raised by the cell updates since the
    // state was set, and the first step that raised one
    unsigned int health = 0u;
    int health_step = 0x7fffffff;
};

// Health bits (the C-ABI's LHYSEA_HEALTH_*): a non-finite value in a new
// state, a layer depth below -eps_h after the flux update (before the clip to
// zero; the positivity limiter keeps depths at or above zero).
constexpr unsigned HEALTH_NONFINITE = 1u, HEALTH_NEGATIVE_DEPTH = 2u;

// Records this thread's health bits for step s (nothing when none is set, the
// common case: the branch is skipped).
__device__ __forceinline__ void raise_health(Ctl* ctl, unsigned bits, int s)
{
    if (bits && ctl) {
        atomicOr(&ctl->health, bits);
        atomicMin(&ctl->health_step, s);
    }
}

// dt from the wave-speed max: min_i dt_i with dt_i = 2 CFL area / A_i, or 1e30
// when A_i < EPSILON (src/GPU/Volumen_kernel.cu:215-229, 255-257); the
// monotone map makes this the min of the reference's per-cell dt.
__host__ __device__ inline double dt_from_maxA(double maxA, double cfl, double area)
{
    return (maxA < LH_EPS) ? LH_BIG : (2.0 * cfl * area) / maxA;
}

// Gauge output of one step: eta = h1 + h2 - H of the new state at the gauge
// cells, written by the workgroup that updates the cell (out null: none).
struct GaugeOut {
    const int* cells;
    int ng;
    double* out;
};

}  // namespace lhysea
