// FP64 kernels of the LandSlide-HySEA solver step on MI355X (gfx950).
//
// Two step paths share the same edge and cell device code (lhysea_fp64.hpp):
//  * four-pass: one launch per edge class in the reference's order
//    (even y, odd y, even x, odd x; src/GPU/ShallowWater.cu:391-427) with
//    accumulators in HBM, then one cell-update launch.  It is the
//    structural twin of the reference and a GPU-side cross-check.
//  * fused: one launch per step.  A workgroup owns a TX x TY tile of cells
//    whose origin is even in x and y, stages the tile plus a 1-cell halo in
//    LDS, evaluates the four edge classes in order with LDS accumulators
//    (the order-dependent positivity limiter reads partial sums, so the
//    class order is reproduced exactly), applies the cell update and writes
//    the new state to the other ping-pong buffer.
// Both end with a wavefront-shuffle max of the per-cell wave-speed sums and
// one 64-bit atomicMax per workgroup; a one-wave finalize launch turns the
// max into the lagged dt of the next step on the device (no host sync).
#pragma clang fp contract(off)
#include <type_traits>
#include "lhysea_internal.hpp"
#include "lhysea_fp64.hpp"
#include "lhysea_tiles.hpp"

namespace lhysea {
namespace f64 {

__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Workgroup max of v (all threads call); thread 0 issues one atomicMax.
template <int NT>
__device__ __forceinline__ void block_max_atomic(double v, unsigned long long* slot)
{
    __shared__ double red[NT / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int i = 1; i < NT / 64; i++) m = fmax(m, red[i]);
        if (m > 0.0) atomicMax(slot, (unsigned long long)__double_as_longlong(m));
    }
}

__device__ __forceinline__ double step_dt(const Ctl* ctl, const double* dt_seq)
{
    return dt_seq ? dt_seq[ctl->step] : ctl->dt_cur;
}

__device__ __forceinline__ void load6(const double* __restrict__ W, size_t N, size_t c, double* o)
{
#pragma unroll
    for (int f = 0; f < 6; f++) o[f] = W[f * N + c];
}

// ---------------------------------------------------------------- initial dt
// Per-cell gather of the four edges' wave-speed terms in class order, so the
// sum rounds like the reference's class-by-class accumulation
// (src/GPU/ShallowWater.cu:238-251).
__global__ void __launch_bounds__(256) k_dt0(Consts k, const double* __restrict__ W, Ctl* ctl)
{
    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const long cidx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double A = 0.0;
    if (cidx < (long)N) {
        const int i = (int)(cidx % nx), j = (int)(cidx / nx);
        double Wc[6], Wo[6];
        load6(W, N, cidx, Wc);
        double v[4];
        // y-edges: top je = j, bottom je = j + 1; x-edges: left ie = i, right ie = i + 1
        for (int s = 0; s < 2; s++) {
            const int je = j + s;
            double val;
            if (je == 0 || je == ny) {
                const double bd = (je == 0) ? k.borde_sup : k.borde_inf;
                for (int f = 0; f < 6; f++) Wo[f] = Wc[f];
                Wo[2] *= bd; Wo[5] *= bd;
                val = edge_dt0(k, Wc, Wo, 0.0, (je == 0) ? -1.0 : 1.0, k.dx);
            } else if (s == 0) {
                load6(W, N, cidx - nx, Wo);
                val = edge_dt0(k, Wo, Wc, 0.0, 1.0, k.dx);
            } else {
                load6(W, N, cidx + nx, Wo);
                val = edge_dt0(k, Wc, Wo, 0.0, 1.0, k.dx);
            }
            v[(je & 1)] = val;
        }
        for (int s = 0; s < 2; s++) {
            const int ie = i + s;
            double val;
            if (ie == 0 || ie == nx) {
                const double bd = (ie == 0) ? k.borde_izq : k.borde_der;
                for (int f = 0; f < 6; f++) Wo[f] = Wc[f];
                Wo[1] *= bd; Wo[4] *= bd;
                val = edge_dt0(k, Wc, Wo, (ie == 0) ? -1.0 : 1.0, 0.0, k.dy);
            } else if (s == 0) {
                load6(W, N, cidx - 1, Wo);
                val = edge_dt0(k, Wo, Wc, 1.0, 0.0, k.dy);
            } else {
                load6(W, N, cidx + 1, Wo);
                val = edge_dt0(k, Wc, Wo, 1.0, 0.0, k.dy);
            }
            v[2 + (ie & 1)] = val;
        }
        A = ((v[0] + v[1]) + v[2]) + v[3];
    }
    block_max_atomic<256>(A, &ctl->maxA_next);
}

__global__ void k_dt0_finalize(Consts k, Ctl* ctl)
{
    if (threadIdx.x == 0) {
        double maxA = __longlong_as_double((long long)ctl->maxA_next);
        ctl->dt_cur = dt_from_maxA(maxA, k.cfl, k.area);
        ctl->maxA_last = maxA;
        ctl->maxA_next = 0ull;
    }
}

// Step finalize: record dt, sample gauges on the new state, advance time and
// derive the next (lagged) dt from this step's wave-speed max
// (src/GPU/ShallowWater.cu:431-442).
__global__ void k_finalize(Consts k, const double* __restrict__ Wn, const double* __restrict__ H, StepArgs a)
{
    Ctl* ctl = a.ctl;
    if (!(ctl->t < a.t_end)) return;
    const size_t N = (size_t)k.nx * k.ny;
    const int s = ctl->step;
    const double dt = step_dt(ctl, a.dt_seq);
    if (a.gseries && s < a.gcap)
        for (int g = threadIdx.x; g < a.ng; g += blockDim.x) {
            const size_t c = (size_t)a.gauges[g];
            a.gseries[(size_t)s * a.ng + g] = Wn[c] + Wn[3 * N + c] - H[c];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (a.dt_hist && s < a.hcap) a.dt_hist[s] = dt;
        double maxA = __longlong_as_double((long long)ctl->maxA_next);
        ctl->t = ctl->t + dt;
        ctl->dt_cur = dt_from_maxA(maxA, k.cfl, k.area);
        ctl->maxA_last = maxA;
        ctl->maxA_next = 0ull;
        ctl->step = s + 1;
    }
}

// ------------------------------------------------------------ four-pass path
// One thread per edge of one class (src/GPU/Arista_kernel.cu:651-949).
__global__ void __launch_bounds__(256) k_edges(Consts k, const double* __restrict__ W,
                                               const double* __restrict__ H, double* __restrict__ acc,
                                               const Ctl* ctl, const double* dt_seq, int cls)
{
    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const double dt = step_dt(ctl, dt_seq);
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double W0[6], W1[6];
    size_t c0, c1 = 0;
    double n1x, n1y, len, H0, H1;
    bool interior;
    if (cls < 2) {
        const int nrow = (ny - cls) / 2 + 1;
        if (e >= (long)nrow * nx) return;
        const int je = cls + 2 * (int)(e / nx), i = (int)(e % nx);
        len = k.dx;
        n1x = 0.0;
        if (je == 0 || je == ny) {
            c0 = (je == 0) ? (size_t)i : (size_t)(ny - 1) * nx + i;
            const double bd = (je == 0) ? k.borde_sup : k.borde_inf;
            load6(W, N, c0, W0);
            for (int f = 0; f < 6; f++) W1[f] = W0[f];
            W1[2] *= bd; W1[5] *= bd;
            n1y = (je == 0) ? -1.0 : 1.0;
            H0 = H1 = H[c0];
            interior = false;
        } else {
            c0 = (size_t)(je - 1) * nx + i;
            c1 = c0 + nx;
            load6(W, N, c0, W0);
            load6(W, N, c1, W1);
            n1y = 1.0;
            H0 = H[c0];
            H1 = H[c1];
            interior = true;
        }
    } else {
        const int par = cls - 2, ncol = (nx - par) / 2 + 1;
        if (e >= (long)ncol * ny) return;
        const int j = (int)(e / ncol), ie = par + 2 * (int)(e % ncol);
        len = k.dy;
        n1y = 0.0;
        if (ie == 0 || ie == nx) {
            c0 = (ie == 0) ? (size_t)j * nx : (size_t)j * nx + nx - 1;
            const double bd = (ie == 0) ? k.borde_izq : k.borde_der;
            load6(W, N, c0, W0);
            for (int f = 0; f < 6; f++) W1[f] = W0[f];
            W1[1] *= bd; W1[4] *= bd;
            n1x = (ie == 0) ? -1.0 : 1.0;
            H0 = H1 = H[c0];
            interior = false;
        } else {
            c0 = (size_t)j * nx + ie - 1;
            c1 = c0 + 1;
            load6(W, N, c0, W0);
            load6(W, N, c1, W1);
            n1x = 1.0;
            H0 = H[c0];
            H1 = H[c1];
            interior = true;
        }
    }
    const double a0h1 = acc[c0], a0h2 = acc[3 * N + c0];
    const double a1h1 = interior ? acc[c1] : 0.0, a1h2 = interior ? acc[3 * N + c1] : 0.0;
    EdgeFlux F;
    const bool wet = (cls < 2) ? edge<1>(k, W0, W1, H0, H1, n1y, len, dt, a0h1, a0h2, a1h1, a1h2, interior, F)
                               : edge<0>(k, W0, W1, H0, H1, n1x, len, dt, a0h1, a0h2, a1h1, a1h2, interior, F);
    if (wet) {
        for (int f = 0; f < 6; f++) acc[f * N + c0] -= F.Fm[f];
        acc[6 * N + c0] += F.lam;
        if (interior) {
            for (int f = 0; f < 6; f++) acc[f * N + c1] -= F.Fp[f];
            acc[6 * N + c1] += F.lam;
        }
    }
}

// Cell update of the four-pass path; clears the accumulators for the next step.
__global__ void __launch_bounds__(256) k_update(Consts k, const double* __restrict__ W, double* __restrict__ Wn,
                                                const double* __restrict__ H, double* __restrict__ acc, StepArgs a)
{
    const size_t N = (size_t)k.nx * k.ny;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double A = 0.0;
    if (c < (long)N) {
        const double dt = step_dt(a.ctl, a.dt_seq);
        double o[6], ac[6], n[6];
        load6(W, N, c, o);
        for (int f = 0; f < 6; f++) ac[f] = acc[f * N + c];
        A = acc[6 * N + c];
        if (a.emax) {
            const double v = o[0] - H[c];
            if (v > a.emax[2 * c]) {
                a.emax[2 * c] = v;
                a.emax[2 * c + 1] = a.ctl->t;
            }
        }
        unsigned hb = 0u;
        cell_update(k, o, ac, dt, n, &hb);
        for (int f = 0; f < 6; f++) Wn[f * N + c] = n[f];
        for (int f = 0; f < 7; f++) acc[f * N + c] = 0.0;
        raise_health(a.ctl, hb, a.ctl->step);
    }
    block_max_atomic<256>(A, &a.ctl->maxA_next);
}

// ----------------------------------------------------------------- fused path
// Both cells of an edge without sediment (h2 = q2x = q2y = 0).  A wavefront
// whose edges all qualify takes edge1(), which yields no layer-2 fluxes, so
// the layer-2 accumulators are left as they are.
__device__ __forceinline__ bool no_sediment(const double* W0, const double* W1)
{
    return (W0[3] == 0.0) && (W0[4] == 0.0) && (W0[5] == 0.0) && (W1[3] == 0.0) && (W1[4] == 0.0) && (W1[5] == 0.0);
}

// sA[f][r][c] -= F[f] for the first NF fields, wave-speed sum += lam
template <int NF, int TY, int CW>
__device__ __forceinline__ void accumulate(double (&sA)[7][TY][CW], int r, int c, const double* F, double lam)
{
#pragma unroll
    for (int f = 0; f < NF; f++) sA[f][r][c] -= F[f];
    sA[6][r][c] += lam;
}

// Tile of TX x TY cells (both even), NT threads, origin (i0, j0).  LDS planes:
//   sW[f][r][c]  f = h1,q1x,q1y,h2,q2x,q2y,H ; r = tile row + 1 ; c = tile col + 1
//   sA[f][r][c]  f = dh1..dq2y, A            ; r = tile row     ; c = tile col + 1
// Returns this thread's max of the wave-speed sums of the cells it updated.
// go: gauges of the new state.  smask (optional): OR of the class-tile halo
// masks of this thread's cells that keep sediment, the class tile at
// (cx0, cy0).  hc (optional): control block whose health word records the
// cell updates' health bits (HEALTH_*) as those of step hs.
// RECOMP: instead of updating the state, write the partial h1/h2 sums that
// classes 2..4 read (P, [3][2][N]) and the final flux sums (ACC, [6][N]) of
// the tile's cells: the forward recomputation of the discrete adjoint.
template <int TX, int TY, int NT, bool RECOMP = false>
__device__ __forceinline__ double tile_step64(const Consts& k, const double* __restrict__ W, double* __restrict__ Wn,
                                              const double* __restrict__ H, double dt, double tnow,
                                              double* __restrict__ emax, int i0, int j0, GaugeOut go, int* smask,
                                              int cx0, int cy0, Ctl* hc = nullptr, int hs = 0,
                                              double* __restrict__ P = nullptr, double* __restrict__ ACC = nullptr)
{
    constexpr int CW = TX + 2;              // staged columns
    constexpr int RW = TY + 2;              // staged rows
    __shared__ double sW[7][RW][CW];
    __shared__ double sA[7][TY][CW];

    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const int tid = threadIdx.x;

    // stage state + depth with a 1-cell halo; zero the accumulators
    for (int e = tid; e < RW * CW; e += NT) {
        const int r = e / CW, c = e % CW;
        const int gj = j0 + r - 1, gi = i0 + c - 1;
        if (gj >= 0 && gj < ny && gi >= 0 && gi < nx) {
            const size_t g = (size_t)gj * nx + gi;
#pragma unroll
            for (int f = 0; f < 6; f++) sW[f][r][c] = W[f * N + g];
            sW[6][r][c] = H[g];
        }
    }
    for (int e = tid; e < TY * CW; e += NT) {
        const int r = e / CW, c = e % CW;
#pragma unroll
        for (int f = 0; f < 7; f++) sA[f][r][c] = 0.0;
    }
    __syncthreads();

    // RECOMP: the tile's own h1/h2 partial sums after class cls (read by class cls + 1)
    auto store_partials = [&](int cls) {
        double* Pc = P + (size_t)cls * 2 * N;
        for (int e = tid; e < TX * TY; e += NT) {
            const int r = e / TX, c = e % TX;
            const int gj = j0 + r, gi = i0 + c;
            if (gj >= ny || gi >= nx) continue;
            const size_t g = (size_t)gj * nx + gi;
            Pc[g] = sA[0][r][c + 1];
            Pc[N + g] = sA[3][r][c + 1];
        }
        __syncthreads();   // before the next class adds to the sums
    };

    // y-edge classes C1 (even) then C2 (odd), on staged columns -1..TX
#pragma unroll 1
    for (int par = 0; par < 2; par++) {
        constexpr int CWc = CW;
        const int nrow = (par == 0) ? (TY / 2 + 1) : (TY / 2);
        for (int e = tid; e < nrow * CWc; e += NT) {
            const int c = e % CWc, jl = par + 2 * (e / CWc);   // tile-local edge row
            const int gi = i0 + c - 1, je = j0 + jl;
            if (gi < 0 || gi >= nx || je > ny) continue;
            double W0[6], W1[6], H0, H1, a1h1 = 0.0, a1h2 = 0.0;
            int r0;                      // tile row of cell 0
            bool interior;
            double n1y;
            if (je == 0 || je == ny) {
                r0 = (je == 0) ? jl : jl - 1;
                const double bd = (je == 0) ? k.borde_sup : k.borde_inf;
#pragma unroll
                for (int f = 0; f < 6; f++) { W0[f] = sW[f][r0 + 1][c]; W1[f] = W0[f]; }
                W1[2] *= bd; W1[5] *= bd;
                H0 = H1 = sW[6][r0 + 1][c];
                n1y = (je == 0) ? -1.0 : 1.0;
                interior = false;
            } else {
                r0 = jl - 1;
#pragma unroll
                for (int f = 0; f < 6; f++) { W0[f] = sW[f][r0 + 1][c]; W1[f] = sW[f][r0 + 2][c]; }
                H0 = sW[6][r0 + 1][c];
                H1 = sW[6][r0 + 2][c];
                n1y = 1.0;
                interior = true;
            }
            const bool has0 = (r0 >= 0 && r0 < TY), has1 = interior && (r0 + 1 < TY);
            const double a0h1 = has0 ? sA[0][r0][c] : 0.0;
            if (has1) a1h1 = sA[0][r0 + 1][c];
            EdgeFlux F;
            if (__all(no_sediment(W0, W1))) {
                if (edge1<1>(k, W0, W1, H0, H1, n1y, k.dx, dt, a0h1, a1h1, interior, F)) {
                    if (has0) accumulate<3>(sA, r0, c, F.Fm, F.lam);
                    if (has1) accumulate<3>(sA, r0 + 1, c, F.Fp, F.lam);
                }
            } else {
                const double a0h2 = has0 ? sA[3][r0][c] : 0.0;
                if (has1) a1h2 = sA[3][r0 + 1][c];
                if (edge<1>(k, W0, W1, H0, H1, n1y, k.dx, dt, a0h1, a0h2, a1h1, a1h2, interior, F)) {
                    if (has0) accumulate<6>(sA, r0, c, F.Fm, F.lam);
                    if (has1) accumulate<6>(sA, r0 + 1, c, F.Fp, F.lam);
                }
            }
        }
        __syncthreads();
        if (RECOMP) store_partials(par);
    }

    // x-edge classes C3 (even) then C4 (odd), on tile rows 0..TY-1
#pragma unroll 1
    for (int par = 0; par < 2; par++) {
        const int ncol = (par == 0) ? (TX / 2 + 1) : (TX / 2);
        for (int e = tid; e < ncol * TY; e += NT) {
            const int r = e / ncol, il = par + 2 * (e % ncol);   // tile-local edge column
            const int gj = j0 + r, ie = i0 + il;
            if (gj >= ny || ie > nx) continue;
            double W0[6], W1[6], H0, H1, a1h1 = 0.0, a1h2 = 0.0;
            int c0;                      // staged column of cell 0
            bool interior;
            double n1x;
            if (ie == 0 || ie == nx) {
                c0 = (ie == 0) ? il + 1 : il;
                const double bd = (ie == 0) ? k.borde_izq : k.borde_der;
#pragma unroll
                for (int f = 0; f < 6; f++) { W0[f] = sW[f][r + 1][c0]; W1[f] = W0[f]; }
                W1[1] *= bd; W1[4] *= bd;
                H0 = H1 = sW[6][r + 1][c0];
                n1x = (ie == 0) ? -1.0 : 1.0;
                interior = false;
            } else {
                c0 = il;
#pragma unroll
                for (int f = 0; f < 6; f++) { W0[f] = sW[f][r + 1][c0]; W1[f] = sW[f][r + 1][c0 + 1]; }
                H0 = sW[6][r + 1][c0];
                H1 = sW[6][r + 1][c0 + 1];
                n1x = 1.0;
                interior = true;
            }
            const double a0h1 = sA[0][r][c0];
            if (interior) a1h1 = sA[0][r][c0 + 1];
            EdgeFlux F;
            if (__all(no_sediment(W0, W1))) {
                if (edge1<0>(k, W0, W1, H0, H1, n1x, k.dy, dt, a0h1, a1h1, interior, F)) {
                    accumulate<3>(sA, r, c0, F.Fm, F.lam);
                    if (interior) accumulate<3>(sA, r, c0 + 1, F.Fp, F.lam);
                }
            } else {
                const double a0h2 = sA[3][r][c0];
                if (interior) a1h2 = sA[3][r][c0 + 1];
                if (edge<0>(k, W0, W1, H0, H1, n1x, k.dy, dt, a0h1, a0h2, a1h1, a1h2, interior, F)) {
                    accumulate<6>(sA, r, c0, F.Fm, F.lam);
                    if (interior) accumulate<6>(sA, r, c0 + 1, F.Fp, F.lam);
                }
            }
        }
        __syncthreads();
        if (RECOMP && par == 0) store_partials(2);
    }

    if (RECOMP) {
        for (int e = tid; e < TX * TY; e += NT) {
            const int r = e / TX, c = e % TX;
            const int gj = j0 + r, gi = i0 + c;
            if (gj >= ny || gi >= nx) continue;
            const size_t g = (size_t)gj * nx + gi;
#pragma unroll
            for (int f = 0; f < 6; f++) ACC[f * N + g] = sA[f][r][c + 1];
        }
        return 0.0;
    }

    // cell update of the tile's own cells
    double Amax = 0.0;
    unsigned hb = 0u;
    for (int e = tid; e < TX * TY; e += NT) {
        const int r = e / TX, c = e % TX;
        const int gj = j0 + r, gi = i0 + c;
        if (gj >= ny || gi >= nx) continue;
        const size_t g = (size_t)gj * nx + gi;
        double o[6], ac[6], n[6];
#pragma unroll
        for (int f = 0; f < 6; f++) { o[f] = sW[f][r + 1][c + 1]; ac[f] = sA[f][r][c + 1]; }
        Amax = fmax(Amax, sA[6][r][c + 1]);
        if (emax) {
            const double v = o[0] - sW[6][r + 1][c + 1];
            if (v > emax[2 * g]) {
                emax[2 * g] = v;
                emax[2 * g + 1] = tnow;
            }
        }
        // a wavefront of cells without sediment and layer-2 sums takes the
        // reduced form (same values)
        const bool bare = (o[3] == 0.0) && (o[4] == 0.0) && (o[5] == 0.0) && (ac[3] == 0.0) && (ac[4] == 0.0) &&
                          (ac[5] == 0.0);
        if (__all(bare))
            cell_update1(k, o, ac, dt, n, &hb);
        else
            cell_update(k, o, ac, dt, n, &hb);
#pragma unroll
        for (int f = 0; f < 6; f++) Wn[f * N + g] = n[f];
        if (go.out) {
            // the new h1, h2 replace this cell's staged ones (no other thread reads them now)
            sW[0][r + 1][c + 1] = n[0];
            sW[3][r + 1][c + 1] = n[3];
        }
        if (smask && (n[3] != 0.0 || n[4] != 0.0 || n[5] != 0.0)) *smask |= halo_mask_of<TX, TY>(gi - cx0, gj - cy0);
    }
    raise_health(hc, hb, hs);
    if (go.out) gauges_of_tile<TX, TY, NT>(go, &sW[0][0][0], &sW[3][0][0], &sW[6][0][0], CW, nx, i0, j0, 0x7fffffff);
    return Amax;
}

constexpr int FTX = 30, FTY = 14, FNT = 256;
// general tiles walked by a fixed grid (four workgroups per CU of an MI355X)
constexpr int GGRID = 1024;

// tile_step64() for a tile whose cells and halo carry no sediment (h2 = q2 =
// 0 in every staged cell): every edge takes edge1() and every cell
// cell_update1(), the arithmetic tile_step64() performs on such a tile, so
// the results are the same; the layer-2 planes are neither read nor kept in
// LDS (zeros are written).  A tile whose class entry *cls equals s belongs to
// the general kernel: the block returns false before its first LDS store.
// From step skip_from + 2 on, Wn holds the state that step s - 2 wrote; a
// tile that step took as sediment-free (entry other than s - 2) has zero
// layer-2 planes there already, and only the layer-1 planes are written.
// hc (optional): control block whose health word records the cell updates'
// health bits as those of step s.
// BUF: the state and depth are read and written with buffer instructions
// (plane offset in a scalar register, a 32-bit byte offset per cell), valid
// while state_buf_fits; otherwise 64-bit addresses.
template <int TX, int TY, int NT, bool BUF = false>
__device__ __forceinline__ bool tile_step_bare64(const Consts& k, const double* __restrict__ W,
                                                 double* __restrict__ Wn, const double* __restrict__ H, double dt,
                                                 double tnow, double* __restrict__ emax, int i0, int j0,
                                                 const int* cls, int s, int skip_from, GaugeOut go, double& Amax,
                                                 Ctl* hc = nullptr)
{
    using Idx = typename std::conditional<BUF, unsigned, size_t>::type;
    constexpr int CW = TX + 2, RW = TY + 2;
    __shared__ double sW[4][RW][CW];   // h1, q1x, q1y, H
    __shared__ double sA[4][TY][CW];   // dh1, dq1x, dq1y, wave-speed sum
    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const int tid = threadIdx.x;
    // FP64 plane p at cell g of the old / new state, depth at cell g
    [[maybe_unused]] __amdgpu_buffer_rsrc_t rW, rWn, rH;
    [[maybe_unused]] unsigned pb = 0u;
    if constexpr (BUF) {
        pb = (unsigned)N * 8u;
        rW = __builtin_amdgcn_make_buffer_rsrc((void*)W, (short)0, (int)(6u * pb), BUF_DWORD3);
        rWn = __builtin_amdgcn_make_buffer_rsrc((void*)Wn, (short)0, (int)(6u * pb), BUF_DWORD3);
        rH = __builtin_amdgcn_make_buffer_rsrc((void*)H, (short)0, (int)pb, BUF_DWORD3);
    }
    auto ldW = [&](int p, Idx g) -> double {
        if constexpr (BUF)
            return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rW, g * 8u, p * pb, 0));
        else
            return W[p * N + g];
    };
    auto ldH = [&](Idx g) -> double {
        if constexpr (BUF)
            return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rH, g * 8u, 0, 0));
        else
            return H[g];
    };
    auto stW = [&](int p, Idx g, double v) {
        if constexpr (BUF) {
            using raw64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(rW, 0u, 0u, 0));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(raw64, v), rWn, g * 8u, p * pb, 0);
        } else {
            Wn[p * N + g] = v;
        }
    };

    constexpr int ST_IT = (RW * CW + NT - 1) / NT;
    double v[ST_IT][4];
    bool in[ST_IT];
#pragma unroll
    for (int it = 0; it < ST_IT; it++) {
        const int e = tid + it * NT;
        const int r = e / CW, c = e % CW;
        const int gj = j0 + r - 1, gi = i0 + c - 1;
        in[it] = e < RW * CW && gj >= 0 && gj < ny && gi >= 0 && gi < nx;
        if (in[it]) {
            const Idx g = (Idx)gj * nx + gi;
            v[it][0] = ldW(0, g);
            v[it][1] = ldW(1, g);
            v[it][2] = ldW(2, g);
            v[it][3] = ldH(g);
        }
    }
    const int cv = *cls;
    if (cv == s) return false;
    const bool zero2 = !(s - 2 >= skip_from && cv != s - 2);
#pragma unroll
    for (int it = 0; it < ST_IT; it++) {
        if (!in[it]) continue;
        const int e = tid + it * NT;
        const int r = e / CW, c = e % CW;
#pragma unroll
        for (int f = 0; f < 4; f++) sW[f][r][c] = v[it][f];
    }
    for (int e = tid; e < TY * CW; e += NT) {
        const int r = e / CW, c = e % CW;
#pragma unroll
        for (int f = 0; f < 4; f++) sA[f][r][c] = 0.0;
    }
    __syncthreads();
    auto acc = [&](int r, int c, const double* F, double lam) {
#pragma unroll
        for (int f = 0; f < 3; f++) sA[f][r][c] -= F[f];
        sA[3][r][c] += lam;
    };

    // y-edge classes C1 then C2
#pragma unroll 1
    for (int par = 0; par < 2; par++) {
        const int nrow = (par == 0) ? (TY / 2 + 1) : (TY / 2);
        for (int e = tid; e < nrow * CW; e += NT) {
            const int c = e % CW, jl = par + 2 * (e / CW);
            const int gi = i0 + c - 1, je = j0 + jl;
            if (gi < 0 || gi >= nx || je > ny) continue;
            double W0[3], W1[3], H0, H1, a1h1 = 0.0, n1y;
            int r0;
            bool interior;
            if (je == 0 || je == ny) {
                r0 = (je == 0) ? jl : jl - 1;
                const double bd = (je == 0) ? k.borde_sup : k.borde_inf;
#pragma unroll
                for (int f = 0; f < 3; f++) { W0[f] = sW[f][r0 + 1][c]; W1[f] = W0[f]; }
                W1[2] *= bd;
                H0 = H1 = sW[3][r0 + 1][c];
                n1y = (je == 0) ? -1.0 : 1.0;
                interior = false;
            } else {
                r0 = jl - 1;
#pragma unroll
                for (int f = 0; f < 3; f++) { W0[f] = sW[f][r0 + 1][c]; W1[f] = sW[f][r0 + 2][c]; }
                H0 = sW[3][r0 + 1][c];
                H1 = sW[3][r0 + 2][c];
                n1y = 1.0;
                interior = true;
            }
            const bool has0 = (r0 >= 0 && r0 < TY), has1 = interior && (r0 + 1 < TY);
            const double a0h1 = has0 ? sA[0][r0][c] : 0.0;
            if (has1) a1h1 = sA[0][r0 + 1][c];
            EdgeFlux F;
            if (edge1<1>(k, W0, W1, H0, H1, n1y, k.dx, dt, a0h1, a1h1, interior, F)) {
                if (has0) acc(r0, c, F.Fm, F.lam);
                if (has1) acc(r0 + 1, c, F.Fp, F.lam);
            }
        }
        __syncthreads();
    }

    // x-edge classes C3 then C4
#pragma unroll 1
    for (int par = 0; par < 2; par++) {
        const int ncol = (par == 0) ? (TX / 2 + 1) : (TX / 2);
        for (int e = tid; e < ncol * TY; e += NT) {
            const int r = e / ncol, il = par + 2 * (e % ncol);
            const int gj = j0 + r, ie = i0 + il;
            if (gj >= ny || ie > nx) continue;
            double W0[3], W1[3], H0, H1, a1h1 = 0.0, n1x;
            int c0;
            bool interior;
            if (ie == 0 || ie == nx) {
                c0 = (ie == 0) ? il + 1 : il;
                const double bd = (ie == 0) ? k.borde_izq : k.borde_der;
#pragma unroll
                for (int f = 0; f < 3; f++) { W0[f] = sW[f][r + 1][c0]; W1[f] = W0[f]; }
                W1[1] *= bd;
                H0 = H1 = sW[3][r + 1][c0];
                n1x = (ie == 0) ? -1.0 : 1.0;
                interior = false;
            } else {
                c0 = il;
#pragma unroll
                for (int f = 0; f < 3; f++) { W0[f] = sW[f][r + 1][c0]; W1[f] = sW[f][r + 1][c0 + 1]; }
                H0 = sW[3][r + 1][c0];
                H1 = sW[3][r + 1][c0 + 1];
                n1x = 1.0;
                interior = true;
            }
            const double a0h1 = sA[0][r][c0];
            if (interior) a1h1 = sA[0][r][c0 + 1];
            EdgeFlux F;
            if (edge1<0>(k, W0, W1, H0, H1, n1x, k.dy, dt, a0h1, a1h1, interior, F)) {
                acc(r, c0, F.Fm, F.lam);
                if (interior) acc(r, c0 + 1, F.Fp, F.lam);
            }
        }
        __syncthreads();
    }

    // cell update (layer 2 stays zero)
    Amax = 0.0;
    unsigned hb = 0u;
    for (int e = tid; e < TX * TY; e += NT) {
        const int r = e / TX, c = e % TX;
        const int gj = j0 + r, gi = i0 + c;
        if (gj >= ny || gi >= nx) continue;
        const Idx g = (Idx)gj * nx + gi;
        double o[6] = {sW[0][r + 1][c + 1], sW[1][r + 1][c + 1], sW[2][r + 1][c + 1], 0.0, 0.0, 0.0};
        double ac[6] = {sA[0][r][c + 1], sA[1][r][c + 1], sA[2][r][c + 1], 0.0, 0.0, 0.0};
        double n[6];
        Amax = fmax(Amax, sA[3][r][c + 1]);
        if (emax) {
            const double vv = o[0] - sW[3][r + 1][c + 1];
            if (vv > emax[2 * g]) {
                emax[2 * g] = vv;
                emax[2 * g + 1] = tnow;
            }
        }
        cell_update1(k, o, ac, dt, n, &hb);
#pragma unroll
        for (int f = 0; f < 3; f++) stW(f, g, n[f]);
        if (zero2) {
#pragma unroll
            for (int f = 3; f < 6; f++) stW(f, g, n[f]);
        }
        if (go.out) sW[0][r + 1][c + 1] = n[0];   // the new h1 (read by no other thread now)
    }
    raise_health(hc, hb, s);
    if (go.out) gauges_of_tile<TX, TY, NT>(go, &sW[0][0][0], nullptr, &sW[3][0][0], CW, nx, i0, j0, 0x7fffffff);
    return true;
}

__device__ __forceinline__ int xcd_tile(int nb)
{
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int q = nb >> 3, rem = nb & 7;
    return xcd * q + (xcd < rem ? xcd : rem) + slot;
}

// Every tile through the general tile function (the classes kept), one
// block per tile in XCD-contiguous order (consecutive workgroups go
// round-robin over the 8 XCDs, so each XCD gets a contiguous band of tiles
// and shared halos stay in its L2).
template <int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_fused_all(Consts k, const double* __restrict__ W, double* __restrict__ Wn,
                                                  const double* __restrict__ H, StepArgs a, TileCls tc)
{
    if (!(a.ctl->t < a.t_end)) return;     // the run has reached its end time
    const int b = xcd_tile(tc.tiles_x * tc.tiles_y);
    const int tx = b % tc.tiles_x, ty = b / tc.tiles_x;
    const int s = a.ctl->step;
    const GaugeOut go{a.gauges, a.ng, (a.gseries && s < a.gcap) ? a.gseries + (size_t)s * a.ng : nullptr};
    int m = 0;
    const double Amax = tile_step64<TX, TY, NT>(k, W, Wn, H, step_dt(a.ctl, a.dt_seq), a.ctl->t, a.emax, tx * TX,
                                                ty * TY, go, &m, tx * TX, ty * TY, a.ctl, s);
    general_tile_end<NT>(Amax, m, tc, a.ctl, tx, ty, s + 1);
}

// Sediment-free tiles of a step, one block per tile (general ones leave).
// Forward recomputation of the discrete adjoint's reverse step: every class
// tile through tile_step64<RECOMP> (the step's own edge arithmetic), writing
// the partial sums and final flux sums instead of the new state.
template <int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_recompute(Consts k, const double* __restrict__ W,
                                                  const double* __restrict__ H, const double* __restrict__ dt_seq,
                                                  int step, double* __restrict__ P, double* __restrict__ ACC,
                                                  int tiles_x, int tiles_y)
{
    const int nb = tiles_x * tiles_y;
    int b;
    {
        const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
        const int q = nb >> 3, rem = nb & 7;
        b = xcd * q + (xcd < rem ? xcd : rem) + slot;
    }
    const int tx = b % tiles_x, ty = b / tiles_x;
    tile_step64<TX, TY, NT, true>(k, W, nullptr, H, dt_seq[step], 0.0, nullptr, tx * TX, ty * TY,
                                  GaugeOut{nullptr, 0, nullptr}, nullptr, 0, 0, nullptr, 0, P, ACC);
}

template <int TX, int TY, int NT, bool BUF>
__global__ void __launch_bounds__(NT) k_fused_bare(Consts k, const double* __restrict__ W, double* __restrict__ Wn,
                                                   const double* __restrict__ H, StepArgs a, TileCls tc)
{
    if (!(a.ctl->t < a.t_end)) return;
    const int nb = tc.tiles_x * tc.tiles_y;
    const int b = xcd_tile(nb);
    const int tx = b % tc.tiles_x, ty = b / tc.tiles_x;
    const int s = a.ctl->step;
    const GaugeOut go{a.gauges, a.ng, (a.gseries && s < a.gcap) ? a.gseries + (size_t)s * a.ng : nullptr};
    double Amax;
    if (!tile_step_bare64<TX, TY, NT, BUF>(k, W, Wn, H, step_dt(a.ctl, a.dt_seq), a.ctl->t, a.emax, tx * TX, ty * TY,
                                      tc.cls + (s & 1) * nb + b, s, a.wn_prev ? a.ctl->cls_from : 0x7fffffff, go,
                                      Amax, a.ctl))
        return;
    block_max_atomic<NT>(Amax, &a.ctl->maxA_next);
}

// General tiles of a step (a fixed grid walks the list), then the step's
// close by the last workgroup (src/GPU/ShallowWater.cu:431-442), as in the
// FP32C general launch: its count increment uses the value its atomicMax
// returned, so it is issued once that max has been performed.
template <int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_fused_gen(Consts k, const double* __restrict__ W, double* __restrict__ Wn,
                                                  const double* __restrict__ H, StepArgs a, TileCls tc)
{
    Ctl* ctl = a.ctl;
    // the step is the launch's own: a workgroup that starts after the last
    // participant closed it would otherwise read the next step's list
    const int s = a.step;
    if (!(ctl->t < a.t_end) || ctl->step != s) return;
    const int n = (int)ctl->gcount[s & 1];
    const unsigned part = (unsigned)max(min(n, (int)gridDim.x), 1);
    if (blockIdx.x >= part) return;
    const int nt = tc.tiles_x * tc.tiles_y;
    const double dt = step_dt(ctl, a.dt_seq);
    const double tnow = ctl->t;
    const GaugeOut go{a.gauges, a.ng, (a.gseries && s < a.gcap) ? a.gseries + (size_t)s * a.ng : nullptr};
    double Amax = 0.0;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int t = tc.list[(s & 1) * nt + i];
        const int tx = t % tc.tiles_x, ty = t / tc.tiles_x;
        int m = 0;
        Amax = fmax(Amax, tile_step64<TX, TY, NT>(k, W, Wn, H, dt, tnow, a.emax, tx * TX, ty * TY, go, &m, tx * TX,
                                                  ty * TY, ctl, s));
        m = block_or(m);   // also orders this tile's LDS reads before the next one's staging
        if (m && threadIdx.x < 64) mark_general(tc, ctl, tx, ty, s + 1, m);
    }
    __shared__ double red[NT / 64];
    double v = Amax;
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / 64; w++) v = fmax(v, red[w]);
        v = fabs(v);   // sign bit clear also for a NaN from a corrupt state
        const unsigned long long old = atomicMax(&ctl->maxA_next, (unsigned long long)__double_as_longlong(v));
        // bit 63 of a non-negative double is clear: the increment is 1
        const unsigned done = atomicAdd(&ctl->blocks_done, 1u + (unsigned)(old >> 63));
        if (done == part - 1) {
            const double maxA = __longlong_as_double((long long)atomicAdd(&ctl->maxA_next, 0ull));
            if (a.dt_hist && s < a.hcap) a.dt_hist[s] = dt;
            ctl->t = tnow + dt;
            ctl->dt_cur = dt_from_maxA(maxA, k.cfl, k.area);
            ctl->maxA_last = maxA;
            ctl->gcount[s & 1] = 0u;
            atomicExch(&ctl->maxA_next, 0ull);
            atomicExch(&ctl->blocks_done, 0u);
            ctl->step = s + 1;
        }
    }
}

// Close of a step run by k_fused_all (src/GPU/ShallowWater.cu:431-442).
__global__ void k_close(Consts k, StepArgs a)
{
    Ctl* ctl = a.ctl;
    if (threadIdx.x != 0 || !(ctl->t < a.t_end)) return;
    const int s = ctl->step;
    const double dt = step_dt(ctl, a.dt_seq);
    if (a.dt_hist && s < a.hcap) a.dt_hist[s] = dt;
    const double maxA = __longlong_as_double((long long)ctl->maxA_next);
    ctl->t = ctl->t + dt;
    ctl->dt_cur = dt_from_maxA(maxA, k.cfl, k.area);
    ctl->maxA_last = maxA;
    ctl->maxA_next = 0ull;
    ctl->gcount[s & 1] = 0u;
    ctl->step = s + 1;
}

// Class table of the state W at step ctl->step: one block per class tile.
__global__ void __launch_bounds__(256) k_classify(Consts k, const double* __restrict__ W, StepArgs a, TileCls tc)
{
    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const int t = blockIdx.x, tx = t % tc.tiles_x, ty = t / tc.tiles_x;
    int m = 0;
    for (int e = threadIdx.x; e < FTX * FTY; e += 256) {
        const int ly = e / FTX, lx = e % FTX;
        const int gj = ty * FTY + ly, gi = tx * FTX + lx;
        if (gj >= ny || gi >= nx) continue;
        const size_t g = (size_t)gj * nx + gi;
        if (W[3 * N + g] != 0.0 || W[4 * N + g] != 0.0 || W[5 * N + g] != 0.0) m |= halo_mask_of<FTX, FTY>(lx, ly);
    }
    m = block_or(m);
    if (m && threadIdx.x < 64) mark_general(tc, a.ctl, tx, ty, a.ctl->step, m);
    if (t == 0 && threadIdx.x == 0) a.ctl->cls_from = a.ctl->step;
}

// eta max of a fresh state: h1 + h2 - H with time 0 (src/GPU/Problema.cxx:541-546)
__global__ void __launch_bounds__(256) k_init_eta_max(const double* __restrict__ W, const double* __restrict__ H,
                                                      double* __restrict__ emax, size_t N)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
        emax[2 * i] = W[i] + W[3 * N + i] - H[i];
        emax[2 * i + 1] = 0.0;
    }
}

}  // namespace f64

// ------------------------------------------------------------------ launchers
hipError_t init_eta_max(const double* W, const double* H, double* emax, size_t N, hipStream_t s)
{
    const size_t b = (N + 255) / 256;
    hipLaunchKernelGGL(f64::k_init_eta_max, dim3((unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192)), dim3(256), 0, s, W,
                       H, emax, N);
    return hipGetLastError();
}

hipError_t f64_initial_dt(const Consts& k, const double* W, Ctl* ctl, hipStream_t s)
{
    const size_t N = (size_t)k.nx * k.ny;
    hipLaunchKernelGGL(f64::k_dt0, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, k, W, ctl);
    hipLaunchKernelGGL(f64::k_dt0_finalize, dim3(1), dim3(64), 0, s, k, ctl);
    return hipGetLastError();
}

hipError_t f64_step_fourpass(const Consts& k, const double* W, double* Wn, const double* H, double* acc,
                             const StepArgs& a, hipStream_t s)
{
    const size_t N = (size_t)k.nx * k.ny;
    for (int cls = 0; cls < 4; cls++) {
        long ne;
        if (cls < 2) ne = (long)((k.ny - cls) / 2 + 1) * k.nx;
        else ne = (long)((k.nx - (cls - 2)) / 2 + 1) * k.ny;
        hipLaunchKernelGGL(f64::k_edges, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, k, W, H, acc,
                           a.ctl, a.dt_seq, cls);
    }
    hipLaunchKernelGGL(f64::k_update, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, k, W, Wn, H, acc, a);
    hipLaunchKernelGGL(f64::k_finalize, dim3(1), dim3(64), 0, s, k, Wn, H, a);
    return hipGetLastError();
}

int f64_tiles(const Consts& k)
{
    return ((k.nx + f64::FTX - 1) / f64::FTX) * ((k.ny + f64::FTY - 1) / f64::FTY);
}

hipError_t f64_classify(const Consts& k, const double* W, const StepArgs& a, hipStream_t s)
{
    const int nt = f64_tiles(k);
    const TileCls tc{a.tcls, a.tlist, (k.nx + f64::FTX - 1) / f64::FTX, (k.ny + f64::FTY - 1) / f64::FTY};
    hipError_t e = clear_classes(a.tcls, 2 * nt, a.ctl, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(f64::k_classify, dim3((unsigned)nt), dim3(256), 0, s, k, W, a, tc);
    return hipGetLastError();
}

// Split (a.split): the sediment-free launch, then the general tiles and the
// step's close; otherwise every tile general, then the close.
hipError_t f64_recompute(const Consts& k, const double* W, const double* H, const double* dt_seq, int step, double* P,
                         double* ACC, hipStream_t s)
{
    const int tiles_x = (k.nx + f64::FTX - 1) / f64::FTX;
    const int tiles_y = (k.ny + f64::FTY - 1) / f64::FTY;
    hipLaunchKernelGGL((f64::k_recompute<f64::FTX, f64::FTY, f64::FNT>), dim3((unsigned)(tiles_x * tiles_y)),
                       dim3(f64::FNT), 0, s, k, W, H, dt_seq, step, P, ACC, tiles_x, tiles_y);
    return hipGetLastError();
}

hipError_t f64_step_fused(const Consts& k, const double* W, double* Wn, const double* H, const StepArgs& a,
                          hipStream_t s)
{
    const int tiles_x = (k.nx + f64::FTX - 1) / f64::FTX;
    const int tiles_y = (k.ny + f64::FTY - 1) / f64::FTY;
    const int nt = tiles_x * tiles_y;
    const TileCls tc{a.tcls, a.tlist, tiles_x, tiles_y};
    hipError_t e = a.ev0 ? hipEventRecord(a.ev0, s) : hipSuccess;
    if (e != hipSuccess) return e;
    if (a.split) {
        if (state_buf_fits((size_t)k.nx * k.ny))
            hipLaunchKernelGGL((f64::k_fused_bare<f64::FTX, f64::FTY, f64::FNT, true>), dim3((unsigned)nt),
                               dim3(f64::FNT), 0, s, k, W, Wn, H, a, tc);
        else
            hipLaunchKernelGGL((f64::k_fused_bare<f64::FTX, f64::FTY, f64::FNT, false>), dim3((unsigned)nt),
                               dim3(f64::FNT), 0, s, k, W, Wn, H, a, tc);
        hipLaunchKernelGGL((f64::k_fused_gen<f64::FTX, f64::FTY, f64::FNT>), dim3(f64::GGRID), dim3(f64::FNT), 0, s,
                           k, W, Wn, H, a, tc);
        if (a.ev1 && (e = hipEventRecord(a.ev1, s)) != hipSuccess) return e;
    } else {
        hipLaunchKernelGGL((f64::k_fused_all<f64::FTX, f64::FTY, f64::FNT>), dim3((unsigned)nt), dim3(f64::FNT), 0,
                           s, k, W, Wn, H, a, tc);
        if (a.ev1 && (e = hipEventRecord(a.ev1, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(f64::k_close, dim3(1), dim3(64), 0, s, k, a);
    }
    return hipGetLastError();
}

}  // namespace lhysea
