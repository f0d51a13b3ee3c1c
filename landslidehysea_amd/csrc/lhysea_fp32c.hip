// FP32-compensated step kernels of the LandSlide-HySEA solver on MI355X.
//
// HBM layout of the state (one ping-pong buffer): 12 FP32 planes of N cells,
// values h1,q1x,q1y,h2,q2x,q2y then their compensations; depth H is one FP64
// plane.  Algorithmic traffic per cell-update: 56 B read (24 values + 24
// compensations + 8 depth) and 48 B written = 104 B.
//
// The fused kernel follows the FP64 one (lhysea_fp64.hip): an even-aligned
// TX x TY tile with a 1-cell halo staged in LDS, the four edge classes in the
// reference's order with LDS accumulators, the compensated cell update, and
// a wavefront-shuffle max of the wave-speed sums for the lagged dt.
#include <algorithm>
#include <type_traits>
#include "lhysea_internal.hpp"
#include "lhysea_fp32.hpp"
#include "lhysea_tiles.hpp"

namespace lhysea {
namespace f32 {

__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

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

// FP32 constants of a step with time step dt: (float)dt, 1 / (dt + EPS) of
// the positivity limiter and dt / area.  FP64 divisions: a kernel derives
// them once per workgroup where it can (tile_step_bare).
struct DtF {
    float dtf, idt_eps, bq;
};

__device__ __forceinline__ DtF dt_floats(const Consts& k, double dt)
{
    DtF d;
    d.dtf = (float)dt;
    d.idt_eps = (float)(1.0 / ((double)d.dtf + (double)LH_EPSF));
    d.bq = (float)(dt / k.area);
    return d;
}

__device__ __forceinline__ K32 make_k32(const Consts& k, const DtF& d)
{
    const ConstsF32& f = k.f;
    K32 c;
    c.eps = f.eps;
    c.eps4 = f.eps4;
    c.g = f.g;
    c.gp = f.gp;
    c.r = f.r;
    c.mu = f.mu;
    c.peso = f.peso;
    c.dtf = d.dtf;
    c.area = f.area;
    c.len_x = f.len_x;
    c.len_y = f.len_y;
    c.igp = f.igp;
    c.ten_eps = f.ten_eps;
    c.idt_eps = d.idt_eps;
    c.bq = d.bq;
    c.pouliquen = k.friction_law == 1;
    c.tan1 = f.tan1;
    c.tan2 = f.tan2;
    c.tan3 = f.tan3;
    c.tan4 = f.tan4;
    c.LoverH = f.LoverH;
    return c;
}

__device__ __forceinline__ KCell make_kcell(const Consts& k, const K32& kc)
{
    const ConstsF32& f = k.f;
    KCell kk;
    kk.eps = kc.eps;
    kk.eps4 = kc.eps4;
    kk.r = kc.r;
    kk.g = kc.g;
    kk.mu = kc.mu;
    kk.mfc = f.mfc;
    kk.mf0 = f.mf0;
    kk.mfs = f.mfs;
    kk.vmax1 = f.vmax1;
    kk.vmax2 = f.vmax2;
    kk.dtf = kc.dtf;
    kk.valf = kc.bq;
    kk.ieps = f.ieps;
    kk.ten_eps = kc.ten_eps;
    kk.pouliquen = kc.pouliquen;
    kk.tan1 = kc.tan1;
    kk.tan2 = kc.tan2;
    kk.tan3 = kc.tan3;
    kk.tan4 = kc.tan4;
    kk.LoverH = kc.LoverH;
    return kk;
}

// ---------------------------------------------------------------- initial dt
__global__ void __launch_bounds__(256) k_dt0(Consts k, const float* __restrict__ W, Ctl* ctl)
{
    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const long cidx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const K32 c = make_k32(k, dt_floats(k, 0.0));
    double A = 0.0;
    if (cidx < (long)N) {
        const int i = (int)(cidx % nx), j = (int)(cidx / nx);
        float Wc[6], Wo[6], v[4];
        for (int f = 0; f < 6; f++) Wc[f] = W[f * N + cidx];
        for (int s = 0; s < 2; s++) {
            const int je = j + s;
            if (je == 0 || je == ny) {
                const float bd = (float)((je == 0) ? k.borde_sup : k.borde_inf);
                for (int f = 0; f < 6; f++) Wo[f] = Wc[f];
                Wo[2] *= bd; Wo[5] *= bd;
                v[je & 1] = edge_dt0(c, Wc, Wo, 0.0f, (je == 0) ? -1.0f : 1.0f, c.len_y);
            } else {
                const long o = (s == 0) ? cidx - nx : cidx + nx;
                for (int f = 0; f < 6; f++) Wo[f] = W[f * N + o];
                v[je & 1] = (s == 0) ? edge_dt0(c, Wo, Wc, 0.0f, 1.0f, c.len_y) : edge_dt0(c, Wc, Wo, 0.0f, 1.0f, c.len_y);
            }
        }
        for (int s = 0; s < 2; s++) {
            const int ie = i + s;
            if (ie == 0 || ie == nx) {
                const float bd = (float)((ie == 0) ? k.borde_izq : k.borde_der);
                for (int f = 0; f < 6; f++) Wo[f] = Wc[f];
                Wo[1] *= bd; Wo[4] *= bd;
                v[2 + (ie & 1)] = edge_dt0(c, Wc, Wo, (ie == 0) ? -1.0f : 1.0f, 0.0f, c.len_x);
            } else {
                const long o = (s == 0) ? cidx - 1 : cidx + 1;
                for (int f = 0; f < 6; f++) Wo[f] = W[f * N + o];
                v[2 + (ie & 1)] = (s == 0) ? edge_dt0(c, Wo, Wc, 1.0f, 0.0f, c.len_x) : edge_dt0(c, Wc, Wo, 1.0f, 0.0f, c.len_x);
            }
        }
        A = (double)(((v[0] + v[1]) + v[2]) + v[3]);
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

// Both cells of an edge without sediment: h2 exactly 0 (FP64) and q2 = 0.
// When every edge of a wavefront is such an edge it takes edge1(), which
// produces no layer-2 fluxes; the layer-2 accumulators then stay untouched.
__device__ __forceinline__ bool no_sediment(const float* W0, const float* W1, double h20, double h21)
{
    return (h20 == 0.0) && (h21 == 0.0) && (W0[4] == 0.0f) && (W0[5] == 0.0f) && (W1[4] == 0.0f) &&
           (W1[5] == 0.0f);
}

// sA[f][r][c] -= F[f] for the first NF fields (h1,q1x,q1y[,h2,q2x,q2y]), and
// the wave-speed sum += lam
template <int NF, int TY, int CW>
__device__ __forceinline__ void accumulate(float (&sA)[7][TY][CW], int r, int c, const float* F, float lam)
{
#pragma unroll
    for (int f = 0; f < NF; f++) sA[f][r][c] -= F[f];
    sA[6][r][c] += lam;
}

constexpr int FTX = 62, FTY = 14, FNT = 512;
// Scalar registers of the sediment-free kernels: a wave that needs more than
// 80 (74 addressable plus the reserved ones) is allocated 96 on gfx950, and
// at 96 a CU no longer holds four 512-thread workgroups (measured: 82% of the
// resident wave-cycles, the 4096^2 step 289 -> 331 us).  The buffer-addressed
// kernels fit in 80; the 64-bit addressed ones are held there by this cap,
// which spills a few scalars to vector lanes.
#define LH_BARE_SGPR_CAP __attribute__((amdgpu_num_sgpr(82)))
// general sub-tiles: width, threads, and the fixed grid walking them (four
// workgroups per CU of an MI355X)
constexpr int GSTX = 16, GNT = 256, GGRID = 1024;
__device__ __forceinline__ int halo_mask(int lx, int ly) { return halo_mask_of<FTX, FTY>(lx, ly); }


// ----------------------------------------------------------------- fused step
// One even-aligned TX x TY tile of one step (origin i0, j0): stage the tile
// and its 1-cell halo, run the four edge classes in the reference's order,
// write the compensated update of the tile's own cells.  Returns this
// thread's max of the wave-speed sums A_i of the cells it updated.
// LDS: sH (depth), sh1/sh2 (full-precision h as FP64), sq (q hi values),
// sA (FP32 accumulators: dh1,dq1x,dq1y,dh2,dq2x,dq2y,A).
// Every cell receives its class-C1..C4 contributions in that order whatever
// the tile shape, so the result does not depend on TX, TY or NT.
// Only cells with column < xend are written.  go: gauges of the new state.
// smask (optional): OR of halo_mask() over this thread's cells that keep
// sediment (h2 or q2 nonzero) in the new state, with the cell's position in
// the FTX x FTY class tile at (cx0, cy0).
// hc (optional): control block whose health word records the cell updates'
// health bits (HEALTH_*) as those of step hs.
// RECOMP: instead of updating the state, write the partial h1/h2 sums that
// classes 2..4 read (P, [3][2][N]) and the final flux sums (ACC, [6][N]) of
// the tile's cells: the forward recomputation of the discrete adjoint.
template <int TX, int TY, int NT, bool RECOMP = false>
__device__ __forceinline__ double tile_step(const Consts& k, const float* __restrict__ W, float* __restrict__ Wn,
                                            const double* __restrict__ H, double dt, double tnow,
                                            double* __restrict__ emax, int i0, int j0,
                                            float* __restrict__ P = nullptr, float* __restrict__ ACC = nullptr,
                                            int xend = 0x7fffffff, GaugeOut go = GaugeOut{nullptr, 0, nullptr},
                                            int* smask = nullptr, int cx0 = 0, int cy0 = 0, Ctl* hc = nullptr,
                                            int hs = 0)
{
    static_assert(TX % 2 == 0 && TY % 2 == 0, "tile origins must stay even");
    // LDS rows are padded to an odd length (CW + 1 words) so that the
    // stride-2 column access of the x-edge classes spreads over all banks
    constexpr int CW = TX + 2, RW = TY + 2, CWP = CW + 1;
    __shared__ double sH[RW][CWP];
    __shared__ double sh1[RW][CWP];
    __shared__ double sh2[RW][CWP];
    __shared__ float sq[4][RW][CWP];
    __shared__ float sA[7][TY][CWP];

    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const K32 kc = make_k32(k, dt_floats(k, dt));
    const int tid = threadIdx.x;

    // HBM loads are issued early and all at once: the q1 compensations the
    // cell update needs, then every staging element of this thread, before the
    // first LDS store waits on them
    constexpr int CU_IT = (TX * TY + NT - 1) / NT, ST_IT = (RW * CW + NT - 1) / NT;
    float lo1x[CU_IT], lo1y[CU_IT];
    if (!RECOMP) {
#pragma unroll
        for (int it = 0; it < CU_IT; it++) {
            const int e = tid + it * NT;
            const int r = e / TX, c = e % TX;
            lo1x[it] = lo1y[it] = 0.0f;
            if (e < TX * TY && j0 + r < ny && i0 + c < nx) {
                const size_t g = (size_t)(j0 + r) * nx + (i0 + c);
                lo1x[it] = W[7 * N + g];
                lo1y[it] = W[8 * N + g];
            }
        }
    }
    {
        double vH[ST_IT], vh1[ST_IT], vh2[ST_IT];
        float vq[ST_IT][4];
        bool in[ST_IT];
#pragma unroll
        for (int it = 0; it < ST_IT; it++) {
            const int e = tid + it * NT;
            const int r = e / CW, c = e % CW;
            const int gj = j0 + r - 1, gi = i0 + c - 1;
            in[it] = e < RW * CW && gj >= 0 && gj < ny && gi >= 0 && gi < nx;
            if (in[it]) {
                const size_t g = (size_t)gj * nx + gi;
                vH[it] = H[g];
                vh1[it] = (double)W[g] + (double)W[6 * N + g];
                vh2[it] = (double)W[3 * N + g] + (double)W[9 * N + g];
                vq[it][0] = W[N + g];
                vq[it][1] = W[2 * N + g];
                vq[it][2] = W[4 * N + g];
                vq[it][3] = W[5 * N + g];
            }
        }
#pragma unroll
        for (int it = 0; it < ST_IT; it++) {
            if (!in[it]) continue;
            const int e = tid + it * NT;
            const int r = e / CW, c = e % CW;
            sH[r][c] = vH[it];
            sh1[r][c] = vh1[it];
            sh2[r][c] = vh2[it];
#pragma unroll
            for (int f = 0; f < 4; f++) sq[f][r][c] = vq[it][f];
        }
    }
    for (int e = tid; e < TY * CW; e += NT) {
        const int r = e / CW, c = e % CW;
#pragma unroll
        for (int f = 0; f < 7; f++) sA[f][r][c] = 0.0f;
    }
    __syncthreads();

    // RECOMP: the tile's own h1/h2 partial sums after class cls (read by class cls + 1)
    auto store_partials = [&](int cls) {
        float* Pc = P + (size_t)cls * 2 * N;
        for (int e = tid; e < TX * TY; e += NT) {
            const int r = e / TX, c = e % TX;
            const int gj = j0 + r, gi = i0 + c;
            if (gj >= ny || gi >= nx) continue;
            const size_t g = (size_t)gj * nx + gi;
            Pc[g] = sA[0][r][c + 1];
            Pc[N + g] = sA[3][r][c + 1];
        }
        __syncthreads();
    };

    auto loadW = [&](int r, int c, float* w) {
        w[0] = (float)sh1[r][c];
        w[1] = sq[0][r][c];
        w[2] = sq[1][r][c];
        w[3] = (float)sh2[r][c];
        w[4] = sq[2][r][c];
        w[5] = sq[3][r][c];
    };

    // y-edge classes C1 then C2
#pragma unroll 1
    for (int par = 0; par < 2; par++) {
        const int nrow = (par == 0) ? (TY / 2 + 1) : (TY / 2);
        for (int e = tid; e < nrow * CW; e += NT) {
            const int c = e % CW, jl = par + 2 * (e / CW);
            const int gi = i0 + c - 1, je = j0 + jl;
            if (gi < 0 || gi >= nx || je > ny) continue;
            float W0[6], W1[6];
            double h10, h20, h11, h21, H0, H1;
            float a1h1 = 0.0f, a1h2 = 0.0f, n1y;
            int r0;
            bool interior;
            if (je == 0 || je == ny) {
                r0 = (je == 0) ? jl : jl - 1;
                const float bd = (float)((je == 0) ? k.borde_sup : k.borde_inf);
                loadW(r0 + 1, c, W0);
#pragma unroll
                for (int f = 0; f < 6; f++) W1[f] = W0[f];
                W1[2] *= bd; W1[5] *= bd;
                h10 = h11 = sh1[r0 + 1][c];
                h20 = h21 = sh2[r0 + 1][c];
                H0 = H1 = sH[r0 + 1][c];
                n1y = (je == 0) ? -1.0f : 1.0f;
                interior = false;
            } else {
                r0 = jl - 1;
                loadW(r0 + 1, c, W0);
                loadW(r0 + 2, c, W1);
                h10 = sh1[r0 + 1][c]; h20 = sh2[r0 + 1][c];
                h11 = sh1[r0 + 2][c]; h21 = sh2[r0 + 2][c];
                H0 = sH[r0 + 1][c];
                H1 = sH[r0 + 2][c];
                n1y = 1.0f;
                interior = true;
            }
            const bool has0 = (r0 >= 0 && r0 < TY), has1 = interior && (r0 + 1 < TY);
            const float a0h1 = has0 ? sA[0][r0][c] : 0.0f;
            if (has1) a1h1 = sA[0][r0 + 1][c];
            EdgeFlux F;
            if (__all(no_sediment(W0, W1, h20, h21))) {
                if (edge1<1>(kc, W0, W1, h10, h11, H0, H1, n1y, kc.len_y, a0h1, a1h1, interior, F)) {
                    if (has0) accumulate<3>(sA, r0, c, F.Fm, F.lam);
                    if (has1) accumulate<3>(sA, r0 + 1, c, F.Fp, F.lam);
                }
            } else {
                const float a0h2 = has0 ? sA[3][r0][c] : 0.0f;
                if (has1) a1h2 = sA[3][r0 + 1][c];
                if (edge<1>(kc, W0, W1, h10, h20, h11, h21, H0, H1, n1y, kc.len_y, a0h1, a0h2, a1h1, a1h2,
                            interior, F)) {
                    if (has0) accumulate<6>(sA, r0, c, F.Fm, F.lam);
                    if (has1) accumulate<6>(sA, r0 + 1, c, F.Fp, F.lam);
                }
            }
        }
        __syncthreads();
        if (RECOMP) store_partials(par);
    }

    // x-edge classes C3 then C4.  Work items go in 16-lane segments of 2 rows x
    // 8 edge columns, two column-adjacent segments per 32 lanes: with the odd
    // row length a 32-lane FP32 access and a 16-lane FP64 access then touch
    // distinct banks.
    static_assert(TY % 2 == 0, "row pairs");
    constexpr int NB = (TX / 2 + 1 + 7) / 8;   // blocks of 8 edge columns
#pragma unroll 1
    for (int par = 0; par < 2; par++) {
        const int ncol = (par == 0) ? (TX / 2 + 1) : (TX / 2);
        for (int e = tid; e < (TY / 2) * NB * 16; e += NT) {
            const int seg = e >> 4, w = e & 15;
            const int r = 2 * (seg / NB) + (w >> 3), j = 8 * (seg % NB) + (w & 7);
            if (j >= ncol) continue;
            const int il = par + 2 * j;
            const int gj = j0 + r, ie = i0 + il;
            if (gj >= ny || ie > nx) continue;
            float W0[6], W1[6];
            double h10, h20, h11, h21, H0, H1;
            float a1h1 = 0.0f, a1h2 = 0.0f, n1x;
            int c0;
            bool interior;
            if (ie == 0 || ie == nx) {
                c0 = (ie == 0) ? il + 1 : il;
                const float bd = (float)((ie == 0) ? k.borde_izq : k.borde_der);
                loadW(r + 1, c0, W0);
#pragma unroll
                for (int f = 0; f < 6; f++) W1[f] = W0[f];
                W1[1] *= bd; W1[4] *= bd;
                h10 = h11 = sh1[r + 1][c0];
                h20 = h21 = sh2[r + 1][c0];
                H0 = H1 = sH[r + 1][c0];
                n1x = (ie == 0) ? -1.0f : 1.0f;
                interior = false;
            } else {
                c0 = il;
                loadW(r + 1, c0, W0);
                loadW(r + 1, c0 + 1, W1);
                h10 = sh1[r + 1][c0]; h20 = sh2[r + 1][c0];
                h11 = sh1[r + 1][c0 + 1]; h21 = sh2[r + 1][c0 + 1];
                H0 = sH[r + 1][c0];
                H1 = sH[r + 1][c0 + 1];
                n1x = 1.0f;
                interior = true;
            }
            const float a0h1 = sA[0][r][c0];
            if (interior) a1h1 = sA[0][r][c0 + 1];
            EdgeFlux F;
            if (__all(no_sediment(W0, W1, h20, h21))) {
                if (edge1<0>(kc, W0, W1, h10, h11, H0, H1, n1x, kc.len_x, a0h1, a1h1, interior, F)) {
                    accumulate<3>(sA, r, c0, F.Fm, F.lam);
                    if (interior) accumulate<3>(sA, r, c0 + 1, F.Fp, F.lam);
                }
            } else {
                const float a0h2 = sA[3][r][c0];
                if (interior) a1h2 = sA[3][r][c0 + 1];
                if (edge<0>(kc, W0, W1, h10, h20, h11, h21, H0, H1, n1x, kc.len_x, a0h1, a0h2, a1h1, a1h2,
                            interior, F)) {
                    accumulate<6>(sA, r, c0, F.Fm, F.lam);
                    if (interior) accumulate<6>(sA, r, c0 + 1, F.Fp, F.lam);
                }
            }
        }
        __syncthreads();
        if (RECOMP && par == 0) store_partials(2);
    }

    if (RECOMP) {
#pragma unroll 1
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

    // compensated cell update of the tile's own cells
    const KCell kk = make_kcell(k, kc);
    double Amax = 0.0;
#pragma unroll
    for (int it = 0; it < CU_IT; it++) {
        const int e = tid + it * NT;
        if (e >= TX * TY) continue;
        const int r = e / TX, c = e % TX;
        const int gj = j0 + r, gi = i0 + c;
        if (gj >= ny || gi >= nx || gi >= xend) continue;
        unsigned hb = 0u;   // raised per cell: nothing stays live across the loop
        const size_t g = (size_t)gj * nx + gi;
        Pair o[6];
        const double h1d = sh1[r + 1][c + 1], h2d = sh2[r + 1][c + 1];
        o[0].hi = (float)h1d; o[0].lo = (float)(h1d - (double)o[0].hi);
        o[1] = Pair{sq[0][r + 1][c + 1], lo1x[it]};
        o[2] = Pair{sq[1][r + 1][c + 1], lo1y[it]};
        float ac[6];
#pragma unroll
        for (int f = 0; f < 6; f++) ac[f] = sA[f][r][c + 1];
        Amax = fmax(Amax, (double)sA[6][r][c + 1]);
        if (emax) {
            const double v = h1d - sH[r + 1][c + 1];
            if (v > emax[2 * g]) {
                emax[2 * g] = v;
                emax[2 * g + 1] = tnow;
            }
        }
        // no sediment and no layer-2 flux sums (a pair whose value is 0 has a
        // zero compensation, so the q2 compensations need not be read)
        const bool bare = (h2d == 0.0) && (sq[2][r + 1][c + 1] == 0.0f) && (sq[3][r + 1][c + 1] == 0.0f) &&
                          (ac[3] == 0.0f) && (ac[4] == 0.0f) && (ac[5] == 0.0f);
        if (__all(bare)) {
            cell_update1(kk, o, ac, &hb);
#pragma unroll
            for (int f = 0; f < 3; f++) {
                Wn[f * N + g] = o[f].hi;
                Wn[(6 + f) * N + g] = o[f].lo;
                Wn[(3 + f) * N + g] = 0.0f;
                Wn[(9 + f) * N + g] = 0.0f;
            }
            // the new h1 (h2 stays 0) replaces this cell's staged one, which
            // no other thread reads any more
            if (go.out) sh1[r + 1][c + 1] = (double)o[0].hi + (double)o[0].lo;
        } else {
            o[3].hi = (float)h2d; o[3].lo = (float)(h2d - (double)o[3].hi);
            o[4] = Pair{sq[2][r + 1][c + 1], W[10 * N + g]};
            o[5] = Pair{sq[3][r + 1][c + 1], W[11 * N + g]};
            cell_update(kk, o, ac, &hb);
#pragma unroll
            for (int f = 0; f < 6; f++) {
                Wn[f * N + g] = o[f].hi;
                Wn[(6 + f) * N + g] = o[f].lo;
            }
            if (go.out) {
                sh1[r + 1][c + 1] = (double)o[0].hi + (double)o[0].lo;
                sh2[r + 1][c + 1] = (double)o[3].hi + (double)o[3].lo;
            }
            if (smask && (o[3].hi != 0.0f || o[4].hi != 0.0f || o[5].hi != 0.0f))
                *smask |= halo_mask(gi - cx0, gj - cy0);
        }
        raise_health(hc, hb, hs);
    }
    if (go.out) gauges_of_tile<TX, TY, NT>(go, &sh1[0][0], &sh2[0][0], &sH[0][0], CWP, nx, i0, j0, xend);
    return Amax;
}

// ------------------------------------------------------- sediment-free tile
// tile_step() for a tile whose cells and halo carry no sediment: h2 = 0
// (value and compensation) and q2 = 0 in every staged cell.  There every
// edge takes edge1() and every cell cell_update1(), exactly the arithmetic
// tile_step() performs on such a tile, so the results are the same; the
// layer-2 planes are neither read nor kept in LDS (zeros are written), which
// halves the LDS footprint and the registers the general two-layer path needs.
// A tile whose class entry *cls equals s belongs to the general kernel: the
// block returns false before its first LDS store (its loads were issued).
// From step skip_from + 2 on, Wn holds the state that step s - 2 wrote; a
// tile that step took as sediment-free (its entry, stamped for s - 2 and not
// raised to s, differs from s - 2) has zero layer-2 planes there already, so
// only the three layer-1 planes and their compensations are written.  hc
// (optional): control block whose health word records the cell updates'
// health bits as those of step s.
// BUF: the state and depth are read and written with buffer instructions
// (plane offset in a scalar register, a 32-bit byte offset per cell), so that
// no vector instruction forms an address; valid while the 12 FP32 planes stay
// below 4 GiB (state_buf_fits).  Otherwise 64-bit addresses.
template <int TX, int TY, int NT, bool BUF = false>
__device__ __forceinline__ bool tile_step_bare(const Consts& k, const float* __restrict__ W, float* __restrict__ Wn,
                                               const double* __restrict__ H, double dt, double tnow,
                                               double* __restrict__ emax, int i0, int j0, const int* cls, int s,
                                               int skip_from, GaugeOut go, double& Amax, Ctl* hc = nullptr)
{
    static_assert(TX % 2 == 0 && TY % 2 == 0, "tile origins must stay even");
    using Idx = typename std::conditional<BUF, unsigned, size_t>::type;
    constexpr int CW = TX + 2, RW = TY + 2, CWP = CW + 1;
    __shared__ double sH[RW][CWP];
    __shared__ double sh1[RW][CWP];
    __shared__ float sq[2][RW][CWP];
    __shared__ float sA[4][TY][CWP];   // dh1, dq1x, dq1y, wave-speed sum

    __shared__ float sdt[2];           // the step's 1 / (dt + EPS) and dt / area

    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const int tid = threadIdx.x;
    // FP32 plane p at cell g of the old / new state, depth at cell g
    [[maybe_unused]] __amdgpu_buffer_rsrc_t rW, rWn, rH;
    [[maybe_unused]] unsigned pb = 0u;
    if constexpr (BUF) {
        pb = (unsigned)N * 4u;
        rW = __builtin_amdgcn_make_buffer_rsrc((void*)W, (short)0, (int)(12u * pb), BUF_DWORD3);
        rWn = __builtin_amdgcn_make_buffer_rsrc((void*)Wn, (short)0, (int)(12u * pb), BUF_DWORD3);
        rH = __builtin_amdgcn_make_buffer_rsrc((void*)H, (short)0, (int)(2u * pb), BUF_DWORD3);
    }
    auto ldW = [&](int p, Idx g) -> float {
        if constexpr (BUF)
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rW, g * 4u, p * pb, 0));
        else
            return W[p * N + g];
    };
    auto ldH = [&](Idx g) -> double {
        if constexpr (BUF)
            return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rH, g * 8u, 0, 0));
        else
            return H[g];
    };
    auto stW = [&](int p, Idx g, float v) {
        if constexpr (BUF)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rWn, g * 4u, p * pb, 0);
        else
            Wn[p * N + g] = v;
    };
    if (tid < 64) {                    // one wave derives them for the workgroup
        const DtF d = dt_floats(k, dt);
        if (tid == 0) {
            sdt[0] = d.idt_eps;
            sdt[1] = d.bq;
        }
    }

    constexpr int CU_IT = (TX * TY + NT - 1) / NT, ST_IT = (RW * CW + NT - 1) / NT;
    float lo1x[CU_IT], lo1y[CU_IT];
    bool zero2;
#pragma unroll
    for (int it = 0; it < CU_IT; it++) {
        const int e = tid + it * NT;
        const int r = e / TX, c = e % TX;
        lo1x[it] = lo1y[it] = 0.0f;
        if (e < TX * TY && j0 + r < ny && i0 + c < nx) {
            const Idx g = (Idx)(j0 + r) * nx + (i0 + c);
            lo1x[it] = ldW(7, g);
            lo1y[it] = ldW(8, g);
        }
    }
    {
        double vH[ST_IT], vh1[ST_IT];
        float vq[ST_IT][2];
        bool in[ST_IT];
#pragma unroll
        for (int it = 0; it < ST_IT; it++) {
            const int e = tid + it * NT;
            const int r = e / CW, c = e % CW;
            const int gj = j0 + r - 1, gi = i0 + c - 1;
            in[it] = e < RW * CW && gj >= 0 && gj < ny && gi >= 0 && gi < nx;
            if (in[it]) {
                const Idx g = (Idx)gj * nx + gi;
                vH[it] = ldH(g);
                vh1[it] = (double)ldW(0, g) + (double)ldW(6, g);
                vq[it][0] = ldW(1, g);
                vq[it][1] = ldW(2, g);
            }
        }
        const int v = *cls;
        if (v == s) return false;
        zero2 = !(s - 2 >= skip_from && v != s - 2);
#pragma unroll
        for (int it = 0; it < ST_IT; it++) {
            if (!in[it]) continue;
            const int e = tid + it * NT;
            const int r = e / CW, c = e % CW;
            sH[r][c] = vH[it];
            sh1[r][c] = vh1[it];
            sq[0][r][c] = vq[it][0];
            sq[1][r][c] = vq[it][1];
        }
    }
    for (int e = tid; e < TY * CW; e += NT) {
        const int r = e / CW, c = e % CW;
#pragma unroll
        for (int f = 0; f < 4; f++) sA[f][r][c] = 0.0f;
    }
    __syncthreads();
    const K32 kc = make_k32(k, DtF{(float)dt, sdt[0], sdt[1]});

    auto loadW = [&](int r, int c, float* w) {
        w[0] = (float)sh1[r][c];
        w[1] = sq[0][r][c];
        w[2] = sq[1][r][c];
    };
    auto acc = [&](int r, int c, const float* F, float lam) {
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
            float W0[3], W1[3];
            double h10, h11, H0, H1;
            float a1h1 = 0.0f, n1y;
            int r0;
            bool interior;
            if (je == 0 || je == ny) {
                r0 = (je == 0) ? jl : jl - 1;
                const float bd = (float)((je == 0) ? k.borde_sup : k.borde_inf);
                loadW(r0 + 1, c, W0);
                W1[0] = W0[0]; W1[1] = W0[1]; W1[2] = W0[2] * bd;
                h10 = h11 = sh1[r0 + 1][c];
                H0 = H1 = sH[r0 + 1][c];
                n1y = (je == 0) ? -1.0f : 1.0f;
                interior = false;
            } else {
                r0 = jl - 1;
                loadW(r0 + 1, c, W0);
                loadW(r0 + 2, c, W1);
                h10 = sh1[r0 + 1][c];
                h11 = sh1[r0 + 2][c];
                H0 = sH[r0 + 1][c];
                H1 = sH[r0 + 2][c];
                n1y = 1.0f;
                interior = true;
            }
            const bool has0 = (r0 >= 0 && r0 < TY), has1 = interior && (r0 + 1 < TY);
            const float a0h1 = has0 ? sA[0][r0][c] : 0.0f;
            if (has1) a1h1 = sA[0][r0 + 1][c];
            EdgeFlux F;
            if (edge1<1>(kc, W0, W1, h10, h11, H0, H1, n1y, kc.len_y, a0h1, a1h1, interior, F)) {
                if (has0) acc(r0, c, F.Fm, F.lam);
                if (has1) acc(r0 + 1, c, F.Fp, F.lam);
            }
        }
        __syncthreads();
    }

    // x-edge classes C3 then C4, 16-lane segments of 2 rows x 8 edge columns
    constexpr int NB = (TX / 2 + 1 + 7) / 8;
#pragma unroll 1
    for (int par = 0; par < 2; par++) {
        const int ncol = (par == 0) ? (TX / 2 + 1) : (TX / 2);
        for (int e = tid; e < (TY / 2) * NB * 16; e += NT) {
            const int seg = e >> 4, w = e & 15;
            const int r = 2 * (seg / NB) + (w >> 3), j = 8 * (seg % NB) + (w & 7);
            if (j >= ncol) continue;
            const int il = par + 2 * j;
            const int gj = j0 + r, ie = i0 + il;
            if (gj >= ny || ie > nx) continue;
            float W0[3], W1[3];
            double h10, h11, H0, H1;
            float a1h1 = 0.0f, n1x;
            int c0;
            bool interior;
            if (ie == 0 || ie == nx) {
                c0 = (ie == 0) ? il + 1 : il;
                const float bd = (float)((ie == 0) ? k.borde_izq : k.borde_der);
                loadW(r + 1, c0, W0);
                W1[0] = W0[0]; W1[1] = W0[1] * bd; W1[2] = W0[2];
                h10 = h11 = sh1[r + 1][c0];
                H0 = H1 = sH[r + 1][c0];
                n1x = (ie == 0) ? -1.0f : 1.0f;
                interior = false;
            } else {
                c0 = il;
                loadW(r + 1, c0, W0);
                loadW(r + 1, c0 + 1, W1);
                h10 = sh1[r + 1][c0];
                h11 = sh1[r + 1][c0 + 1];
                H0 = sH[r + 1][c0];
                H1 = sH[r + 1][c0 + 1];
                n1x = 1.0f;
                interior = true;
            }
            const float a0h1 = sA[0][r][c0];
            if (interior) a1h1 = sA[0][r][c0 + 1];
            EdgeFlux F;
            if (edge1<0>(kc, W0, W1, h10, h11, H0, H1, n1x, kc.len_x, a0h1, a1h1, interior, F)) {
                acc(r, c0, F.Fm, F.lam);
                if (interior) acc(r, c0 + 1, F.Fp, F.lam);
            }
        }
        __syncthreads();
    }

    // cell update of the tile's own cells (layer 2 stays zero)
    const KCell kk = make_kcell(k, kc);
    Amax = 0.0;
    unsigned hb = 0u;
#pragma unroll
    for (int it = 0; it < CU_IT; it++) {
        const int e = tid + it * NT;
        if (e >= TX * TY) continue;
        const int r = e / TX, c = e % TX;
        const int gj = j0 + r, gi = i0 + c;
        if (gj >= ny || gi >= nx) continue;
        const Idx g = (Idx)gj * nx + gi;
        Pair o[3];
        const double h1d = sh1[r + 1][c + 1];
        o[0].hi = (float)h1d; o[0].lo = (float)(h1d - (double)o[0].hi);
        o[1] = Pair{sq[0][r + 1][c + 1], lo1x[it]};
        o[2] = Pair{sq[1][r + 1][c + 1], lo1y[it]};
        const float ac[3] = {sA[0][r][c + 1], sA[1][r][c + 1], sA[2][r][c + 1]};
        Amax = fmax(Amax, (double)sA[3][r][c + 1]);
        if (emax) {
            const double v = h1d - sH[r + 1][c + 1];
            if (v > emax[2 * g]) {
                emax[2 * g] = v;
                emax[2 * g + 1] = tnow;
            }
        }
        cell_update1(kk, o, ac, &hb);
#pragma unroll
        for (int f = 0; f < 3; f++) {
            stW(f, g, o[f].hi);
            stW(6 + f, g, o[f].lo);
        }
        if (zero2) {
#pragma unroll
            for (int f = 0; f < 3; f++) {
                stW(3 + f, g, 0.0f);
                stW(9 + f, g, 0.0f);
            }
        }
        // the new h1 replaces this cell's staged one (read by no other thread now)
        if (go.out) sh1[r + 1][c + 1] = (double)o[0].hi + (double)o[0].lo;
    }
    raise_health(hc, hb, s);
    if (go.out) gauges_of_tile<TX, TY, NT>(go, &sh1[0][0], nullptr, &sH[0][0], CWP, nx, i0, j0, 0x7fffffff);
    return true;
}

// ------------------------------------------------------------ tile classes
// The single-simulation step is split by tile class (lhysea_tiles.hpp); the
// class tiles are the FTX x FTY tiles of the grid.

// Sediment-free tiles of a step: one block per class tile, tiles of one XCD
// contiguous (blocks are dispatched round-robin over the 8 XCDs, so XCD x
// runs blocks x, x+8, ...; giving it one run of tiles keeps shared halo rows
// in its L2).
template <int TX, int TY, int NT, bool BUF>
__device__ __forceinline__ void fused_bare(const Consts& k, const float* __restrict__ W, float* __restrict__ Wn,
                                           const double* __restrict__ H, const StepArgs& a, const TileCls& tc)
{
    if (!(a.ctl->t < a.t_end)) return;     // the run has reached its end time
    const int nb = tc.tiles_x * tc.tiles_y;
    int b;
    {
        const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
        const int q = nb >> 3, rem = nb & 7;
        b = xcd * q + (xcd < rem ? xcd : rem) + slot;
    }
    const int tx = b % tc.tiles_x, ty = b / tc.tiles_x;
    const int s = a.ctl->step;
    const double dt = step_dt(a.ctl, a.dt_seq);
    const GaugeOut go{a.gauges, a.ng, (a.gseries && s < a.gcap) ? a.gseries + (size_t)s * a.ng : nullptr};
    double Amax;
    if (!tile_step_bare<TX, TY, NT, BUF>(k, W, Wn, H, dt, a.ctl->t, a.emax, tx * TX, ty * TY, tc.cls + (s & 1) * nb + b,
                                    s, a.wn_prev ? a.ctl->cls_from : 0x7fffffff, go, Amax, a.ctl))
        return;
    block_max_atomic<NT>(Amax, &a.ctl->maxA_next);
}

template <int TX, int TY, int NT, bool BUF>
__global__ void __launch_bounds__(NT) k_fused_bare(Consts k, const float* __restrict__ W, float* __restrict__ Wn,
                                                   const double* __restrict__ H, StepArgs a, TileCls tc)
{
    fused_bare<TX, TY, NT, BUF>(k, W, Wn, H, a, tc);
}

// The same with 64-bit addresses, for states beyond state_buf_fits.
template <int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) LH_BARE_SGPR_CAP k_fused_bare_a64(Consts k, const float* __restrict__ W,
                                                                        float* __restrict__ Wn,
                                                                        const double* __restrict__ H, StepArgs a,
                                                                        TileCls tc)
{
    fused_bare<TX, TY, NT, false>(k, W, Wn, H, a, tc);
}

// General tiles of a step, then the step's close.  Each general class tile
// is split into STX-wide sub-tiles (one work item each, the last one clipped
// at the class tile's edge) so that the few general tiles spread over many
// workgroups; a fixed grid walks the work items.  Cells that keep sediment
// mark the next step's general tiles.  Every workgroup then adds its
// wave-speed max and counts itself done; the last one closes the step
// (src/GPU/ShallowWater.cu:431-442): dt history, t += dt, the lagged dt of
// the next step from the max of both launches, the consumed list and the
// counters cleared.  Workgroups beyond the work items leave at once.  Its count increment uses the value its atomicMax
// returned, so it is issued once that max has been performed, and the
// closing workgroup reads the max with an atomic.  The sediment-free launch
// ran before this one, so its maxima are in place.
// The marks of a workgroup's last work item go out from its second wavefront
// while the first one counts the workgroup done and maybe closes the step,
// so their atomics' latencies overlap: the marks write the next step's class
// table, list and count, which the close does not touch, and the next step's
// launches start after this one has ended.
template <int STX, int TY, int NT>
__global__ void __launch_bounds__(NT, 3) k_fused_gen(Consts k, const float* __restrict__ W, float* __restrict__ Wn,
                                                  const double* __restrict__ H, StepArgs a, TileCls tc)
{
    constexpr int NSUB = (FTX + STX - 1) / STX;
    static_assert(TY == FTY && STX % 2 == 0 && NT >= 128, "sub-tiles of a class tile; two wavefronts");
    Ctl* ctl = a.ctl;
    // the step is the launch's own: a workgroup that starts after the last
    // participant closed it would otherwise read the next step's list
    const int s = a.step;
    if (!(ctl->t < a.t_end) || ctl->step != s) return;
    const int n = NSUB * (int)ctl->gcount[s & 1];
    // workgroups with work items (at least one, which then closes the step)
    const unsigned part = (unsigned)max(min(n, (int)gridDim.x), 1);
    if (blockIdx.x >= part) return;
    const int nt = tc.tiles_x * tc.tiles_y;
    const int* list = tc.list + (s & 1) * nt;
    const double dt = step_dt(ctl, a.dt_seq);
    const double tnow = ctl->t;
    const GaugeOut go{a.gauges, a.ng, (a.gseries && s < a.gcap) ? a.gseries + (size_t)s * a.ng : nullptr};
    double Amax = 0.0;
    int mlast = 0, txl = 0, tyl = 0;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int t = list[i / NSUB], sub = i % NSUB;
        const int tx = t % tc.tiles_x, ty = t / tc.tiles_x;
        const int cx0 = tx * FTX, cy0 = ty * FTY;
        int m = 0;
        Amax = fmax(Amax, tile_step<STX, TY, NT>(k, W, Wn, H, dt, tnow, a.emax, cx0 + sub * STX, cy0, nullptr,
                                                 nullptr, cx0 + FTX, go, &m, cx0, cy0, ctl, s));
        // also orders this sub-tile's LDS reads before the next one's staging
        m = block_or(m);
        if (i + (int)gridDim.x < n) {
            if (m && threadIdx.x < 64) mark_general(tc, ctl, tx, ty, s + 1, m);
        } else {
            mlast = m;
            txl = tx;
            tyl = ty;
        }
    }
    // close of the step
    __shared__ double red[NT / 64];
    double v = Amax;
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (mlast && threadIdx.x >= 64 && threadIdx.x < 128) mark_general(tc, ctl, txl, tyl, s + 1, mlast);
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / 64; w++) v = fmax(v, red[w]);
        v = fabs(fmax(v, red[0]));   // sign bit clear also for a NaN from a corrupt state
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

// Every class tile through the general tile function, one block per tile
// (XCD-contiguous order), for grids where the sediment-free tiles are too
// few to pay for the split; the classes are kept up to date all the same.
template <int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_fused_all(Consts k, const float* __restrict__ W, float* __restrict__ Wn,
                                                  const double* __restrict__ H, StepArgs a, TileCls tc)
{
    static_assert(TX == FTX && TY == FTY, "class tiles");
    if (!(a.ctl->t < a.t_end)) return;
    const int nb = tc.tiles_x * tc.tiles_y;
    int b;
    {
        const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
        const int q = nb >> 3, rem = nb & 7;
        b = xcd * q + (xcd < rem ? xcd : rem) + slot;
    }
    const int tx = b % tc.tiles_x, ty = b / tc.tiles_x;
    const int s = a.ctl->step;
    const double dt = step_dt(a.ctl, a.dt_seq);
    const GaugeOut go{a.gauges, a.ng, (a.gseries && s < a.gcap) ? a.gseries + (size_t)s * a.ng : nullptr};
    int m = 0;
    const double Amax = tile_step<TX, TY, NT>(k, W, Wn, H, dt, a.ctl->t, a.emax, tx * TX, ty * TY, nullptr, nullptr,
                                              0x7fffffff, go, &m, tx * TX, ty * TY, a.ctl, s);
    general_tile_end<NT>(Amax, m, tc, a.ctl, tx, ty, s + 1);
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

// Class table of the state W at step ctl->step: one block per class tile;
// cells with sediment mark as in a step.  The table and both counts are
// cleared beforehand.
template <int TX, int TY>
__global__ void __launch_bounds__(256) k_classify(Consts k, const float* __restrict__ W, StepArgs a, TileCls tc)
{
    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const int t = blockIdx.x, tx = t % tc.tiles_x, ty = t / tc.tiles_x;
    int m = 0;
    for (int e = threadIdx.x; e < TX * TY; e += 256) {
        const int ly = e / TX, lx = e % TX;
        const int gj = ty * TY + ly, gi = tx * TX + lx;
        if (gj >= ny || gi >= nx) continue;
        const size_t g = (size_t)gj * nx + gi;
        if (W[3 * N + g] != 0.0f || W[9 * N + g] != 0.0f || W[4 * N + g] != 0.0f || W[5 * N + g] != 0.0f)
            m |= halo_mask(lx, ly);
    }
    m = block_or(m);
    if (m && threadIdx.x < 64) mark_general(tc, a.ctl, tx, ty, a.ctl->step, m);
    if (t == 0 && threadIdx.x == 0) a.ctl->cls_from = a.ctl->step;
}

// Forward recomputation for the discrete adjoint: the step's partial and
// final flux sums at state W with the recorded dt of step `step`.
template <int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_recompute(Consts k, const float* __restrict__ W, const double* __restrict__ H,
                                                  const double* __restrict__ dt_seq, int step, float* __restrict__ P,
                                                  float* __restrict__ ACC, int tiles_x, int tiles_y)
{
    const int nb = tiles_x * tiles_y;
    int b;
    {
        const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
        const int q = nb >> 3, rem = nb & 7;
        b = xcd * q + (xcd < rem ? xcd : rem) + slot;
    }
    const int tx = b % tiles_x, ty = b / tiles_x;
    tile_step<TX, TY, NT, true>(k, W, nullptr, H, dt_seq[step], 0.0, nullptr, tx * TX, ty * TY, P, ACC);
}

// A pointer read from a descriptor in memory is generic to the compiler, so
// its accesses would be flat instructions (which also count against the LDS
// wait counter); the batch buffers are all device memory, and passing them
// through the global address space gives global loads and stores.
template <typename T>
__device__ __forceinline__ T* gmem(T* p)
{
    return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}

// ------------------------------------------------------------- ensemble step
// One launch steps every scenario of a batch.  tab[blockIdx.x] = (scenario
// << 20) | tile, built on the host with the same XCD-contiguous order over the
// concatenated tile lists.  A scenario whose step budget or end time is
// reached skips its tiles; its ping-pong parity is its own step count.
__device__ __forceinline__ bool scen_active(const BatchScen& d, int step, double t)
{
    return step < d.max_steps && t < d.t_end;
}

template <int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_batch(const BatchScen* __restrict__ sd, const unsigned* __restrict__ tab)
{
    const unsigned e = tab[blockIdx.x];
    const BatchScen& d = sd[e >> 20];
    Ctl* ctl = gmem(d.ctl);
    const int step = ctl->step;
    if (!scen_active(d, step, ctl->t)) return;
    const int tile = (int)(e & 0xFFFFFu);
    const int tx = tile % d.tiles_x, ty = tile / d.tiles_x;
    const int cur = step & 1;
    const Consts k = d.k;
    const double Amax = tile_step<TX, TY, NT>(k, gmem(d.W[cur]), gmem(d.W[cur ^ 1]), gmem(d.H), ctl->dt_cur, ctl->t,
                                              d.emax ? gmem(d.emax) : nullptr, tx * TX, ty * TY, nullptr, nullptr,
                                              0x7fffffff, GaugeOut{nullptr, 0, nullptr}, nullptr, 0, 0, ctl, step);
    block_max_atomic<NT>(Amax, &ctl->maxA_next);
}

// Split shape of the batch: the sediment-free class tiles of every scenario
// (the block table lists all 62 x 14 class tiles; general ones leave), then
// the general class tiles of every scenario cut into sub-tiles, each
// scenario walked by its own group of gps workgroups.  Classes as in the
// single-simulation step, per scenario.
template <int TX, int TY, int NT, bool BUF>
__device__ __forceinline__ void batch_bare(const BatchScen* __restrict__ sd, const unsigned* __restrict__ tab)
{
    const unsigned e = tab[blockIdx.x];
    const BatchScen& d = sd[e >> 20];
    Ctl* ctl = gmem(d.ctl);
    const int step = ctl->step;
    if (!scen_active(d, step, ctl->t)) return;
    const int tile = (int)(e & 0xFFFFFu);
    const int tx = tile % d.ctiles_x, ty = tile / d.ctiles_x;
    const int nct = d.ctiles_x * d.ctiles_y;
    const int cur = step & 1;
    const Consts k = d.k;
    double Amax;
    if (!tile_step_bare<TX, TY, NT, BUF>(k, gmem(d.W[cur]), gmem(d.W[cur ^ 1]), gmem(d.H), ctl->dt_cur, ctl->t,
                                    d.emax ? gmem(d.emax) : nullptr, tx * TX, ty * TY,
                                    gmem(d.cls) + (step & 1) * nct + tile, step, ctl->cls_from,
                                    GaugeOut{nullptr, 0, nullptr}, Amax, ctl))
        return;
    block_max_atomic<NT>(Amax, &ctl->maxA_next);
}

template <int TX, int TY, int NT, bool BUF>
__global__ void __launch_bounds__(NT) k_batch_bare(const BatchScen* __restrict__ sd, const unsigned* __restrict__ tab)
{
    batch_bare<TX, TY, NT, BUF>(sd, tab);
}

// The same with 64-bit addresses, for a batch with a state beyond state_buf_fits.
template <int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) LH_BARE_SGPR_CAP k_batch_bare_a64(const BatchScen* __restrict__ sd,
                                                                        const unsigned* __restrict__ tab)
{
    batch_bare<TX, TY, NT, false>(sd, tab);
}

// General sub-tiles of every scenario as one item space: item i of the
// concatenation of the scenarios' lists (a scenario holds 4 x its general
// tile count of items while active, none when stopped) goes to workgroup
// i mod gridDim.x.  Each workgroup forms the prefix sums of the counts in LDS
// and finds an item's scenario by binary search.
template <int STX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_batch_gen(const BatchScen* __restrict__ sd, int nscen)
{
    constexpr int NSUB = (FTX + STX - 1) / STX;
    static_assert(TY == FTY && STX % 2 == 0, "sub-tiles of a class tile");
    __shared__ int pre[kMaxBatchScen + 1];
    __shared__ int part[NT];
    // per-thread run of scenarios, then the block scan of the run totals
    const int per = (nscen + NT - 1) / NT, lo = threadIdx.x * per, hi = min(lo + per, nscen);
    int run = 0;
    for (int i = lo; i < hi; i++) {
        const BatchScen& d = sd[i];
        const int st = d.ctl->step;
        const int c = scen_active(d, st, d.ctl->t) ? NSUB * (int)d.ctl->gcount[st & 1] : 0;
        pre[i + 1] = c;
        run += c;
    }
    part[threadIdx.x] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < NT; t++) {
            const int v = part[t];
            part[t] = acc;
            acc += v;
        }
        pre[0] = 0;
    }
    __syncthreads();
    {
        int acc = part[threadIdx.x];
        for (int i = lo; i < hi; i++) {
            acc += pre[i + 1];
            pre[i + 1] = acc;
        }
    }
    __syncthreads();
    const int total = pre[nscen];
    double Amax = 0.0;
    int sid_max = -1;
    for (int i = blockIdx.x; i < total; i += gridDim.x) {
        // scenario: the last sid with pre[sid] <= i
        int l = 0, h = nscen - 1;
        while (l < h) {
            const int m = (l + h + 1) >> 1;
            if (pre[m] <= i) l = m; else h = m - 1;
        }
        const int sid = l, item = i - pre[sid];
        const BatchScen& d = sd[sid];
        Ctl* ctl = gmem(d.ctl);
        const int step = ctl->step;
        const int nct = d.ctiles_x * d.ctiles_y;
        const int t = gmem(d.list)[(step & 1) * nct + item / NSUB], sub = item % NSUB;
        const int tx = t % d.ctiles_x, ty = t / d.ctiles_x;
        const int cx0 = tx * FTX, cy0 = ty * FTY;
        const int cur = step & 1;
        int m = 0;
        const double A = tile_step<STX, TY, NT>(d.k, gmem(d.W[cur]), gmem(d.W[cur ^ 1]), gmem(d.H), ctl->dt_cur,
                                                ctl->t, d.emax ? gmem(d.emax) : nullptr, cx0 + sub * STX, cy0,
                                                nullptr, nullptr, cx0 + FTX, GaugeOut{nullptr, 0, nullptr}, &m, cx0,
                                                cy0, ctl, step);
        m = block_or(m);   // also orders this sub-tile's LDS reads before the next one's staging
        if (m && threadIdx.x < 64)
            mark_general(TileCls{gmem(d.cls), gmem(d.list), d.ctiles_x, d.ctiles_y}, ctl, tx, ty, step + 1, m);
        // the wave-speed max goes to the item's scenario (flushed when the scenario changes)
        if (sid != sid_max && sid_max >= 0) {
            block_max_atomic<NT>(Amax, &sd[sid_max].ctl->maxA_next);
            Amax = 0.0;
        }
        sid_max = sid;
        Amax = fmax(Amax, A);
    }
    if (sid_max >= 0) block_max_atomic<NT>(Amax, &sd[sid_max].ctl->maxA_next);
}

// Class table of one scenario at its current step: one block per class tile.
__global__ void __launch_bounds__(256) k_batch_classify(const BatchScen* __restrict__ sd, int sid)
{
    const BatchScen& d = sd[sid];
    const Consts& k = d.k;
    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const int step = d.ctl->step;
    const float* W = d.W[step & 1];
    const int t = blockIdx.x, tx = t % d.ctiles_x, ty = t / d.ctiles_x;
    int m = 0;
    for (int e = threadIdx.x; e < FTX * FTY; e += 256) {
        const int ly = e / FTX, lx = e % FTX;
        const int gj = ty * FTY + ly, gi = tx * FTX + lx;
        if (gj >= ny || gi >= nx) continue;
        const size_t g = (size_t)gj * nx + gi;
        if (W[3 * N + g] != 0.0f || W[9 * N + g] != 0.0f || W[4 * N + g] != 0.0f || W[5 * N + g] != 0.0f)
            m |= halo_mask(lx, ly);
    }
    m = block_or(m);
    if (m && threadIdx.x < 64) mark_general(TileCls{d.cls, d.list, d.ctiles_x, d.ctiles_y}, d.ctl, tx, ty, step, m);
    if (t == 0 && threadIdx.x == 0) d.ctl->cls_from = step;
}

// One block per scenario: gauges of the new state, dt history, t += dt, the
// lagged dt of the next step from this step's wave-speed max.
__global__ void k_batch_finalize(const BatchScen* __restrict__ sd)
{
    const BatchScen& d = sd[blockIdx.x];
    Ctl* ctl = gmem(d.ctl);
    const int s = ctl->step;
    if (!scen_active(d, s, ctl->t)) return;
    const double dt = ctl->dt_cur;
    const size_t N = (size_t)d.k.nx * d.k.ny;
    const float* Wn = d.W[(s & 1) ^ 1];
    if (d.gseries && s < d.gcap)
        for (int g = threadIdx.x; g < d.ng; g += blockDim.x) {
            const size_t c = (size_t)d.gauges[g];
            const double h1 = (double)Wn[c] + (double)Wn[6 * N + c];
            const double h2 = (double)Wn[3 * N + c] + (double)Wn[9 * N + c];
            d.gseries[(size_t)s * d.ng + g] = h1 + h2 - d.H[c];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (d.dt_hist && s < d.hcap) d.dt_hist[s] = dt;
        const double maxA = __longlong_as_double((long long)ctl->maxA_next);
        ctl->t = ctl->t + dt;
        ctl->dt_cur = dt_from_maxA(maxA, d.k.cfl, d.k.area);
        ctl->maxA_last = maxA;
        ctl->maxA_next = 0ull;
        ctl->gcount[s & 1] = 0u;      // the general-tile list of this step is consumed
        ctl->step = s + 1;
    }
}

// State conversion between FP64 planes (6 of N) and the compensated layout
// (6 value planes, then 6 compensation planes): value = the FP64 number
// rounded to FP32, compensation = the FP32 rounding of what is left.
__global__ void __launch_bounds__(256) k_split_state(const double* __restrict__ src, float* __restrict__ W, size_t n6)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n6; i += (size_t)gridDim.x * blockDim.x) {
        const double x = src[i];
        const float hi = (float)x;
        W[i] = hi;
        W[n6 + i] = (float)(x - (double)hi);
    }
}

__global__ void __launch_bounds__(256) k_merge_state(const float* __restrict__ W, double* __restrict__ dst, size_t n6)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n6; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (double)W[i] + (double)W[n6 + i];
}

}  // namespace f32

static unsigned stream_blocks(size_t n)
{
    const size_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

hipError_t f32c_split_state(const double* src, float* W, size_t N, hipStream_t s)
{
    hipLaunchKernelGGL(f32::k_split_state, dim3(stream_blocks(6 * N)), dim3(256), 0, s, src, W, 6 * N);
    return hipGetLastError();
}

hipError_t f32c_merge_state(const float* W, double* dst, size_t N, hipStream_t s)
{
    hipLaunchKernelGGL(f32::k_merge_state, dim3(stream_blocks(6 * N)), dim3(256), 0, s, W, dst, 6 * N);
    return hipGetLastError();
}

hipError_t f32c_initial_dt(const Consts& k, const float* W, Ctl* ctl, hipStream_t s)
{
    const size_t N = (size_t)k.nx * k.ny;
    hipLaunchKernelGGL(f32::k_dt0, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, k, W, ctl);
    hipLaunchKernelGGL(f32::k_dt0_finalize, dim3(1), dim3(64), 0, s, k, ctl);
    return hipGetLastError();
}

int f32c_tiles(const Consts& k)
{
    return ((k.nx + f32::FTX - 1) / f32::FTX) * ((k.ny + f32::FTY - 1) / f32::FTY);
}

hipError_t f32c_classify(const Consts& k, const float* W, const StepArgs& a, hipStream_t s)
{
    const int nt = f32c_tiles(k);
    const TileCls tc{a.tcls, a.tlist, (k.nx + f32::FTX - 1) / f32::FTX, (k.ny + f32::FTY - 1) / f32::FTY};
    hipError_t e = clear_classes(a.tcls, 2 * nt, a.ctl, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((f32::k_classify<f32::FTX, f32::FTY>), dim3((unsigned)nt), dim3(256), 0, s, k, W, a, tc);
    return hipGetLastError();
}

// Split (a.split): the sediment-free launch, then the general sub-tiles and
// the step's close, back to back on one stream.  The general launch takes
// about one general sub-tile's latency (~25 us on an MI355X) however few
// general tiles there are, so the split pays only when the sediment-free
// tiles are many; otherwise every tile takes the general tile function.
hipError_t f32c_step_fused(const Consts& k, const float* W, float* Wn, const double* H, const StepArgs& a,
                           hipStream_t s)
{
    const int tiles_x = (k.nx + f32::FTX - 1) / f32::FTX;
    const int tiles_y = (k.ny + f32::FTY - 1) / f32::FTY;
    const int nt = tiles_x * tiles_y;
    const TileCls tc{a.tcls, a.tlist, tiles_x, tiles_y};
    hipError_t e = a.ev0 ? hipEventRecord(a.ev0, s) : hipSuccess;
    if (e != hipSuccess) return e;
    if (a.split) {
        if (state_buf_fits((size_t)k.nx * k.ny))
            hipLaunchKernelGGL((f32::k_fused_bare<f32::FTX, f32::FTY, f32::FNT, true>), dim3((unsigned)nt),
                               dim3(f32::FNT), 0, s, k, W, Wn, H, a, tc);
        else
            hipLaunchKernelGGL((f32::k_fused_bare_a64<f32::FTX, f32::FTY, f32::FNT>), dim3((unsigned)nt),
                               dim3(f32::FNT), 0, s, k, W, Wn, H, a, tc);
        hipLaunchKernelGGL((f32::k_fused_gen<f32::GSTX, f32::FTY, f32::GNT>), dim3(f32::GGRID), dim3(f32::GNT), 0, s,
                           k, W, Wn, H, a, tc);
        if (a.ev1 && (e = hipEventRecord(a.ev1, s)) != hipSuccess) return e;
    } else {
        hipLaunchKernelGGL((f32::k_fused_all<f32::FTX, f32::FTY, f32::FNT>), dim3((unsigned)nt), dim3(f32::FNT), 0, s,
                           k, W, Wn, H, a, tc);
        if (a.ev1 && (e = hipEventRecord(a.ev1, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(f32::k_close, dim3(1), dim3(64), 0, s, k, a);
    }
    return hipGetLastError();
}

// measured on an MI355X (single slide scenarios, step kernels with and without
// the split): FP32C 1024^2 (1238 sediment-free tiles) 47.2 vs 47.3 us, 1536^2
// (2712) 67.6 vs 79.1 us; FP64 (30 x 14 tiles) 1024^2 (2559) 99.2 vs 102.8 us
int f32c_split_min_bare() { return 1000; }

hipError_t f32c_recompute(const Consts& k, const float* W, const double* H, const double* dt_seq, int step, float* P,
                          float* ACC, hipStream_t s)
{
    const int tiles_x = (k.nx + f32::FTX - 1) / f32::FTX;
    const int tiles_y = (k.ny + f32::FTY - 1) / f32::FTY;
    hipLaunchKernelGGL((f32::k_recompute<f32::FTX, f32::FTY, f32::FNT>), dim3((unsigned)(tiles_x * tiles_y)),
                       dim3(f32::FNT), 0, s, k, W, H, dt_seq, step, P, ACC, tiles_x, tiles_y);
    return hipGetLastError();
}

// shape 0 (kBatchSplit) is the class-tile split: 62 x 14 class tiles of 512
// threads for the sediment-free ones, general ones in 16 x 14 sub-tiles
const BatchTile kBatchTiles[kNumBatchTiles] = {{62, 14, 512}, {30, 14, 256}, {62, 30, 1024}, {126, 14, 1024}};

hipError_t f32c_batch_classify(const BatchScen& d, const BatchScen* sd_dev, int sid, hipStream_t s)
{
    const int nct = d.ctiles_x * d.ctiles_y;
    hipError_t e = clear_classes(d.cls, 2 * nct, d.ctl, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(f32::k_batch_classify, dim3((unsigned)nct), dim3(256), 0, s, sd_dev, sid);
    return hipGetLastError();
}

hipError_t f32c_batch_step(int tile, const BatchScen* sd, int nscen, const unsigned* tab, int nblocks, bool buf,
                           hipStream_t s)
{
    switch (tile) {
    case kBatchSplit: {
        // tab lists every 62 x 14 class tile; a fixed grid walks the general
        // sub-tiles of all scenarios
        if (buf)
            hipLaunchKernelGGL((f32::k_batch_bare<f32::FTX, f32::FTY, f32::FNT, true>), dim3(nblocks), dim3(f32::FNT),
                               0, s, sd, tab);
        else
            hipLaunchKernelGGL((f32::k_batch_bare_a64<f32::FTX, f32::FTY, f32::FNT>), dim3(nblocks),
                               dim3(f32::FNT), 0, s, sd, tab);
        hipLaunchKernelGGL((f32::k_batch_gen<f32::GSTX, f32::FTY, f32::GNT>), dim3(f32::GGRID), dim3(f32::GNT), 0, s,
                           sd, nscen);
        break;
    }
    case 1: hipLaunchKernelGGL((f32::k_batch<30, 14, 256>), dim3(nblocks), dim3(256), 0, s, sd, tab); break;
    case 2: hipLaunchKernelGGL((f32::k_batch<62, 30, 1024>), dim3(nblocks), dim3(1024), 0, s, sd, tab); break;
    case 3: hipLaunchKernelGGL((f32::k_batch<126, 14, 1024>), dim3(nblocks), dim3(1024), 0, s, sd, tab); break;
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(f32::k_batch_finalize, dim3(nscen), dim3(64), 0, s, sd);
    return hipGetLastError();
}

}  // namespace lhysea
