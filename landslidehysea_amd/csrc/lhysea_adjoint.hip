// Discrete adjoint of the LandSlide-HySEA step on MI355X, with binomial
// checkpointing of state snapshots in HBM.
//
// One reverse step n (state W^n in the working buffer, frozen dt^n, adjoint
// lambda^{n+1} of the step's output state):
//   1. recompute the four edge classes (even y, odd y, even x, odd x) with the
//      forward step's own tile kernel, keeping the partial h-sums each class
//      reads (the positivity limiter's inputs) and the final flux sums;
//   2. cell adjoint: lambda^{n+1} -> adjoint of the flux sums, the direct part
//      of lambda^n, and d/dmu (Coulomb coefficient) of the cell update;
//   3. edge adjoints in reverse class order: each edge adds its state
//      adjoints to lambda^n and its partial-sum adjoints to the flux-sum
//      adjoints of its cells, which the earlier classes then see.
// Within one class no two edges share a cell, so every update is a plain
// read-modify-write.  The loss is the L2 misfit of eta = h1 + h2 - H at the
// gauges after every step; its gradient is injected into lambda^n.
#include <cmath>
#include <vector>
#include "lhysea_internal.hpp"
#include "lhysea_adjoint.hpp"

namespace lhysea {
namespace adj {

template <typename F>
struct SV;
template <>
struct SV<float> {
    const float* W;
    size_t N;
    __device__ double h1(size_t c) const { return (double)W[c] + (double)W[6 * N + c]; }
    __device__ double h2(size_t c) const { return (double)W[3 * N + c] + (double)W[9 * N + c]; }
    __device__ float q(int k, size_t c) const { return W[(k < 2 ? 1 + k : 2 + k) * N + c]; }
};
template <>
struct SV<double> {
    const double* W;
    size_t N;
    __device__ double h1(size_t c) const { return W[c]; }
    __device__ double h2(size_t c) const { return W[3 * N + c]; }
    __device__ double q(int k, size_t c) const { return W[(k < 2 ? 1 + k : 2 + k) * N + c]; }
};

template <typename F>
struct Work {
    F* P;          // [3][2][N] partial h1/h2 sums read by classes 1..3
    F* ACC;        // [6][N] flux sums
    F* AB;         // [6][N] adjoints of the flux sums
    double* LAM;   // [6][N] state adjoint (h1,q1x,q1y,h2,q2x,q2y)
    double* gmu;   // d/dmu, one slot per workgroup (Pouliquen: d/dtan(angle i) at gmu[i * gstride + slot])
    size_t gstride;
};

// Edge geometry of edge e in class cls (the reference's class layout,
// src/GPU/Arista_kernel.cu:651-949).
struct EdgeRef {
    size_t c0, c1;
    bool interior, xedge;
    float n1x, n1y;   // 0 or +-1
    double len, bd;   // taken in the adjoint's precision
};

__device__ __forceinline__ bool edge_ref(const Consts& k, long e, int cls, EdgeRef& r)
{
    const int nx = k.nx, ny = k.ny;
    if (cls < 2) {
        const int nrow = (ny - cls) / 2 + 1;
        if (e >= (long)nrow * nx) return false;
        const int je = cls + 2 * (int)(e / nx), i = (int)(e % nx);
        r.xedge = false;
        r.len = k.dx;
        r.n1x = 0.0f;
        if (je == 0 || je == ny) {
            r.c0 = (je == 0) ? (size_t)i : (size_t)(ny - 1) * nx + i;
            r.c1 = r.c0;
            r.interior = false;
            r.n1y = (je == 0) ? -1.0f : 1.0f;
            r.bd = (je == 0) ? k.borde_sup : k.borde_inf;
        } else {
            r.c0 = (size_t)(je - 1) * nx + i;
            r.c1 = r.c0 + nx;
            r.interior = true;
            r.n1y = 1.0f;
            r.bd = 1.0f;
        }
    } else {
        const int par = cls - 2, ncol = (nx - par) / 2 + 1;
        if (e >= (long)ncol * ny) return false;
        const int j = (int)(e / ncol), ie = par + 2 * (int)(e % ncol);
        r.xedge = true;
        r.len = k.dy;
        r.n1y = 0.0f;
        if (ie == 0 || ie == nx) {
            r.c0 = (ie == 0) ? (size_t)j * nx : (size_t)j * nx + nx - 1;
            r.c1 = r.c0;
            r.interior = false;
            r.n1x = (ie == 0) ? -1.0f : 1.0f;
            r.bd = (ie == 0) ? k.borde_izq : k.borde_der;
        } else {
            r.c0 = (size_t)j * nx + ie - 1;
            r.c1 = r.c0 + 1;
            r.interior = true;
            r.n1x = 1.0f;
            r.bd = 1.0f;
        }
    }
    return true;
}

// Inputs of one edge: cell 0, and cell 1 or the ghost of cell 0 (normal
// discharge times borde).
template <typename F>
struct EdgeIn {
    double h10, h20, h11, h21, H0, H1;
    F q0[4], q1[4];
};

template <typename F>
__device__ __forceinline__ void edge_inputs(const SV<F>& sv, const double* H, const EdgeRef& r, EdgeIn<F>& in)
{
    const size_t c1 = r.interior ? r.c1 : r.c0;
    in.h10 = sv.h1(r.c0);
    in.h20 = sv.h2(r.c0);
    in.h11 = sv.h1(c1);
    in.h21 = sv.h2(c1);
    in.H0 = H[r.c0];
    in.H1 = H[c1];
    for (int k = 0; k < 4; k++) {
        in.q0[k] = sv.q(k, r.c0);
        in.q1[k] = sv.q(k, c1);
    }
    if (!r.interior) {
        const int nk = r.xedge ? 0 : 1;   // normal component within (qx, qy)
        in.q1[nk] *= (F)r.bd;
        in.q1[2 + nk] *= (F)r.bd;
    }
}

// constants of the edge functions in the adjoint's precision F
template <typename F>
struct EC {
    F eps, eps4, g, gp, r, mu, dt, area;
    F tan1, tan2, tan3, tan4, LoverH;   // Pouliquen law
};

template <typename F>
__device__ __forceinline__ EC<F> edge_consts(const Consts& k, double dt)
{
    EC<F> c;
    c.eps = (F)k.eps_h;
    c.eps4 = c.eps * c.eps * c.eps * c.eps;
    c.g = (F)k.g;
    c.gp = (F)(k.g * (1.0 - k.r));
    c.r = (F)k.r;
    c.mu = (F)k.mu;
    c.dt = (F)dt;
    c.area = (F)k.area;
    c.tan1 = (F)k.tan1;
    c.tan2 = (F)k.tan2;
    c.tan3 = (F)k.tan3;
    c.tan4 = (F)k.tan4;
    c.LoverH = (F)k.LoverH;
    return c;
}

__device__ __forceinline__ double block_sum256(double v)
{
    __shared__ double red[4];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = red[0] + red[1] + red[2] + red[3];
    return s;
}

// 2. cell adjoint: lambda^{n+1} -> flux-sum adjoints, direct lambda^n, d/dmu
//    (PQ, the Pouliquen law: d/dtan of the four angles)
template <typename F, bool PQ>
__global__ void __launch_bounds__(256) k_cell_adj(Consts k, SV<F> sv, Work<F> w, const double* dt_hist, int step)
{
    constexpr int NG = PQ ? 4 : 1;
    const size_t N = sv.N;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double gm[NG] = {};
    if (c < N) {
        const double dt = dt_hist[step];
        const F eps = (F)k.eps_h;
        F o[6] = {(F)sv.h1(c), sv.q(0, c), sv.q(1, c), (F)sv.h2(c), sv.q(2, c), sv.q(3, c)};
        F a[6], gn[6], n[6], go[6], ga[6], gmu;
        for (int f = 0; f < 6; f++) {
            a[f] = w.ACC[f * N + c];
            gn[f] = (F)w.LAM[f * N + c];
        }
#define LH_CELL_ARGS                                                                                           \
    eps, eps * eps * eps * eps, (F)k.r, (F)k.g, (F)k.mfc, (F)k.mf0, (F)k.mfs, (F)dt, (F)(dt / k.area), (F)k.vmax1, \
        (F)k.vmax2, o[0], o[1], o[2], o[3], o[4], o[5], a[0], a[1], a[2], a[3], a[4], a[5], (F)k.mu, n[0], n[1],   \
        n[2], n[3], n[4], n[5], gn[0], gn[1], gn[2], gn[3], gn[4], gn[5], go[0], go[1], go[2], go[3], go[4], go[5], \
        ga[0], ga[1], ga[2], ga[3], ga[4], ga[5], gmu
        // a wavefront of cells without sediment and layer-2 sums takes the
        // specialisation (same values and adjoints)
        const bool bare = (o[3] == F(0)) && (o[4] == F(0)) && (o[5] == F(0)) && (a[3] == F(0)) && (a[4] == F(0)) &&
                          (a[5] == F(0));
        // without sediment the Coulomb scaling multiplies zeros: either law
        // takes the specialisation, whose coefficient adjoint is zero
        if (__all(bare)) {
            adj_cell_ns<F>(LH_CELL_ARGS);
        } else if constexpr (PQ) {
            F gt[4];
            adj_cell_pq<F>(eps, eps * eps * eps * eps, (F)k.r, (F)k.g, (F)k.mfc, (F)k.mf0, (F)k.mfs, (F)dt,
                           (F)(dt / k.area), (F)k.vmax1, (F)k.vmax2, (F)k.LoverH, o[0], o[1], o[2], o[3], o[4], o[5],
                           a[0], a[1], a[2], a[3], a[4], a[5], (F)k.tan1, (F)k.tan2, (F)k.tan3, (F)k.tan4, n[0], n[1],
                           n[2], n[3], n[4], n[5], gn[0], gn[1], gn[2], gn[3], gn[4], gn[5], go[0], go[1], go[2],
                           go[3], go[4], go[5], ga[0], ga[1], ga[2], ga[3], ga[4], ga[5], gt[0], gt[1], gt[2], gt[3]);
            for (int i = 0; i < NG; i++) gm[i] = (double)gt[i];
        } else {
            adj_cell<F>(LH_CELL_ARGS);
            gm[0] = (double)gmu;
        }
#undef LH_CELL_ARGS
        for (int f = 0; f < 6; f++) {
            w.LAM[f * N + c] = (double)go[f];
            w.AB[f * N + c] = ga[f];
        }
    }
    // one slot per workgroup (and component), summed in slot order at the end
    // of the sweep: the result does not depend on the workgroups' timing
    for (int i = 0; i < NG; i++) {
        const double s = block_sum256(gm[i]);
        if (threadIdx.x == 0 && s != 0.0) w.gmu[i * w.gstride + blockIdx.x] += s;
        __syncthreads();   // block_sum256's LDS slots are reused
    }
}

// 3. edge adjoints of one class
template <typename F, bool PQ>
__global__ void __launch_bounds__(256) k_edge_adj(Consts k, SV<F> sv, const double* H, Work<F> w,
                                                  const double* dt_hist, int step, int cls)
{
    const size_t N = sv.N;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    EdgeRef r;
    if (!edge_ref(k, e, cls, r)) return;
    const EC<F> c = edge_consts<F>(k, dt_hist[step]);
    F a0h1 = F(0), a0h2 = F(0), a1h1 = F(0), a1h2 = F(0);
    if (cls > 0) {
        const F* P = w.P + (size_t)(cls - 1) * 2 * N;
        a0h1 = P[r.c0];
        a0h2 = P[N + r.c0];
        if (r.interior) {
            a1h1 = P[r.c1];
            a1h2 = P[N + r.c1];
        }
    }
    EdgeIn<F> in;
    edge_inputs(sv, H, r, in);
    F gfm[6], gfp[6];
    for (int f = 0; f < 6; f++) {
        gfm[f] = -w.AB[f * N + r.c0];
        gfp[f] = r.interior ? -w.AB[f * N + r.c1] : F(0);
    }
    F fm[6], fp[6], gq0[4], gq1[4], ga0h1, ga0h2, ga1h1, ga1h2;
    double gh10, gh20, gh11, gh21;
#define LH_EDGE_ARGS                                                                                            \
    (F)c.eps, (F)c.eps4, (F)c.g, (F)c.gp, (F)c.r, (F)c.mu, (F)c.dt, (F)c.area, (F)r.len, (F)r.n1x, (F)r.n1y,        \
        r.interior ? F(1) : F(0), in.H0, in.H1, in.h10, in.h20, in.h11, in.h21, in.q0[0], in.q0[1], in.q0[2],        \
        in.q0[3], in.q1[0], in.q1[1], in.q1[2], in.q1[3], a0h1, a0h2, a1h1, a1h2, fm[0], fm[1], fm[2], fm[3], fm[4], \
        fm[5], fp[0], fp[1], fp[2], fp[3], fp[4], fp[5], gfm[0], gfm[1], gfm[2], gfm[3], gfm[4], gfm[5], gfp[0],     \
        gfp[1], gfp[2], gfp[3], gfp[4], gfp[5], gh10, gh20, gh11, gh21, gq0[0], gq0[1], gq0[2], gq0[3], gq1[0],      \
        gq1[1], gq1[2], gq1[3], ga0h1, ga0h2, ga1h1, ga1h2
    // a wavefront of edges whose cells carry no sediment (and no layer-2
    // partial sums) takes the specialisation (same values and adjoints)
    const bool bare = (in.h20 == 0.0) && (in.h21 == 0.0) && (in.q0[2] == F(0)) && (in.q0[3] == F(0)) &&
                      (in.q1[2] == F(0)) && (in.q1[3] == F(0)) && (a0h2 == F(0)) && (a1h2 == F(0));
    // without sediment the Coulomb flag is false under either law
    if (__all(bare)) {
        adj_edge_ns<F>(LH_EDGE_ARGS);
    } else if constexpr (PQ) {
        adj_edge_pq<F>((F)c.eps, (F)c.eps4, (F)c.g, (F)c.gp, (F)c.r, (F)c.mu, (F)c.dt, (F)c.area, (F)r.len, (F)r.n1x,
                       (F)r.n1y, r.interior ? F(1) : F(0), in.H0, in.H1, c.tan1, c.tan2, c.tan3, c.tan4, c.LoverH,
                       in.h10, in.h20, in.h11, in.h21, in.q0[0], in.q0[1], in.q0[2], in.q0[3], in.q1[0], in.q1[1],
                       in.q1[2], in.q1[3], a0h1, a0h2, a1h1, a1h2, fm[0], fm[1], fm[2], fm[3], fm[4], fm[5], fp[0],
                       fp[1], fp[2], fp[3], fp[4], fp[5], gfm[0], gfm[1], gfm[2], gfm[3], gfm[4], gfm[5], gfp[0],
                       gfp[1], gfp[2], gfp[3], gfp[4], gfp[5], gh10, gh20, gh11, gh21, gq0[0], gq0[1], gq0[2], gq0[3],
                       gq1[0], gq1[1], gq1[2], gq1[3], ga0h1, ga0h2, ga1h1, ga1h2);
    } else {
        adj_edge<F>(LH_EDGE_ARGS);
    }
#undef LH_EDGE_ARGS
    double* L = w.LAM;
    if (r.interior) {
        L[r.c0] += gh10;
        L[3 * N + r.c0] += gh20;
        L[N + r.c0] += (double)gq0[0];
        L[2 * N + r.c0] += (double)gq0[1];
        L[4 * N + r.c0] += (double)gq0[2];
        L[5 * N + r.c0] += (double)gq0[3];
        L[r.c1] += gh11;
        L[3 * N + r.c1] += gh21;
        L[N + r.c1] += (double)gq1[0];
        L[2 * N + r.c1] += (double)gq1[1];
        L[4 * N + r.c1] += (double)gq1[2];
        L[5 * N + r.c1] += (double)gq1[3];
        w.AB[r.c1] += ga1h1;
        w.AB[3 * N + r.c1] += ga1h2;
    } else {
        // the ghost copies cell 0 with its normal discharge times borde
        const double bd = (double)(F)r.bd;
        const int nk = r.xedge ? 0 : 1;
        double gq[4];
        for (int q = 0; q < 4; q++) gq[q] = (double)gq0[q] + (double)gq1[q] * (((q & 1) == nk) ? bd : 1.0);
        L[r.c0] += gh10 + gh11;
        L[3 * N + r.c0] += gh20 + gh21;
        L[N + r.c0] += gq[0];
        L[2 * N + r.c0] += gq[1];
        L[4 * N + r.c0] += gq[2];
        L[5 * N + r.c0] += gq[3];
    }
    w.AB[r.c0] += ga0h1;
    w.AB[3 * N + r.c0] += ga0h2;
}

// loss gradient at the gauges: eta = h1 + h2 - H
__global__ void k_inject(double* LAM, size_t N, const int* gauges, int ng, const double* G)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ng) {
        const size_t c = (size_t)gauges[g];
        // gauges on the same cell add up; one thread handles all duplicates of a cell
        bool first = true;
        for (int h = 0; h < g; h++)
            if (gauges[h] == gauges[g]) first = false;
        if (!first) return;
        double s = 0.0;
        for (int h = g; h < ng; h++)
            if (gauges[h] == gauges[g]) s += G[h];
        LAM[c] += s;
        LAM[3 * N + c] += s;
    }
}

__global__ void k_set_step(Ctl* ctl, int step)
{
    if (threadIdx.x == 0) ctl->step = step;
}

// ------------------------------------------------------- fused reverse step
// One workgroup computes lambda^n of a TX x TY tile (even origin) in one pass,
// FP32 adjoint.  lambda^n of a cell depends on the cell adjoints and on the
// edge adjoints of a cone that reaches exactly two cells beyond the tile on
// every side (each class-k edge needs its cells' flux-sum adjoints after the
// later classes' partial-sum adjoints, which come from the later-class edges
// of those cells), so the workgroup stages the (TX+4) x (TY+4) box from HBM:
// state, depth, the partial sums P and flux sums ACC of the recomputed
// forward step, and lambda^{n+1}.  It then runs the cell adjoint of every box
// cell (flux-sum adjoints in LDS; the direct part of lambda^n and d/dmu for the
// tile's own cells), the four edge classes in reverse order over every edge
// inside the box (LDS updates; within a class no two edges share a cell), and
// writes lambda^n of its own cells.  lambda^{n+1} is read from one buffer and
// lambda^n written to another.
// PQ: the Pouliquen law; the workgroup's d/dtan of the four angles go to
// gmu[i * gstride + blockIdx.x] (Coulomb: d/dmu to gmu[blockIdx.x]).
template <int TX, int TY, int NT, bool PQ>
__global__ void __launch_bounds__(NT) k_adj_fused(Consts k, const float* __restrict__ W,
                                                  const double* __restrict__ H, const float* __restrict__ P,
                                                  const float* __restrict__ ACC, const double* __restrict__ LIN,
                                                  double* __restrict__ LOUT, double* gmu, size_t gstride,
                                                  const double* __restrict__ dt_hist, int step, int tiles_x,
                                                  int tiles_y, const int* __restrict__ gauges, int ng,
                                                  const double* __restrict__ Gk)
{
    static_assert(TX % 2 == 0 && TY % 2 == 0, "tile origins must stay even");
    constexpr int NG = PQ ? 4 : 1;
    constexpr int HW = 2, CW = TX + 2 * HW, RW = TY + 2 * HW;
    __shared__ double sh1[RW][CW], sh2[RW][CW], sH[RW][CW];
    __shared__ float sq[4][RW][CW];
    __shared__ float sP[3][2][RW][CW];
    __shared__ float sAB[6][RW][CW];
    __shared__ double sL[6][TY][TX];
    __shared__ double red[NG][NT / 64];

    const int nx = k.nx, ny = k.ny;
    const size_t N = (size_t)nx * ny;
    const int nb = tiles_x * tiles_y;
    int b;
    {
        const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
        const int q = nb >> 3, rem = nb & 7;
        b = xcd * q + (xcd < rem ? xcd : rem) + slot;
    }
    const int i0 = (b % tiles_x) * TX, j0 = (b / tiles_x) * TY;
    const int bi = i0 - HW, bj = j0 - HW;          // global index of box cell (0, 0)
    const int tid = threadIdx.x;
    const double dt = dt_hist[step];
    const EC<float> c = edge_consts<float>(k, dt);
    const float eps = (float)k.eps_h;

    // 1. stage the box; cell adjoint of every box cell
    double gm[NG] = {};
    for (int e = tid; e < RW * CW; e += NT) {
        const int r = e / CW, cc = e % CW;
        const int gj = bj + r, gi = bi + cc;
        if (gj < 0 || gj >= ny || gi < 0 || gi >= nx) {
#pragma unroll
            for (int f = 0; f < 6; f++) sAB[f][r][cc] = 0.0f;
            continue;
        }
        const size_t g = (size_t)gj * nx + gi;
        const double h1 = (double)W[g] + (double)W[6 * N + g], h2 = (double)W[3 * N + g] + (double)W[9 * N + g];
        float o[6] = {(float)h1, W[N + g], W[2 * N + g], (float)h2, W[4 * N + g], W[5 * N + g]};
        sh1[r][cc] = h1;
        sh2[r][cc] = h2;
        sH[r][cc] = H[g];
        sq[0][r][cc] = o[1];
        sq[1][r][cc] = o[2];
        sq[2][r][cc] = o[4];
        sq[3][r][cc] = o[5];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            sP[m][0][r][cc] = P[(size_t)m * 2 * N + g];
            sP[m][1][r][cc] = P[(size_t)m * 2 * N + N + g];
        }
        float a[6], gn[6], n[6], go[6], ga[6], gmuc, gt[NG];
#pragma unroll
        for (int f = 0; f < 6; f++) {
            a[f] = ACC[f * N + g];
            gn[f] = (float)LIN[f * N + g];
        }
#define LH_CELL_ARGS                                                                                               \
    eps, eps * eps * eps * eps, (float)k.r, (float)k.g, (float)k.mfc, (float)k.mf0, (float)k.mfs, (float)dt,           \
        (float)(dt / k.area), (float)k.vmax1, (float)k.vmax2, o[0], o[1], o[2], o[3], o[4], o[5], a[0], a[1], a[2],    \
        a[3], a[4], a[5], (float)k.mu, n[0], n[1], n[2], n[3], n[4], n[5], gn[0], gn[1], gn[2], gn[3], gn[4], gn[5],  \
        go[0], go[1], go[2], go[3], go[4], go[5], ga[0], ga[1], ga[2], ga[3], ga[4], ga[5], gmuc
        const bool bare = (o[3] == 0.0f) && (o[4] == 0.0f) && (o[5] == 0.0f) && (a[3] == 0.0f) && (a[4] == 0.0f) &&
                          (a[5] == 0.0f);
        // without sediment the Coulomb scaling multiplies zeros under either law
        if (__all(bare)) {
            adj_cell_ns<float>(LH_CELL_ARGS);
            gt[0] = PQ ? 0.0f : gmuc;   // the coefficient adjoint is zero without sediment
#pragma unroll
            for (int i = 1; i < NG; i++) gt[i] = 0.0f;
        } else if constexpr (PQ) {
            adj_cell_pq<float>(eps, eps * eps * eps * eps, (float)k.r, (float)k.g, (float)k.mfc, (float)k.mf0,
                               (float)k.mfs, (float)dt, (float)(dt / k.area), (float)k.vmax1, (float)k.vmax2,
                               c.LoverH, o[0], o[1], o[2], o[3], o[4], o[5], a[0], a[1], a[2], a[3], a[4], a[5],
                               c.tan1, c.tan2, c.tan3, c.tan4, n[0], n[1], n[2], n[3], n[4], n[5], gn[0], gn[1],
                               gn[2], gn[3], gn[4], gn[5], go[0], go[1], go[2], go[3], go[4], go[5], ga[0], ga[1],
                               ga[2], ga[3], ga[4], ga[5], gt[0], gt[1], gt[2], gt[3]);
        } else {
            adj_cell<float>(LH_CELL_ARGS);
            gt[0] = gmuc;
        }
#undef LH_CELL_ARGS
#pragma unroll
        for (int f = 0; f < 6; f++) sAB[f][r][cc] = ga[f];
        const int tr = r - HW, tc = cc - HW;
        if (tr >= 0 && tr < TY && tc >= 0 && tc < TX) {
#pragma unroll
            for (int f = 0; f < 6; f++) sL[f][tr][tc] = (double)go[f];
#pragma unroll
            for (int i = 0; i < NG; i++) gm[i] += (double)gt[i];
        }
    }
    __syncthreads();

    // 2. edge adjoints, classes 3, 2, 1, 0.  A class-cls edge joins box cells
    //    (r0, c0) and (r1, c1), or is the ghost edge of (r0, c0) on the domain
    //    boundary.
    auto edge_adj = [&](int cls, int r0, int c0, int r1, int c1, bool interior, float n1x, float n1y, float len,
                        float bd) {
        const int rr1 = interior ? r1 : r0, cc1 = interior ? c1 : c0;
        EdgeIn<float> in;
        in.h10 = sh1[r0][c0];
        in.h20 = sh2[r0][c0];
        in.h11 = sh1[rr1][cc1];
        in.h21 = sh2[rr1][cc1];
        in.H0 = sH[r0][c0];
        in.H1 = sH[rr1][cc1];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            in.q0[q] = sq[q][r0][c0];
            in.q1[q] = sq[q][rr1][cc1];
        }
        const bool xedge = cls >= 2;
        if (!interior) {
            const int nk = xedge ? 0 : 1;
            in.q1[nk] *= bd;
            in.q1[2 + nk] *= bd;
        }
        float a0h1 = 0.0f, a0h2 = 0.0f, a1h1 = 0.0f, a1h2 = 0.0f;
        if (cls > 0) {
            a0h1 = sP[cls - 1][0][r0][c0];
            a0h2 = sP[cls - 1][1][r0][c0];
            if (interior) {
                a1h1 = sP[cls - 1][0][r1][c1];
                a1h2 = sP[cls - 1][1][r1][c1];
            }
        }
        float gfm[6], gfp[6];
#pragma unroll
        for (int f = 0; f < 6; f++) {
            gfm[f] = -sAB[f][r0][c0];
            gfp[f] = interior ? -sAB[f][r1][c1] : 0.0f;
        }
        float fm[6], fp[6], gq0[4], gq1[4], ga0h1, ga0h2, ga1h1, ga1h2;
        double gh10, gh20, gh11, gh21;
#define LH_EDGE_ARGS                                                                                                \
    c.eps, c.eps4, c.g, c.gp, c.r, c.mu, c.dt, c.area, len, n1x, n1y, interior ? 1.0f : 0.0f, in.H0, in.H1, in.h10,   \
        in.h20, in.h11, in.h21, in.q0[0], in.q0[1], in.q0[2], in.q0[3], in.q1[0], in.q1[1], in.q1[2], in.q1[3], a0h1,  \
        a0h2, a1h1, a1h2, fm[0], fm[1], fm[2], fm[3], fm[4], fm[5], fp[0], fp[1], fp[2], fp[3], fp[4], fp[5], gfm[0], \
        gfm[1], gfm[2], gfm[3], gfm[4], gfm[5], gfp[0], gfp[1], gfp[2], gfp[3], gfp[4], gfp[5], gh10, gh20, gh11,     \
        gh21, gq0[0], gq0[1], gq0[2], gq0[3], gq1[0], gq1[1], gq1[2], gq1[3], ga0h1, ga0h2, ga1h1, ga1h2
        const bool bare = (in.h20 == 0.0) && (in.h21 == 0.0) && (in.q0[2] == 0.0f) && (in.q0[3] == 0.0f) &&
                          (in.q1[2] == 0.0f) && (in.q1[3] == 0.0f) && (a0h2 == 0.0f) && (a1h2 == 0.0f);
        // without sediment the Coulomb flag is false under either law
        if (__all(bare)) {
            adj_edge_ns<float>(LH_EDGE_ARGS);
        } else if constexpr (PQ) {
            adj_edge_pq<float>(c.eps, c.eps4, c.g, c.gp, c.r, c.mu, c.dt, c.area, len, n1x, n1y,
                               interior ? 1.0f : 0.0f, in.H0, in.H1, c.tan1, c.tan2, c.tan3, c.tan4, c.LoverH, in.h10,
                               in.h20, in.h11, in.h21, in.q0[0], in.q0[1], in.q0[2], in.q0[3], in.q1[0], in.q1[1],
                               in.q1[2], in.q1[3], a0h1, a0h2, a1h1, a1h2, fm[0], fm[1], fm[2], fm[3], fm[4], fm[5],
                               fp[0], fp[1], fp[2], fp[3], fp[4], fp[5], gfm[0], gfm[1], gfm[2], gfm[3], gfm[4],
                               gfm[5], gfp[0], gfp[1], gfp[2], gfp[3], gfp[4], gfp[5], gh10, gh20, gh11, gh21, gq0[0],
                               gq0[1], gq0[2], gq0[3], gq1[0], gq1[1], gq1[2], gq1[3], ga0h1, ga0h2, ga1h1, ga1h2);
        } else {
            adj_edge<float>(LH_EDGE_ARGS);
        }
#undef LH_EDGE_ARGS
        // lambda^n of the tile's own cells
        auto add_lambda = [&](int r, int cc, double dh1, double dh2, const double* dq) {
            const int tr = r - HW, tc = cc - HW;
            if (tr < 0 || tr >= TY || tc < 0 || tc >= TX) return;
            sL[0][tr][tc] += dh1;
            sL[3][tr][tc] += dh2;
            sL[1][tr][tc] += dq[0];
            sL[2][tr][tc] += dq[1];
            sL[4][tr][tc] += dq[2];
            sL[5][tr][tc] += dq[3];
        };
        if (interior) {
            const double d0[4] = {(double)gq0[0], (double)gq0[1], (double)gq0[2], (double)gq0[3]};
            const double d1[4] = {(double)gq1[0], (double)gq1[1], (double)gq1[2], (double)gq1[3]};
            add_lambda(r0, c0, gh10, gh20, d0);
            add_lambda(r1, c1, gh11, gh21, d1);
            sAB[0][r1][c1] += ga1h1;
            sAB[3][r1][c1] += ga1h2;
        } else {
            // the ghost copies cell 0 with its normal discharge times borde
            const int nk = xedge ? 0 : 1;
            double d0[4];
#pragma unroll
            for (int q = 0; q < 4; q++) d0[q] = (double)gq0[q] + (double)gq1[q] * (((q & 1) == nk) ? (double)bd : 1.0);
            add_lambda(r0, c0, gh10 + gh11, gh20 + gh21, d0);
        }
        sAB[0][r0][c0] += ga0h1;
        sAB[3][r0][c0] += ga0h2;
    };

#pragma unroll 1
    for (int cls = 3; cls >= 0; cls--) {
        if (cls >= 2) {
            // x-edges: box edge column cl joins box columns cl-1 and cl (global ie = bi + cl)
            const int par = cls - 2;
            const int first = ((bi + 1) & 1) == par ? 1 : 2;       // first cl >= 1 with (bi + cl) % 2 == par
            const int ncol = (CW - 1 - first) / 2 + 1;             // cl = first, first + 2, ... <= CW - 1
            for (int e = tid; e < ncol * RW; e += NT) {
                const int r = e / ncol, cl = first + 2 * (e % ncol);
                const int gj = bj + r, ie = bi + cl;
                if (gj < 0 || gj >= ny || ie < 0 || ie > nx) continue;
                if (ie == 0)
                    edge_adj(cls, r, cl, r, cl, false, -1.0f, 0.0f, (float)k.dy, (float)k.borde_izq);
                else if (ie == nx)
                    edge_adj(cls, r, cl - 1, r, cl - 1, false, 1.0f, 0.0f, (float)k.dy, (float)k.borde_der);
                else
                    edge_adj(cls, r, cl - 1, r, cl, true, 1.0f, 0.0f, (float)k.dy, 1.0f);
            }
            // ghost edges on the box's outer edge columns (cl = 0: ie = 0; cl = CW: ie = nx)
            const bool gl = (bi == 0) && (par == 0), gr = (bi + CW == nx) && ((nx & 1) == par);
            if (gl || gr)
                for (int r = tid; r < RW; r += NT) {
                    const int gj = bj + r;
                    if (gj < 0 || gj >= ny) continue;
                    if (gl) edge_adj(cls, r, 0, r, 0, false, -1.0f, 0.0f, (float)k.dy, (float)k.borde_izq);
                    if (gr) edge_adj(cls, r, CW - 1, r, CW - 1, false, 1.0f, 0.0f, (float)k.dy, (float)k.borde_der);
                }
        } else {
            // y-edges: box edge row rl joins box rows rl-1 and rl (global je = bj + rl)
            const int first = ((bj + 1) & 1) == cls ? 1 : 2;
            const int nrow = (RW - 1 - first) / 2 + 1;
            for (int e = tid; e < nrow * CW; e += NT) {
                const int cc = e % CW, rl = first + 2 * (e / CW);
                const int gi = bi + cc, je = bj + rl;
                if (gi < 0 || gi >= nx || je < 0 || je > ny) continue;
                if (je == 0)
                    edge_adj(cls, rl, cc, rl, cc, false, 0.0f, -1.0f, (float)k.dx, (float)k.borde_sup);
                else if (je == ny)
                    edge_adj(cls, rl - 1, cc, rl - 1, cc, false, 0.0f, 1.0f, (float)k.dx, (float)k.borde_inf);
                else
                    edge_adj(cls, rl - 1, cc, rl, cc, true, 0.0f, 1.0f, (float)k.dx, 1.0f);
            }
            // ghost edges on the box's outer edge rows (rl = 0: je = 0; rl = RW: je = ny)
            const bool gt = (bj == 0) && (cls == 0), gb = (bj + RW == ny) && ((ny & 1) == cls);
            if (gt || gb)
                for (int cc = tid; cc < CW; cc += NT) {
                    const int gi = bi + cc;
                    if (gi < 0 || gi >= nx) continue;
                    if (gt) edge_adj(cls, 0, cc, 0, cc, false, 0.0f, -1.0f, (float)k.dx, (float)k.borde_sup);
                    if (gb) edge_adj(cls, RW - 1, cc, RW - 1, cc, false, 0.0f, 1.0f, (float)k.dx, (float)k.borde_inf);
                }
        }
        __syncthreads();
    }

    // the loss gradient of the gauge series at step n (Gk, one per gauge) adds
    // to the h1 and h2 adjoints of the gauge cells in this tile (eta = h1 + h2
    // - H); one thread per cell sums the gauges that share it
    if (Gk) {
        for (int q = tid; q < ng; q += NT) {
            const int c = gauges[q], gj = c / nx, gi = c % nx;
            const int tr = gj - j0, tc = gi - i0;
            if (tr < 0 || tr >= TY || tc < 0 || tc >= TX) continue;
            bool first = true;
            for (int h = 0; h < q; h++)
                if (gauges[h] == c) first = false;
            if (!first) continue;
            double sum = 0.0;
            for (int h = q; h < ng; h++)
                if (gauges[h] == c) sum += Gk[h];
            sL[0][tr][tc] += sum;
            sL[3][tr][tc] += sum;
        }
        __syncthreads();
    }

    // 3. lambda^n of the tile's own cells, d/dmu
    for (int e = tid; e < TX * TY; e += NT) {
        const int tr = e / TX, tc = e % TX;
        const int gj = j0 + tr, gi = i0 + tc;
        if (gj >= ny || gi >= nx) continue;
        const size_t g = (size_t)gj * nx + gi;
#pragma unroll
        for (int f = 0; f < 6; f++) LOUT[f * N + g] = sL[f][tr][tc];
    }
#pragma unroll
    for (int i = 0; i < NG; i++) {
        double v = gm[i];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((tid & 63) == 0) red[i][tid >> 6] = v;
    }
    __syncthreads();
    if (tid == 0) {
        for (int i = 0; i < NG; i++) {
            double s = 0.0;
            for (int w = 0; w < NT / 64; w++) s += red[i][w];
            if (s != 0.0) gmu[i * gstride + blockIdx.x] += s;   // this workgroup's slot
        }
    }
}

}  // namespace adj

// ----------------------------------------------------------------- driver
// Binomial checkpointing: with c snapshots (the range start included) the
// fewest re-advances for l steps is T(l, c) = t l - C(c+1+t-1, t-1), t minimal
// with C(c+t, t) >= l (Griewank & Walther, "revolve", ACM TOMS 26, 2000);
// the first snapshot goes to the split m minimising m + T(l-m, c-1) + T(m, c).
static double binom(int c, long t)
{
    if (t < 0) return 0.0;
    double v = 1.0;
    for (long i = 1; i <= t; i++) {
        v = v * (double)(c + i) / (double)i;
        if (v > 1e18) return 1e18;
    }
    return v;
}

long revolve_advances(long l, int c)
{
    if (l <= 1) return 0;
    if (c <= 1) return l * (l - 1) / 2;
    long t = 0;
    while (binom(c, t) < (double)l) t++;
    return (long)(t * l - binom(c + 1, t - 1));
}

long revolve_split(long l, int c)
{
    long best = 1, bestv = -1;
    for (long m = 1; m < l; m++) {
        long v = m + revolve_advances(l - m, c - 1) + revolve_advances(m, c);
        if (bestv < 0 || v < bestv) {
            bestv = v;
            best = m;
        }
    }
    return best;
}

// reverse steps end-1 ... start; the state at start is in slot (S - c)
static hipError_t treeverse(AdjointHooks& h, int S, long start, long end, int c, bool on_chain)
{
    hipError_t e;
    const long l = end - start;
    if (l <= 0) return hipSuccess;
    if (l == 1) {
        if ((e = h.restore(h.ctx, S - c)) != hipSuccess) return e;
        h.reverses++;
        return h.reverse(h.ctx, (int)start);
    }
    if (c <= 1) {
        for (long k = end - 1; k >= start; k--) {
            if ((e = h.restore(h.ctx, S - c)) != hipSuccess) return e;
            if (k > start) {
                if ((e = h.advance(h.ctx, (int)start, (int)(k - start))) != hipSuccess) return e;
                h.advances += k - start;
            }
            h.reverses++;
            if ((e = h.reverse(h.ctx, (int)k)) != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const long m = revolve_split(l, c);
    if (!on_chain) {
        if ((e = h.restore(h.ctx, S - c)) != hipSuccess) return e;
        if ((e = h.advance(h.ctx, (int)start, (int)m)) != hipSuccess) return e;
        h.advances += m;
        if ((e = h.store(h.ctx, S - c + 1)) != hipSuccess) return e;
    }
    if ((e = treeverse(h, S, start + m, end, c - 1, on_chain)) != hipSuccess) return e;
    return treeverse(h, S, start, start + m, c, false);
}

// positions of the snapshots the first descent stores: (step, slot)
void revolve_chain(long N, int S, std::vector<std::pair<long, int>>& out)
{
    long x = 0;
    int c = S;
    while (N - x > 1 && c > 1) {
        long m = revolve_split(N - x, c);
        x += m;
        out.push_back({x, S - c + 1});
        c--;
    }
}

hipError_t run_treeverse(AdjointHooks& h, int S, long N) { return treeverse(h, S, 0, N, S, true); }

// tile of the fused FP32 reverse step: the (TX + 4) x (TY + 4) box holds 1008
// cells (two per thread of 512) and every edge class of the box at most 504
// edges (one pass); of the shapes with those properties this one recomputes
// the fewest box cells per own cell (1.33, against 1.42 for 60 x 12)
constexpr int kAdjTX = 38, kAdjTY = 20;

// launch wrappers for one reverse step (four-pass structure)
template <typename F, bool PQ>
hipError_t reverse_step_t(const Consts& k, const F* W, const double* H, F* P, F* ACC, F* AB, double* LAM, double* gmu,
                          size_t gstride, const double* dt_hist, int step, hipStream_t s)
{
    const size_t N = (size_t)k.nx * k.ny;
    adj::SV<F> sv{W, N};
    adj::Work<F> w{P, ACC, AB, LAM, gmu, gstride};
    long ne[4];
    for (int cls = 0; cls < 4; cls++)
        ne[cls] = (cls < 2) ? (long)((k.ny - cls) / 2 + 1) * k.nx : (long)((k.nx - (cls - 2)) / 2 + 1) * k.ny;
    // the forward step's own tile kernel recomputes the partial and final sums
    hipError_t e;
    if constexpr (sizeof(F) == 4)
        e = f32c_recompute(k, W, H, dt_hist, step, P, ACC, s);
    else
        e = f64_recompute(k, W, H, dt_hist, step, P, ACC, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((adj::k_cell_adj<F, PQ>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, k, sv, w, dt_hist,
                       step);
    for (int cls = 3; cls >= 0; cls--)
        hipLaunchKernelGGL((adj::k_edge_adj<F, PQ>), dim3((unsigned)((ne[cls] + 255) / 256)), dim3(256), 0, s, k, sv,
                           H, w, dt_hist, step, cls);
    return hipGetLastError();
}

hipError_t adj_reverse_step_f32(const Consts& k, const float* W, const double* H, float* P, float* ACC, float* AB,
                                double* LAM, double* gmu, size_t gstride, const double* dt_hist, int step,
                                hipStream_t s)
{
    if (k.friction_law == 1)
        return reverse_step_t<float, true>(k, W, H, P, ACC, AB, LAM, gmu, gstride, dt_hist, step, s);
    return reverse_step_t<float, false>(k, W, H, P, ACC, AB, LAM, gmu, gstride, dt_hist, step, s);
}

hipError_t adj_reverse_step_f32_fused(const Consts& k, const float* W, const double* H, float* P, float* ACC,
                                      const double* LAMin, double* LAMout, double* gmu, size_t gstride,
                                      const double* dt_hist, int step, const int* gauges, int ng, const double* Gk,
                                      hipStream_t s)
{
    constexpr int TX = kAdjTX, TY = kAdjTY, NT = 512;
    hipError_t e = f32c_recompute(k, W, H, dt_hist, step, P, ACC, s);
    if (e != hipSuccess) return e;
    const int tiles_x = (k.nx + TX - 1) / TX, tiles_y = (k.ny + TY - 1) / TY;
    const dim3 grid((unsigned)(tiles_x * tiles_y));
    if (k.friction_law == 1)
        hipLaunchKernelGGL((adj::k_adj_fused<TX, TY, NT, true>), grid, dim3(NT), 0, s, k, W, H, P, ACC, LAMin, LAMout,
                           gmu, gstride, dt_hist, step, tiles_x, tiles_y, gauges, ng > 0 ? ng : 0,
                           ng > 0 ? Gk : nullptr);
    else
        hipLaunchKernelGGL((adj::k_adj_fused<TX, TY, NT, false>), grid, dim3(NT), 0, s, k, W, H, P, ACC, LAMin, LAMout,
                           gmu, gstride, dt_hist, step, tiles_x, tiles_y, gauges, ng > 0 ? ng : 0,
                           ng > 0 ? Gk : nullptr);
    return hipGetLastError();
}

hipError_t adj_reverse_step_f64(const Consts& k, const double* W, const double* H, double* P, double* ACC, double* AB,
                                double* LAM, double* gmu, size_t gstride, const double* dt_hist, int step,
                                hipStream_t s)
{
    if (k.friction_law == 1)
        return reverse_step_t<double, true>(k, W, H, P, ACC, AB, LAM, gmu, gstride, dt_hist, step, s);
    return reverse_step_t<double, false>(k, W, H, P, ACC, AB, LAM, gmu, gstride, dt_hist, step, s);
}

size_t adj_gmu_slots(const Consts& k)
{
    const size_t N = (size_t)k.nx * k.ny;
    const size_t cells = (N + 255) / 256;
    const size_t tiles = (size_t)((k.nx + kAdjTX - 1) / kAdjTX) * ((k.ny + kAdjTY - 1) / kAdjTY);
    return cells > tiles ? cells : tiles;
}

hipError_t adj_inject(double* LAM, size_t N, const int* gauges, int ng, const double* G, hipStream_t s)
{
    if (ng <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj::k_inject, dim3((ng + 63) / 64), dim3(64), 0, s, LAM, N, gauges, ng, G);
    return hipGetLastError();
}

hipError_t adj_set_step(Ctl* ctl, int step, hipStream_t s)
{
    hipLaunchKernelGGL(adj::k_set_step, dim3(1), dim3(64), 0, s, ctl, step);
    return hipGetLastError();
}

}  // namespace lhysea
