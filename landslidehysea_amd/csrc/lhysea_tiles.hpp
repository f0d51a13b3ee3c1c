// Tile classes of the split step, shared by the FP32C and FP64 kernels.
//
// The grid is cut into CTX x CTY class tiles.  A class tile is general at
// step s when its tile + 1-cell halo box holds a cell with sediment (h2 or q2
// nonzero) in the state of step s: its own cells, or the edge and corner
// cells of its neighbours that form its halo.  cls[(s & 1) * nt + t] == s
// marks class tile t general at step s and list[(s & 1) * nt + i],
// i < ctl->gcount[s & 1], enumerates those tiles; the sediment-free kernel
// skips them and the general kernel takes them.  A cell that keeps sediment
// through a general update marks, for step s + 1 (the other parity, so the
// running step's entries stay intact), its own tile and the neighbours whose
// halo it lies in.  A sediment-free tile cannot gain sediment in one step (no
// edge of its box carries any), so it marks nothing.
#pragma once
#include "lhysea_common.hpp"

namespace lhysea {

struct TileCls {
    int* cls;
    int* list;
    int tiles_x, tiles_y;
};

// Clears a class table before a classification: n entries -1 (no tile
// general at any step) and both general counts 0.  A kernel rather than a
// memset, so that a launch sequence that contains it can be captured into a
// graph and replayed.
template <int NT>
__global__ void __launch_bounds__(NT) k_clear_classes(int* __restrict__ cls, int n, Ctl* ctl)
{
    for (int i = (int)(blockIdx.x * NT + threadIdx.x); i < n; i += (int)(gridDim.x * NT)) cls[i] = -1;
    if (blockIdx.x == 0 && threadIdx.x < 2) ctl->gcount[threadIdx.x] = 0u;
}

inline hipError_t clear_classes(int* cls, int n, Ctl* ctl, hipStream_t s)
{
    const int b = (n + 255) / 256;
    hipLaunchKernelGGL(k_clear_classes<256>, dim3((unsigned)(b < 1 ? 1 : (b > 1024 ? 1024 : b))), dim3(256), 0, s,
                       cls, n, ctl);
    return hipGetLastError();
}

// Class tiles whose tile + halo box holds the cell at (lx, ly) of a CTX x CTY
// class tile, as bits (dy + 1) * 3 + (dx + 1) for the neighbour at (dx, dy):
// the tile itself and the neighbours across the edges and corners the cell
// lies on.
template <int CTX, int CTY>
__device__ __forceinline__ int halo_mask_of(int lx, int ly)
{
    const int x0 = (lx == 0) ? 0 : 1, x1 = (lx == CTX - 1) ? 2 : 1;
    const int y0 = (ly == 0) ? 0 : 1, y1 = (ly == CTY - 1) ? 2 : 1;
    int m = 0;
    for (int y = y0; y <= y1; y++)
        for (int x = x0; x <= x1; x++) m |= 1 << (3 * y + x);
    return m;
}

// Marks, for step s, the class tiles of the 3 x 3 neighbourhood of (tx, ty)
// whose bit is set in mask.  Called by the first 9 threads of a wavefront
// with the same mask, thread b taking bit b, so the atomics of the nine
// tiles are in flight together.
__device__ __forceinline__ void mark_general(const TileCls& tc, Ctl* ctl, int tx, int ty, int s, int mask)
{
    const int bit = threadIdx.x & 63;
    if (bit >= 9 || !(mask & (1 << bit))) return;
    const int nt = tc.tiles_x * tc.tiles_y, base = (s & 1) * nt;
    const int x = tx + bit % 3 - 1, y = ty + bit / 3 - 1;
    if (x < 0 || y < 0 || x >= tc.tiles_x || y >= tc.tiles_y) return;
    const int t = y * tc.tiles_x + x;
    // the first marker of t for step s appends it
    if (atomicMax(&tc.cls[base + t], s) < s) tc.list[base + (int)atomicAdd(&ctl->gcount[s & 1], 1u)] = t;
}

// Block-wide OR of m; every thread calls it.
__device__ __forceinline__ int block_or(int m)
{
    __shared__ int acc;
    if (threadIdx.x == 0) acc = 0;
    __syncthreads();
    if (m) atomicOr(&acc, m);
    __syncthreads();
    const int r = acc;
    __syncthreads();
    return r;
}

// Workgroup end of a general tile: the atomicMax of the wave-speed max v and
// the marks of the next step's general tiles from the OR of the halo masks m
// (one barrier for both reductions).
template <int NT>
__device__ __forceinline__ void general_tile_end(double v, int m, const TileCls& tc, Ctl* ctl, int tx, int ty,
                                                 int s_next)
{
    __shared__ double red[NT / 64];
    __shared__ int ored[NT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        v = fmax(v, __shfl_xor(v, off, 64));
        m |= __shfl_xor(m, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[wid] = v;
        ored[wid] = m;
    }
    __syncthreads();
    if (wid == 0) {
        // every lane of wavefront 0 forms both reductions
        for (int i = 1; i < NT / 64; i++) {
            v = fmax(v, red[i]);
            m |= ored[i];
        }
        if (lane == 0 && v > 0.0) atomicMax(&ctl->maxA_next, (unsigned long long)__double_as_longlong(v));
        if (m) mark_general(tc, ctl, tx, ty, s_next, m);
    }
}

// Gauges inside a tile (columns below xend): eta = h1 + h2 - H from the new
// h1, h2 (zero when sh2 is null) and the depth the block staged in LDS
// (row-major planes of row length cwp, the tile's cell (r, c) at row r + 1,
// column c + 1), after a barrier that publishes the threads' new values.
template <int TX, int TY, int NT>
__device__ __forceinline__ void gauges_of_tile(const GaugeOut& go, const double* sh1, const double* sh2,
                                               const double* sH, int cwp, int nx, int i0, int j0, int xend)
{
    __syncthreads();
    for (int gi = threadIdx.x; gi < go.ng; gi += NT) {
        const int cell = go.cells[gi];
        const int r = cell / nx - j0, c = cell % nx - i0;
        if (r >= 0 && r < TY && c >= 0 && c < TX && i0 + c < xend) {
            const int at = (r + 1) * cwp + c + 1;
            go.out[gi] = (sh1[at] + (sh2 ? sh2[at] : 0.0)) - sH[at];
        }
    }
}

}  // namespace lhysea
