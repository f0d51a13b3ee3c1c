// FP32 device pipeline of the LandSlide-HySEA step with a compensated state.
//
// Arithmetic split:
//  * FP64: depth H, the hydrostatic reconstructions h1+h2-H+Hm and h2-H+Hm
//    (lake-at-rest cancellation), the dry/wet depth comparisons, dt, time.
//  * FP32: the Riemann solver (eigenvalue estimates, PVM viscosity, fluxes,
//    tangential upwinding, positivity) and the friction coefficients, with
//    hardware reciprocal / square root / exp.
//  * State: each field is an unevaluated FP32 pair (value, compensation);
//    every operation that changes the state (flux increment, damping,
//    implicit friction, Coulomb scaling) is linear in the state and is applied
//    with error-free transformations (TwoSum, FMA-based TwoProd).
//
// Algebraic identities used (exact in real arithmetic):
//  * U(h,q) = sqrt2 h q / sqrt(h^4 + max(h,eps)^4) = q * s(h) with
//    s(h) = 1/h for h >= eps and sqrt2 h / sqrt(h^4 + eps^4) otherwise;
//  * in a single edge the mass fluxes satisfy F+[h] = -F-[h]
//    (src/GPU/Arista_kernel.cu:496-504), used to make mass exactly conservative.
#pragma once
#include "lhysea_common.hpp"

namespace lhysea {
namespace f32 {

constexpr float SQRT2F = 1.41421356237309505f;

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fexp(float x) { return __expf(x); }
// x^y for x >= 0 as exp2(y log2 x) with the hardware log2 / exp2 (x = 0 gives 0)
__device__ __forceinline__ float fpow(float x, float y)
{
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// s(h): U(h,q) = q * s(h) = sqrt2 h / sqrt(h^4 + max(h,eps)^4)  (= 1/h for h >= eps);
// one reciprocal square root, no branch
__device__ __forceinline__ float sdes(float h, float eps, float eps4)
{
    float h2 = h * h, m = fmaxf(h, eps), m2 = m * m;
    return SQRT2F * h * rsq(h2 * h2 + m2 * m2);
}

// h * s(h): the factor by which q -> U(h,q) * h shrinks a discharge; exactly 1
// for h >= eps, sqrt2 h^2 / sqrt(h^4 + eps^4) < 1 below
__device__ __forceinline__ float hsdes(float h, float eps, float eps4)
{
    if (h >= eps) return 1.0f;
    float h2 = h * h;
    return SQRT2F * h2 * rsq(h2 * h2 + eps4);
}

// Two-layer eigenvalue estimate + the reference's compare-swap network
// (src/GPU/Arista_kernel.cu:152-237), FLAG 0: (h1,u1,h2,u2), FLAG 1: (h1,q1,h2,q2).
// s0, s2: s(f0), s(f2) of the layers (FLAG 1 only); igp = 1/gp
template <int FLAG>
__device__ __forceinline__ void eig(float f0, float f1, float f2, float f3, float s0, float s2, float g, float gp,
                                    float igp, float eps, float eps4, float& Dx, float& Dy, float& Dz, float& Dw)
{
    float h = f0 + f2;
    float sh = sdes(h, eps, eps4);
    float q = (FLAG == 1) ? (f1 + f3) : (f0 * f1 + f2 * f3);
    float u = q * sh;
    float cg = fsqrt(g * h);
    float uu, u1, u2;
    Dx = u - cg;
    Dw = u + cg;
    if (FLAG == 0) {
        uu = (f0 * f3 + f1 * f2) * sh;
        u1 = f1;
        u2 = f3;
    } else {
        u1 = f1 * s0;
        u2 = f3 * s2;
        uu = (u1 * f2 + u2 * f0) * sh;
    }
    float du = u1 - u2;
    float aux = 1.0f - du * du * sh * igp;
    float ci = fsqrt(gp * f0 * f2 * sh * fabsf(aux));
    Dy = uu - ci;
    Dz = uu + ci;
    float lo, hi;
    lo = fminf(Dx, Dy); hi = fmaxf(Dx, Dy); Dx = lo; Dy = hi;
    lo = fminf(Dx, Dz); hi = fmaxf(Dx, Dz); Dx = lo; Dz = hi;
    lo = fminf(Dx, Dw); hi = fmaxf(Dx, Dw); Dx = lo; Dw = hi;
    lo = fminf(Dy, Dz); hi = fmaxf(Dy, Dz); Dy = lo; Dz = hi;
    lo = fminf(Dy, Dw); hi = fmaxf(Dy, Dw); Dy = lo; Dw = hi;
    lo = fminf(Dy, Dz); hi = fmaxf(Dy, Dz); Dy = hi; Dz = lo;   // inverted (2,1) step
    lo = fminf(Dz, Dw); hi = fmaxf(Dz, Dw); Dz = lo; Dw = hi;
}

struct EdgeFlux {
    float Fm[6], Fp[6];
    float lam;
};

// Per-run FP32 constants.
struct K32 {
    float eps, eps4, g, gp, r, mu, peso, dtf, area, len_x, len_y;
    float igp, ten_eps, idt_eps, bq;   // 1/gp, 10/eps, 1/(dt+EPS), dt/area
    int pouliquen;                     // friction law 1: the Pouliquen coefficient replaces mu
    float tan1, tan2, tan3, tan4, LoverH;
};

// Pouliquen friction coefficient (src/GPU/Arista_kernel.cu:20-58) in FP32:
// the reduced-gravity Froude numbers u_i^2 s(h_i) / g' with s the
// desingularised 1/h, combined as sqrt(Fr1 + Fr2 + (1 - rr) Fr1 Fr2); the
// stop / start coefficients from the four angles' tangents.
template <typename KC>
__device__ __forceinline__ float friction32(const KC& k, float h1, float u1, float h2, float u2)
{
    const float x = h1 * k.ten_eps;
    const float ex = (x < 9.5f) ? fexp(-(x * x)) : 0.0f;
    const float rr = 1.0f - k.r * (1.0f - ex);
    const float igp = rcp(k.g * rr);
    const float fr1 = u1 * u1 * sdes(h1, k.eps, k.eps4) * igp;
    const float fr2 = u2 * u2 * sdes(h2, k.eps, k.eps4) * igp;
    const float fr = fsqrt(fr1 + fr2 + (1.0f - rr) * fr1 * fr2);
    const float beta = 0.136f, L2 = 8.0e-4f, chi = 1.0e-3f;
    const float dstop = k.tan2 - k.tan1, dstart = k.tan4 - k.tan3;
    const float den = rcp(1.0f + h2 * (1.0f / L2));
    float muf;
    if (fr > beta)
        muf = k.tan1 + dstop * rcp(1.0f + beta * h2 * rcp(fr * L2));
    else if (fabsf(fr) < LH_EPSF)
        muf = k.tan3 + dstart * den;
    else {
        const float mustop = k.tan1 + dstop * den, mustart = k.tan3 + dstart * den;
        muf = mustart + fpow(fr * (1.0f / beta), chi) * (mustop - mustart);
    }
    return muf * k.LoverH;
}

// Edge normals are axis-aligned: AX 0 is an x-edge with normal (sn, 0), AX 1
// a y-edge with normal (0, sn), sn = +-1.  With the normal's zero component
// dropped the projections and rotations below equal the general forms
// (q . n, q . t with t = (-ny, nx), and back) exactly, up to the sign of
// zero results; the tangential flux F[h] u.t does not depend on sn (the two
// sign flips cancel), so it is formed without it.
// Normal component of the discharge (qx, qy) = (W[f], W[f + 1]).
template <int AX>
__device__ __forceinline__ float qn(const float* W, int f, float sn)
{
    return sn * W[f + AX];
}

// The edge's fluxes in the x/y frame from the normal fluxes Fm, Fp, the
// tangential flux Ft = F[h] u.t (all per unit length) and lsn = len * sn:
// n-frame (F, Ft) -> (F nx - Ft ny, F ny + Ft nx), and F+ rotated with -Ft.
template <int AX>
__device__ __forceinline__ void rotate_out(float* Mn, float* P, float Fm, float Fp, float Ft, float len, float lsn)
{
    const float lt = len * Ft;
    Mn[1 + AX] = lsn * Fm;
    Mn[2 - AX] = lt;
    P[1 + AX] = lsn * Fp;
    P[2 - AX] = -lt;
}

// One edge (src/GPU/Arista_kernel.cu:324-645) with FP32 arithmetic and FP64
// hydrostatic terms.  W0/W1: FP32 states (ghost when !interior); hd*: full
// precision h1/h2 of both cells; H0/H1 depths; normal of axis AX and sign sn.
template <int AX>
__device__ __forceinline__ bool edge(const K32& k, const float* W0, const float* W1, double h10, double h20,
                                     double h11, double h21, double H0, double H1, float sn, float len, float a0h1,
                                     float a0h2, float a1h1, float a1h2, bool interior, EdgeFlux& out)
{
    const float eps = k.eps, eps4 = k.eps4, g = k.g, r = k.r;
    float w00 = W0[0], w01 = qn<AX>(W0, 1, sn), w02 = W0[3], w03 = qn<AX>(W0, 4, sn);
    float w10 = W1[0], w11 = qn<AX>(W1, 1, sn), w12 = W1[3], w13 = qn<AX>(W1, 4, sn);

    // dry/wet treatment with FP64 depth comparisons (src/GPU/Arista_kernel.cu:301-321)
    if ((w02 < eps) && (H1 - h21 > H0)) w13 = 0.0f;
    if ((w12 < eps) && (H0 - h20 > H1)) w03 = 0.0f;
    if ((w00 < eps) && (H1 - h11 - h21 > H0 - h20)) w11 = 0.0f;
    if ((w10 < eps) && (H0 - h10 - h20 > H1 - h21)) w01 = 0.0f;
    if ((w02 + w00 < eps) && (H1 - h11 - h21 > H0)) { w11 = 0.0f; w13 = 0.0f; }
    if ((w12 + w10 < eps) && (H0 - h10 - h20 > H1)) { w01 = 0.0f; w03 = 0.0f; }

    float h1ij = 0.5f * (w00 + w10);
    float h2ij = 0.5f * (w02 + w12);
    float s00 = sdes(w00, eps, eps4), s10 = sdes(w10, eps, eps4);
    float s02 = sdes(w02, eps, eps4), s12 = sdes(w12, eps, eps4);
    float r00 = fsqrt(w00), r10 = fsqrt(w10), r02 = fsqrt(w02), r12 = fsqrt(w12);
    float u1ij = (r00 * (w01 * s00) + r10 * (w11 * s10)) * rcp(r00 + r10 + LH_EPSF);
    float u2ij = (r02 * (w03 * s02) + r12 * (w13 * s12)) * rcp(r02 + r12 + LH_EPSF);

    if (!((h1ij >= LH_EPSF) || (h2ij >= LH_EPSF))) return false;

    // hydrostatic reconstruction in FP64 (src/GPU/Arista_kernel.cu:94-150)
    const double Hm = fmin(H0, H1);
    const double e10 = fmax(h10 + h20 - H0 + Hm, 0.0), e11 = fmax(h11 + h21 - H1 + Hm, 0.0);
    const double e20 = fmax(h20 - H0 + Hm, 0.0), e21 = fmax(h21 - H1 + Hm, 0.0);
    const float deta1 = (float)(e11 - e10), deta2 = (float)(e21 - e20);

    float tp1 = g * h1ij * deta1;
    float tp3 = g * h2ij * ((1.0f - r) * deta2 + r * deta1);
    float x = h1ij * k.ten_eps;
    float ex = 0.0f;
    if (x < 9.5f) ex = fexp(-(x * x));         // exp(-x^2) is 0 in FP32 beyond x = 9.4
    float fsc = 1.0f - r * (1.0f - ex);
    const float muc = k.pouliquen ? friction32(k, h1ij, u1ij, h2ij, u2ij) : k.mu;
    float sc = fsc * muc * g * h2ij;
    bool coul = fabsf(h2ij * u2ij) < k.peso * sc * k.dtf;
    float t23 = coul ? (g * h2ij * r * deta1) : tp3;

    float Dx, Dy, Dz, Dw, Lx, Ly, Lz, Lw, Rx, Ry, Rz, Rw;
    eig<0>(h1ij, u1ij, h2ij, u2ij, 0.0f, 0.0f, g, k.gp, k.igp, eps, eps4, Dx, Dy, Dz, Dw);
    eig<1>(w00, w01, w02, w03, s00, s02, g, k.gp, k.igp, eps, eps4, Lx, Ly, Lz, Lw);
    eig<1>(w10, w11, w12, w13, s10, s12, g, k.gp, k.igp, eps, eps4, Rx, Ry, Rz, Rw);
    float aut1 = fminf(Dx, fminf(Lx, Rx));
    float aut3 = fmaxf(Dw, fmaxf(Lw, Rw));
    float maxl = fmaxf(fmaxf(fabsf(aut3), fabsf(aut1)), fmaxf(fabsf(u1ij), fabsf(u2ij)));
    float aut2 = fmaxf(fabsf(Dy), fabsf(Dz));
    float sg = aut1 + aut3;
    if (!(fabsf(sg) < LH_EPSF) && !(sg > 0.0f)) aut2 = -aut2;

    // PVM coefficients (src/GPU/Arista_kernel.cu:446-465)
    float c0, c1, c2;
    float d13 = aut1 - aut3, d12 = aut1 - aut2, d23 = aut2 - aut3;
    if ((fabsf(d13) >= LH_EPSF) && (fabsf(d12) >= LH_EPSF) && (fabsf(d23) >= LH_EPSF)) {
        if ((h1ij >= eps) && (h2ij >= eps)) {
            // A = |a1|/(d13 d12), B = -|a2|/(d12 d23), C = |a3|/(d13 d23) over one reciprocal
            float iD = rcp(d13 * d12 * d23);
            float A = fabsf(aut1) * d23 * iD;
            float B = -fabsf(aut2) * d13 * iD;
            float C = fabsf(aut3) * d12 * iD;
            c0 = A * aut2 * aut3 + B * aut1 * aut3 + C * aut1 * aut2;
            c1 = -A * (aut2 + aut3) - B * (aut1 + aut3) - C * (aut1 + aut2);
            c2 = A + B + C;
        } else {
            float id = rcp(-d13);
            c0 = (aut3 * fabsf(aut1) - aut1 * fabsf(aut3)) * id;
            c1 = (fabsf(aut3) - fabsf(aut1)) * id;
            c2 = 0.0f;
        }
    } else {
        c0 = maxl;
        c1 = 0.0f;
        c2 = 0.0f;
    }

    // physical fluxes and path-conservative viscosity (src/GPU/Arista_kernel.cu:467-504)
    float FL1 = (w00 < LH_EPSF) ? 0.0f : w01 * w01 * s00;
    float FL3 = (w02 < LH_EPSF) ? 0.0f : w03 * w03 * s02;
    float FR1 = (w10 < LH_EPSF) ? 0.0f : w11 * w11 * s10;
    float FR3 = (w12 < LH_EPSF) ? 0.0f : w13 * w13 * s12;
    float dF0 = w11 - w01, dF1 = FR1 - FL1, dF2 = w13 - w03, dF3 = FR3 - FL3;
    float t20 = dF0, t21 = tp1 + dF1, t22 = dF2;
    t23 = t23 + dF3;
    float ga = g * h1ij, gb = g * h2ij;
    float A0 = t21;
    float A1 = (ga - u1ij * u1ij) * t20 + 2.0f * u1ij * t21 + ga * t22;
    float A2 = t23;
    float A3 = r * gb * t20 + (gb - u2ij * u2ij) * t22 + 2.0f * u2ij * t23;
    float I0 = deta1, I2 = 0.0f;
    if (!coul) {
        if ((w00 > eps) && (w10 > eps)) I0 -= deta2;
        I2 = deta2;
    }
    // identity modification: I = (I0, dq1n, I2, dq2n) with dq1n = dF0, dq2n = dF2
    float V0 = 0.5f * (c0 * I0 + c1 * t20 + c2 * A0);
    float V1 = 0.5f * (c0 * dF0 + c1 * t21 + c2 * A1);
    float V2 = 0.5f * (c0 * I2 + c1 * t22 + c2 * A2);
    float V3 = 0.5f * (c0 * dF2 + c1 * t23 + c2 * A3);
    float M0 = 0.5f * dF0, M1 = 0.5f * (dF1 + tp1), M2 = 0.5f * dF2, M3 = 0.5f * (dF3 + tp3);
    float Fm0 = M0 + (w01 - V0), Fm1 = M1 + (FL1 - V1), Fm2 = M2 + (w03 - V2), Fm3 = M3 + (FL3 - V3);
    float Fp1 = M1 + (V1 - FR1), Fp3 = M3 + (V3 - FR3);

    // tangential upwinding (src/GPU/Arista_kernel.cu:506-532): the upwind
    // cell's tangential velocity, up to the normal's sign (see rotate_out)
    float ut1 = 0.0f, ut2 = 0.0f;
    if (!(fabsf(Fm0) < LH_EPSF)) ut1 = (Fm0 > 0.0f) ? W0[2 - AX] * s00 : W1[2 - AX] * s10;
    if (!(fabsf(Fm2) < LH_EPSF)) ut2 = (Fm2 > 0.0f) ? W0[5 - AX] * s02 : W1[5 - AX] * s12;
    float* P = out.Fp;
    float* Mn = out.Fm;
    const float lsn = len * sn;
    Mn[0] = len * Fm0;
    Mn[3] = len * Fm2;
    P[0] = -Mn[0];                                  // F+[h] = -F-[h]
    P[3] = -Mn[3];
    rotate_out<AX>(Mn, P, Fm1, Fp1, Fm0 * ut1, len, lsn);
    rotate_out<AX>(Mn + 3, P + 3, Fm3, Fp3, Fm2 * ut2, len, lsn);

    // positivity with the partial accumulators (src/GPU/Arista_kernel.cu:554-609)
    // P[h] = -Mn[h]: at most one of the two outflow candidates exists per layer
    const float area = k.area, dt = k.dtf, bq = k.bq;
    float dt0 = 1e30f, dt1 = 1e30f;
    {
        const bool out0 = Mn[0] > 0.0f, in0 = interior && (Mn[0] < 0.0f);
        if (out0 || in0) dt0 = (out0 ? (W0[0] + bq * a0h1) : (W1[0] + bq * a1h1)) * area * rcp(fabsf(Mn[0]) + LH_EPSF);
        const bool out1 = Mn[3] > 0.0f, in1 = interior && (Mn[3] < 0.0f);
        if (out1 || in1) dt1 = (out1 ? (W0[3] + bq * a0h2) : (W1[3] + bq * a1h2)) * area * rcp(fabsf(Mn[3]) + LH_EPSF);
    }
    if (!(dt <= dt0)) {
        float al = dt0 * k.idt_eps;
        Mn[0] *= al; Mn[1] *= al; Mn[2] *= al;
        P[0] *= al; P[1] *= al; P[2] *= al;
    }
    if (!(dt <= dt1)) {
        float al = dt1 * k.idt_eps;
        Mn[3] *= al; Mn[4] *= al; Mn[5] *= al;
        P[3] *= al; P[4] *= al; P[5] *= al;
    }
    if (maxl < eps) maxl += eps;
    out.lam = len * maxl * rcp(k.peso);
    return true;
}

// Sorting network of eig() for a state whose internal pair is zero: on
// (Dx, 0, 0, Dw) with Dx <= Dw (Dx = u - c, Dw = u + c, c >= 0) the seven
// compare-swaps, the inverted (2,1) step included, give Dx' = min(Dx, 0),
// Dw' = max(Dw, 0) and an inner pair with max(|Dy'|, |Dz'|) = max(Dx, -Dw, 0)
// (case by case: Dx >= 0 gives (0, Dx, 0, Dw); Dw <= 0 gives (Dx, 0, Dw, 0);
// otherwise (Dx, 0, 0, Dw)).
__device__ __forceinline__ void net_zero_pair(float& Dx, float& Dw, float* aut2)
{
    if (aut2) *aut2 = fmaxf(fmaxf(Dx, -Dw), 0.0f);
    Dx = fminf(Dx, 0.0f);
    Dw = fmaxf(Dw, 0.0f);
}

// edge() for an edge whose two cells carry no sediment (h2 = 0 and q2 = 0 on
// both sides, h2 exactly 0 in FP64 too).  Every term of edge() that such an
// edge makes zero is dropped and the rest is evaluated in the same order, so
// the fluxes equal edge()'s up to the sign of zero results; the layer-2
// fluxes are zero and are not produced (Fm[3..5], Fp[3..5] left unset):
//  * dry/wet: the layer-2 tests only zero momenta that are already 0, the
//    mixed tests reduce to the layer-1 ones;
//  * h2ij = u2ij = 0, so tp3, the Coulomb flag (false), deta2 (0) and the
//    internal eigenvalue pair (0, 0) vanish; the PVM takes the two-wave or
//    Rusanov form because h2ij < eps;
//  * eig(FLAG 1) of a cell uses s(f0 + f2) = s(w) of that cell's water.
template <int AX>
__device__ __forceinline__ bool edge1(const K32& k, const float* W0, const float* W1, double h10, double h11,
                                      double H0, double H1, float sn, float len, float a0h1, float a1h1,
                                      bool interior, EdgeFlux& out)
{
    const float eps = k.eps, eps4 = k.eps4, g = k.g;
    float w00 = W0[0], w01 = qn<AX>(W0, 1, sn);
    float w10 = W1[0], w11 = qn<AX>(W1, 1, sn);

    // dry/wet treatment (src/GPU/Arista_kernel.cu:301-321) with h2 = 0
    if ((w00 < eps) && (H1 - h11 > H0)) w11 = 0.0f;
    if ((w10 < eps) && (H0 - h10 > H1)) w01 = 0.0f;

    float h1ij = 0.5f * (w00 + w10);
    float s00 = sdes(w00, eps, eps4), s10 = sdes(w10, eps, eps4);
    float r00 = fsqrt(w00), r10 = fsqrt(w10);
    float u1ij = (r00 * (w01 * s00) + r10 * (w11 * s10)) * rcp(r00 + r10 + LH_EPSF);
    if (!(h1ij >= LH_EPSF)) return false;

    // hydrostatic reconstruction in FP64 (src/GPU/Arista_kernel.cu:94-150)
    const double Hm = fmin(H0, H1);
    const double e10 = fmax(h10 - H0 + Hm, 0.0), e11 = fmax(h11 - H1 + Hm, 0.0);
    const float deta1 = (float)(e11 - e10);
    float tp1 = g * h1ij * deta1;

    // averaged state, then both cells, internal pair zero
    float Dx, Dw, aut2, Lx, Lw, Rx, Rw;
    {
        float sh = sdes(h1ij, eps, eps4);
        float u = (h1ij * u1ij) * sh;
        float cg = fsqrt(g * h1ij);
        Dx = u - cg;
        Dw = u + cg;
        net_zero_pair(Dx, Dw, &aut2);
    }
    {
        float u = w01 * s00, cg = fsqrt(g * w00);
        Lx = u - cg;
        Lw = u + cg;
        net_zero_pair(Lx, Lw, nullptr);
        u = w11 * s10;
        cg = fsqrt(g * w10);
        Rx = u - cg;
        Rw = u + cg;
        net_zero_pair(Rx, Rw, nullptr);
    }
    float aut1 = fminf(Dx, fminf(Lx, Rx));
    float aut3 = fmaxf(Dw, fmaxf(Lw, Rw));
    float maxl = fmaxf(fmaxf(fabsf(aut3), fabsf(aut1)), fabsf(u1ij));
    float sg = aut1 + aut3;
    if (!(fabsf(sg) < LH_EPSF) && !(sg > 0.0f)) aut2 = -aut2;

    // PVM coefficients (src/GPU/Arista_kernel.cu:446-465), h2ij < eps
    float c0, c1;
    float d13 = aut1 - aut3, d12 = aut1 - aut2, d23 = aut2 - aut3;
    if ((fabsf(d13) >= LH_EPSF) && (fabsf(d12) >= LH_EPSF) && (fabsf(d23) >= LH_EPSF)) {
        float id = rcp(-d13);
        c0 = (aut3 * fabsf(aut1) - aut1 * fabsf(aut3)) * id;
        c1 = (fabsf(aut3) - fabsf(aut1)) * id;
    } else {
        c0 = maxl;
        c1 = 0.0f;
    }

    float FL1 = (w00 < LH_EPSF) ? 0.0f : w01 * w01 * s00;
    float FR1 = (w10 < LH_EPSF) ? 0.0f : w11 * w11 * s10;
    float dF0 = w11 - w01, dF1 = FR1 - FL1;
    float t20 = dF0, t21 = tp1 + dF1;
    float I0 = deta1;
    float V0 = 0.5f * (c0 * I0 + c1 * t20);
    float V1 = 0.5f * (c0 * dF0 + c1 * t21);
    float M0 = 0.5f * dF0, M1 = 0.5f * (dF1 + tp1);
    float Fm0 = M0 + (w01 - V0), Fm1 = M1 + (FL1 - V1);
    float Fp1 = M1 + (V1 - FR1);

    // tangential upwinding (src/GPU/Arista_kernel.cu:506-532)
    float ut1 = 0.0f;
    if (!(fabsf(Fm0) < LH_EPSF)) ut1 = (Fm0 > 0.0f) ? W0[2 - AX] * s00 : W1[2 - AX] * s10;
    float* P = out.Fp;
    float* Mn = out.Fm;
    Mn[0] = len * Fm0;
    P[0] = -Mn[0];
    rotate_out<AX>(Mn, P, Fm1, Fp1, Fm0 * ut1, len, len * sn);

    // positivity with the partial accumulators (src/GPU/Arista_kernel.cu:554-609)
    const float area = k.area, dt = k.dtf, bq = k.bq;
    float dt0 = 1e30f;
    {
        const bool out0 = Mn[0] > 0.0f, in0 = interior && (Mn[0] < 0.0f);
        if (out0 || in0) dt0 = (out0 ? (W0[0] + bq * a0h1) : (W1[0] + bq * a1h1)) * area * rcp(fabsf(Mn[0]) + LH_EPSF);
    }
    if (!(dt <= dt0)) {
        float al = dt0 * k.idt_eps;
        Mn[0] *= al; Mn[1] *= al; Mn[2] *= al;
        P[0] *= al; P[1] *= al; P[2] *= al;
    }
    if (maxl < eps) maxl += eps;
    out.lam = len * maxl * rcp(k.peso);
    return true;
}

// Initial wave-speed edge (src/GPU/Arista_kernel.cu:1040-1127) in FP32.
__device__ __forceinline__ float edge_dt0(const K32& k, const float* W0, const float* W1, float n1x, float n1y,
                                          float len)
{
    const float eps = k.eps, eps4 = k.eps4;
    float w00 = W0[0], w01 = W0[1] * n1x + W0[2] * n1y, w02 = W0[3], w03 = W0[4] * n1x + W0[5] * n1y;
    float w10 = W1[0], w11 = W1[1] * n1x + W1[2] * n1y, w12 = W1[3], w13 = W1[4] * n1x + W1[5] * n1y;
    float h1ij = 0.5f * (w00 + w10), h2ij = 0.5f * (w02 + w12);
    float r00 = fsqrt(w00), r10 = fsqrt(w10), r02 = fsqrt(w02), r12 = fsqrt(w12);
    float u1ij = (r00 * (w01 * sdes(w00, eps, eps4)) + r10 * (w11 * sdes(w10, eps, eps4))) * rcp(r00 + r10 + LH_EPSF);
    float u2ij = (r02 * (w03 * sdes(w02, eps, eps4)) + r12 * (w13 * sdes(w12, eps, eps4))) * rcp(r02 + r12 + LH_EPSF);
    float Dx, Dy, Dz, Dw;
    eig<0>(h1ij, u1ij, h2ij, u2ij, 0.0f, 0.0f, k.g, k.gp, k.igp, eps, eps4, Dx, Dy, Dz, Dw);
    float maxl = fmaxf(fmaxf(Dx, Dw), fmaxf(fabsf(u1ij), fabsf(u2ij)));
    if (maxl < eps) maxl += eps;
    return len * maxl;
}

// ----------------------------------------------------------- compensated state
struct Pair {
    float hi, lo;
};

// (a.hi + a.lo) + b, renormalised
__device__ __forceinline__ Pair add(Pair a, float b)
{
#pragma clang fp contract(off)
    float s = a.hi + b;
    float bb = s - a.hi;
    float e = (a.hi - (s - bb)) + (b - bb);
    float lo = a.lo + e;
    float hi = s + lo;
    return Pair{hi, lo - (hi - s)};
}

// c * a, with the product's rounding error kept (FMA TwoProd)
__device__ __forceinline__ Pair scale(Pair a, float c)
{
#pragma clang fp contract(off)
    float p = c * a.hi;
    float e = __builtin_fmaf(c, a.hi, -p);
    float lo = e + c * a.lo;
    float hi = p + lo;
    return Pair{hi, lo - (hi - p)};
}

// ca * a + cb * b
__device__ __forceinline__ Pair lin2(Pair a, float ca, Pair b, float cb)
{
#pragma clang fp contract(off)
    float p1 = ca * a.hi, e1 = __builtin_fmaf(ca, a.hi, -p1);
    float p2 = cb * b.hi, e2 = __builtin_fmaf(cb, b.hi, -p2);
    float s = p1 + p2;
    float bb = s - p1;
    float es = (p1 - (s - bb)) + (p2 - bb);
    float lo = ((e1 + e2) + es) + (ca * a.lo + cb * b.lo);
    float hi = s + lo;
    return Pair{hi, lo - (hi - s)};
}

__device__ __forceinline__ bool negative(Pair a) { return a.hi < 0.0f || (a.hi == 0.0f && a.lo < 0.0f); }

struct KCell {
    float eps, eps4, r, g, mu, mfc, mf0, mfs, vmax1, vmax2, dtf, valf;
    float ieps, ten_eps;
    int pouliquen;
    float tan1, tan2, tan3, tan4, LoverH;
};

// Cell update (src/GPU/Volumen_kernel.cu:7-180, 231-284) on compensated
// pairs: o = old state (h1,q1x,q1y,h2,q2x,q2y), acc = flux sums.  hb (if
// given) collects the health bits: a depth below -eps_h before the clip, a
// non-finite new state.
__device__ __forceinline__ void cell_update(const KCell& k, Pair* o, const float* acc, unsigned* hb = nullptr)
{
    const float eps = k.eps, eps4 = k.eps4, dt = k.dtf;
    // old-state velocities for the implicit friction (speeds taken where needed)
    const float so1 = sdes(o[0].hi, eps, eps4), so2 = sdes(o[3].hi, eps, eps4);
    const float uo1x = o[1].hi * so1, uo1y = o[2].hi * so1, uo2x = o[4].hi * so2, uo2y = o[5].hi * so2;

    Pair n[6];
#pragma unroll
    for (int f = 0; f < 6; f++) n[f] = add(o[f], k.valf * acc[f]);

    if (hb && (n[0].hi < -eps || n[3].hi < -eps)) *hb |= HEALTH_NEGATIVE_DEPTH;
    // filtroEstado: clip, damp thin layers, speed caps
    if (negative(n[0])) n[0] = Pair{0.0f, 0.0f};
    if (negative(n[3])) n[3] = Pair{0.0f, 0.0f};
    {
        // exp(-dt / ((h/eps)^4 + EPS)) is exactly 1 in FP32 once the exponent is below 2^-26
        const float lim = dt * 6.7108864e7f;
        float z = n[0].hi * k.ieps;
        float z4 = (z * z) * (z * z);
        if (z4 < lim) {
            float aux = fexp(-dt * rcp(z4 + LH_EPSF));
            n[1] = scale(n[1], aux);
            n[2] = scale(n[2], aux);
        }
        z = n[3].hi * k.ieps;
        z4 = (z * z) * (z * z);
        if (z4 < lim) {
            float aux = fexp(-dt * rcp(z4 + LH_EPSF));
            n[4] = scale(n[4], aux);
            n[5] = scale(n[5], aux);
        }
    }
    // speed caps: q <- (q s(h)) (vmax / u) h, i.e. q scaled by (vmax / u) h s(h)
    if (k.vmax1 > 0.0f) {
        float h = n[0].hi, s = sdes(h, eps, eps4);
        float ux = n[1].hi * s, uy = n[2].hi * s, u = fsqrt(ux * ux + uy * uy);
        if (u > k.vmax1) {
            float c = k.vmax1 * rcp(u) * hsdes(h, eps, eps4);
            n[1] = scale(n[1], c); n[2] = scale(n[2], c);
        }
    }
    if (k.vmax2 > 0.0f) {
        float h = n[3].hi, s = sdes(h, eps, eps4);
        float ux = n[4].hi * s, uy = n[5].hi * s, u = fsqrt(ux * ux + uy * uy);
        if (u > k.vmax2) {
            float c = k.vmax2 * rcp(u) * hsdes(h, eps, eps4);
            n[4] = scale(n[4], c); n[5] = scale(n[5], c);
        }
    }

    // disImplicita: three sequential, non-exclusive cases, each linear in q
    {
        const float h1 = n[0].hi, h2 = n[3].hi, r = k.r;
        const float s1 = sdes(h1, eps, eps4), s2 = sdes(h2, eps, eps4);
        const float hmod = h2 + r * h1;
        const float em = eps * (1.0f + r);
        float phi;   // sqrt2 hmod / sqrt(hmod^4 + max(hmod, em)^4)
        {
            float m2 = hmod * hmod;
            phi = (hmod >= em) ? rcp(hmod) : SQRT2F * hmod * rsq(m2 * m2 + (em * em) * (em * em));
        }
        if ((h1 > 0.0f) && (h2 > 0.0f)) {
            const float ddx = uo1x - uo2x, ddy = uo1y - uo2y;
            const float du = fsqrt(ddx * ddx + ddy * ddy), uo2 = fsqrt(uo2x * uo2x + uo2y * uo2y);
            float c1 = dt * h2 * phi * k.mfc * du;
            float c2 = dt * r * h1 * phi * k.mfc * du;
            float c3 = dt * k.g * k.mfs * k.mfs * uo2 * rcp(fpow(h2, 4.0f / 3.0f) + LH_EPSF);
            float det = rcp((1.0f + c1) * (1.0f + c2 + c3) - c1 * c2);
            float a11 = h1 * s1 * (1.0f + c2 + c3) * det, a12 = h1 * c1 * s2 * det;
            float a21 = h2 * c2 * s1 * det, a22 = h2 * s2 * (1.0f + c1) * det;
            Pair q1x = lin2(n[1], a11, n[4], a12), q2x = lin2(n[4], a22, n[1], a21);
            Pair q1y = lin2(n[2], a11, n[5], a12), q2y = lin2(n[5], a22, n[2], a21);
            n[1] = q1x; n[4] = q2x; n[2] = q1y; n[5] = q2y;
        }
        if ((h1 > 0.0f) && (h2 < eps)) {
            const float uo1 = fsqrt(uo1x * uo1x + uo1y * uo1y);
            float c = dt * k.g * k.mf0 * k.mf0 * uo1 * rcp(fpow(h1, 4.0f / 3.0f) + LH_EPSF);
            float f = h1 * s1 * rcp(1.0f + c);
            n[1] = scale(n[1], f);
            n[2] = scale(n[2], f);
        }
        if ((h2 > 0.0f) && (h1 < eps)) {
            const float uo2 = fsqrt(uo2x * uo2x + uo2y * uo2y);
            float c = dt * k.g * k.mfs * k.mfs * uo2 * rcp(fpow(h2, 4.0f / 3.0f) + LH_EPSF);
            float f = h2 * s2 * rcp(1.0f + c);
            n[4] = scale(n[4], f);
            n[5] = scale(n[5], f);
        }
    }

    // coulomb
    {
        float x = n[0].hi * k.ten_eps;
        float ex = 0.0f;
        if (x < 9.5f) ex = fexp(-(x * x));
        float fsc = 1.0f - k.r * (1.0f - ex);
        float muc = k.mu;
        if (k.pouliquen) {
            const float u1 = fsqrt(n[1].hi * n[1].hi + n[2].hi * n[2].hi) * sdes(n[0].hi, k.eps, k.eps4);
            const float u2 = fsqrt(n[4].hi * n[4].hi + n[5].hi * n[5].hi) * sdes(n[3].hi, k.eps, k.eps4);
            muc = friction32(k, n[0].hi, u1, n[3].hi, u2);
        }
        float sc = fsc * muc * k.g * n[3].hi * dt;
        float nq = fsqrt(n[4].hi * n[4].hi + n[5].hi * n[5].hi);
        float f = ((nq + sc >= LH_EPSF) && (n[3].hi > 0.0f) && (nq >= sc)) ? nq * rcp(nq + sc) : 0.0f;
        if (f != 1.0f) { n[4] = scale(n[4], f); n[5] = scale(n[5], f); }
    }
    if (hb && !(__builtin_fabsf(((n[0].hi + n[1].hi) + (n[2].hi + n[3].hi)) + (n[4].hi + n[5].hi)) <= 3.40282347e38f))
        *hb |= HEALTH_NONFINITE;
#pragma unroll
    for (int f = 0; f < 6; f++) o[f] = n[f];
}

// cell_update() for a cell without sediment: o[3..5] (h2, q2x, q2y) are zero
// pairs and acc[3..5] are zero.  The layer-2 sums stay zero, their damping
// and the Coulomb scaling multiply zeros, and disImplicita can only take its
// water-bottom case, so those steps are dropped; the layer-1 steps are
// cell_update()'s, in its order.  o[3..5] are left as they are (zero).
__device__ __forceinline__ void cell_update1(const KCell& k, Pair* o, const float* acc, unsigned* hb = nullptr)
{
    const float eps = k.eps, eps4 = k.eps4, dt = k.dtf;
    const float so1 = sdes(o[0].hi, eps, eps4);
    const float uo1x = o[1].hi * so1, uo1y = o[2].hi * so1;

    Pair n[3];
#pragma unroll
    for (int f = 0; f < 3; f++) n[f] = add(o[f], k.valf * acc[f]);

    if (hb && n[0].hi < -eps) *hb |= HEALTH_NEGATIVE_DEPTH;
    if (negative(n[0])) n[0] = Pair{0.0f, 0.0f};
    {
        const float lim = dt * 6.7108864e7f;
        float z = n[0].hi * k.ieps;
        float z4 = (z * z) * (z * z);
        if (z4 < lim) {
            float aux = fexp(-dt * rcp(z4 + LH_EPSF));
            n[1] = scale(n[1], aux);
            n[2] = scale(n[2], aux);
        }
    }
    if (k.vmax1 > 0.0f) {
        float h = n[0].hi, s = sdes(h, eps, eps4);
        float ux = n[1].hi * s, uy = n[2].hi * s, u = fsqrt(ux * ux + uy * uy);
        if (u > k.vmax1) {
            float c = k.vmax1 * rcp(u) * hsdes(h, eps, eps4);
            n[1] = scale(n[1], c); n[2] = scale(n[2], c);
        }
    }
    {
        const float h1 = n[0].hi;
        const float s1 = sdes(h1, eps, eps4);
        if (h1 > 0.0f) {
            const float uo1 = fsqrt(uo1x * uo1x + uo1y * uo1y);
            float c = dt * k.g * k.mf0 * k.mf0 * uo1 * rcp(fpow(h1, 4.0f / 3.0f) + LH_EPSF);
            float f = h1 * s1 * rcp(1.0f + c);
            n[1] = scale(n[1], f);
            n[2] = scale(n[2], f);
        }
    }
    if (hb && !(__builtin_fabsf((n[0].hi + n[1].hi) + n[2].hi) <= 3.40282347e38f)) *hb |= HEALTH_NONFINITE;
#pragma unroll
    for (int f = 0; f < 3; f++) o[f] = n[f];
}

}  // namespace f32
}  // namespace lhysea
