// FP64 device pipeline of the LandSlide-HySEA solver step.
//
// Every expression keeps the reference's evaluation order so that, compiled
// without contraction (the including file sets `#pragma clang fp contract(off)`),
// the FP64 path rounds one operation at a time like the FP64 oracle.
// powf(x,4) of the reference is (x*x)*(x*x); tan() is hoisted to the host;
// expf and powf are the deterministic dm:: functions (lhysea_dmath.hpp), whose
// restatement in the oracle gives the same bits.
#pragma once
#include "lhysea_common.hpp"
#include "lhysea_dmath.hpp"

namespace lhysea {
namespace f64 {

__device__ __forceinline__ double p4(double x) { double x2 = x * x; return x2 * x2; }

// U(h,q) = sqrt2 h q / sqrt(h^4 + max(h,eps_h)^4)   (src/GPU/Arista_kernel.cu:378-379)
__device__ __forceinline__ double desing_u(double h, double q, double eps)
{
    return LH_SQRT2 * h * q / sqrt(p4(h) + p4(fmax(h, eps)));
}

// Friction coefficient: Coulomb (src/GPU/Arista_kernel.cu:11-16) or Pouliquen (:20-58).
__device__ __forceinline__ double friction(const Consts& k, double h1ij, double u1ij, double h2ij, double u2ij)
{
    if (k.friction_law == 0) return k.mu;
    double x = 10.0 * h1ij / k.eps_h;
    double rr = 1.0 - k.r * (1.0 - dm::exp(-(x * x)));
    double gp = k.g * rr;
    double fr1 = LH_SQRT2 * (u1ij * u1ij) * h1ij / (gp * sqrt(p4(h1ij) + p4(fmax(h1ij, k.eps_h))));
    double fr2 = LH_SQRT2 * (u2ij * u2ij) * h2ij / (gp * sqrt(p4(h2ij) + p4(fmax(h2ij, k.eps_h))));
    double fr = sqrt(fr1 + fr2 + (1.0 - rr) * fr1 * fr2);
    const double beta = 0.136, L2 = 8.0e-4, chi = 1.0e-3;
    double mustop = k.tan1 + (k.tan2 - k.tan1) / (1.0 + h2ij / L2);
    double mustart = k.tan3 + (k.tan4 - k.tan3) / (1.0 + h2ij / L2);
    double muf;
    if (fr > beta) muf = k.tan1 + (k.tan2 - k.tan1) / (1.0 + beta * h2ij / (fr * L2));
    else if (fabs(fr) < LH_EPS) muf = k.tan3 + (k.tan4 - k.tan3) / (1.0 + h2ij / L2);
    else muf = mustart + dm::powpos(fr / beta, chi) * (mustop - mustart);
    return muf * k.LoverH;
}

// Eigenvalue estimate + the reference's compare-swap network (src/GPU/Arista_kernel.cu:152-237).
template <int FLAG>
__device__ __forceinline__ void eig(double f0, double f1, double f2, double f3, double r, double g, double eps,
                                    double& Dx, double& Dy, double& Dz, double& Dw)
{
    double gp = g * (1.0 - r);
    double q = (FLAG == 1) ? (f1 + f3) : (f0 * f1 + f2 * f3);
    double h = f0 + f2;
    double hd = sqrt(p4(h) + p4(fmax(h, eps)));
    double u = LH_SQRT2 * q * h / hd;
    double cg = sqrt(g * h);
    double uu, u1, u2;
    Dx = u - cg;
    Dw = u + cg;
    if (FLAG == 0) {
        uu = LH_SQRT2 * (f0 * f3 + f1 * f2) * h / hd;
        u1 = f1;
        u2 = f3;
    } else {
        u1 = desing_u(f0, f1, eps);
        u2 = desing_u(f2, f3, eps);
        uu = LH_SQRT2 * (u1 * f2 + u2 * f0) * h / hd;
    }
    double du = u1 - u2;
    double aux = 1.0 - LH_SQRT2 * (du * du) * h / (gp * hd);
    cg = sqrt(gp * f0 * f2 * LH_SQRT2 * h / hd * fabs(aux));
    Dy = uu - cg;
    Dz = uu + cg;
    double t;
    if (Dx > Dy) { t = Dy; Dy = Dx; Dx = t; }
    if (Dx > Dz) { t = Dz; Dz = Dx; Dx = t; }
    if (Dx > Dw) { t = Dw; Dw = Dx; Dx = t; }
    if (Dy > Dz) { t = Dz; Dz = Dy; Dy = t; }
    if (Dy > Dw) { t = Dw; Dw = Dy; Dy = t; }
    if (Dz > Dy) { t = Dy; Dy = Dz; Dz = t; }   // the reference's inverted (2,1) step
    if (Dz > Dw) { t = Dw; Dw = Dz; Dz = t; }
}

// Result of one edge: contributions to subtract from the two cells' accumulators
// (Fm for cell 0, Fp for cell 1) and the wave-speed term len*lambda_max.
struct EdgeFlux {
    double Fm[6], Fp[6];
    double lam;
};

// Edge normals are axis-aligned: AX 0 is an x-edge with normal (sn, 0), AX 1
// a y-edge with normal (0, sn), sn = +-1.  With the normal's zero component
// dropped the projections and rotations equal the general forms (q . n,
// q . t with t = (-ny, nx), and back) exactly, up to the sign of zero
// results; the tangential flux F[h] u.t does not depend on sn (u.t is odd in
// the tangential discharge, and the two sign flips cancel).
// Normal component of the discharge (qx, qy) = (W[f], W[f + 1]).
template <int AX>
__device__ __forceinline__ double qn(const double* W, int f, double sn)
{
    return sn * W[f + AX];
}

// The edge's fluxes in the x/y frame from the normal fluxes Fm, Fp and the
// tangential flux Ft (per unit length), lsn = len * sn:
// n-frame (F, Ft) -> (F nx - Ft ny, F ny + Ft nx), and F+ rotated with -Ft.
template <int AX>
__device__ __forceinline__ void rotate_out(double* Mn, double* P, double Fm, double Fp, double Ft, double len,
                                           double lsn)
{
    const double lt = len * Ft;
    Mn[1 + AX] = lsn * Fm;
    Mn[2 - AX] = lt;
    P[1 + AX] = lsn * Fp;
    P[2 - AX] = -lt;
}

// One edge of the two-layer path-conservative PVM solver
// (src/GPU/Arista_kernel.cu:324-645).  W0/W1 unrotated states; W1 is the
// ghost when !interior.  Normal of axis AX with sign sn.
// a0h1.. are the partial accumulators of h1/h2 read by the positivity limiter.
// Returns false for a dry edge (no contribution at all).
template <int AX>
__device__ __forceinline__ bool edge(const Consts& k, const double* W0, const double* W1, double H0, double H1,
                                     double sn, double len, double dt, double a0h1, double a0h2, double a1h1,
                                     double a1h2, bool interior, EdgeFlux& out)
{
    const double r = k.r, g = k.g, eps = k.eps_h, area = k.area;
    double w00 = W0[0], w01 = qn<AX>(W0, 1, sn), w02 = W0[3], w03 = qn<AX>(W0, 4, sn);
    double w10 = W1[0], w11 = qn<AX>(W1, 1, sn), w12 = W1[3], w13 = qn<AX>(W1, 4, sn);

    // dry/wet treatment (src/GPU/Arista_kernel.cu:301-321)
    if ((w02 < eps) && (H1 - w12 > H0)) w13 = 0.0;
    if ((w12 < eps) && (H0 - w02 > H1)) w03 = 0.0;
    if ((w00 < eps) && (H1 - w10 - w12 > H0 - w02)) w11 = 0.0;
    if ((w10 < eps) && (H0 - w00 - w02 > H1 - w12)) w01 = 0.0;
    if ((w02 + w00 < eps) && (H1 - w10 - w12 > H0)) { w11 = 0.0; w13 = 0.0; }
    if ((w12 + w10 < eps) && (H0 - w00 - w02 > H1)) { w01 = 0.0; w03 = 0.0; }

    double h1ij = 0.5 * (w00 + w10);
    double u0n = desing_u(w00, w01, eps), u1n = desing_u(w10, w11, eps);
    double s0 = sqrt(w00), s1 = sqrt(w10);
    double u1ij = (s0 * u0n + s1 * u1n) / (s0 + s1 + LH_EPS);
    double h2ij = 0.5 * (w02 + w12);
    u0n = desing_u(w02, w03, eps);
    u1n = desing_u(w12, w13, eps);
    s0 = sqrt(w02);
    s1 = sqrt(w12);
    double u2ij = (s0 * u0n + s1 * u1n) / (s0 + s1 + LH_EPS);

    if (!((h1ij >= LH_EPS) || (h2ij >= LH_EPS))) return false;

    // pressure terms (src/GPU/Arista_kernel.cu:94-150)
    double Hm = fmin(H0, H1);
    double e0 = fmax(w00 + w02 - H0 + Hm, 0.0), e1 = fmax(w10 + w12 - H1 + Hm, 0.0);
    double deta1 = e1 - e0;
    e0 = fmax(w02 - H0 + Hm, 0.0);
    e1 = fmax(w12 - H1 + Hm, 0.0);
    double deta2 = e1 - e0;
    double tp1 = g * h1ij * deta1;
    double tp3 = g * h2ij * ((1 - r) * deta2 + r * deta1);
    double muc = friction(k, h1ij, u1ij, h2ij, u2ij);
    double x = 10.0 * h1ij / eps;
    double fsc = 1.0 - r * (1.0 - dm::exp(-(x * x)));
    double sc = fsc * muc * g * h2ij;
    bool coul = fabs(h2ij * u2ij) < k.peso * sc * dt;
    double t20 = 0.0, t21 = tp1, t22 = 0.0, t23 = coul ? (g * h2ij * r * deta1) : tp3;

    // wave speeds (src/GPU/Arista_kernel.cu:427-444)
    double Dx, Dy, Dz, Dw, Lx, Ly, Lz, Lw, Rx, Ry, Rz, Rw;
    eig<0>(h1ij, u1ij, h2ij, u2ij, r, g, eps, Dx, Dy, Dz, Dw);
    eig<1>(w00, w01, w02, w03, r, g, eps, Lx, Ly, Lz, Lw);
    eig<1>(w10, w11, w12, w13, r, g, eps, Rx, Ry, Rz, Rw);
    double aut1 = fmin(Dx, fmin(Lx, Rx));
    double aut3 = fmax(Dw, fmax(Lw, Rw));
    double maxl = fmax(fabs(aut3), fabs(aut1));
    double a = fmax(fabs(u1ij), fabs(u2ij));
    if (a > maxl) maxl = a;
    double aut2 = fmax(fabs(Dy), fabs(Dz));
    double sg = aut1 + aut3;
    if (!(fabs(sg) < LH_EPS) && !(sg > 0)) aut2 = -aut2;   // SGN(aut1+aut3) < 0

    // PVM coefficients (src/GPU/Arista_kernel.cu:446-465)
    double c0, c1, c2;
    if ((fabs(aut1 - aut3) >= LH_EPS) && (fabs(aut1 - aut2) >= LH_EPS) && (fabs(aut2 - aut3) >= LH_EPS)) {
        if ((h1ij >= eps) && (h2ij >= eps)) {
            double A = fabs(aut1) / ((aut1 - aut3) * (aut1 - aut2));
            double B = fabs(aut2) / ((aut2 - aut1) * (aut2 - aut3));
            double C = fabs(aut3) / ((aut3 - aut1) * (aut3 - aut2));
            c0 = A * aut2 * aut3 + B * aut1 * aut3 + C * aut1 * aut2;
            c1 = -A * (aut2 + aut3) - B * (aut1 + aut3) - C * (aut1 + aut2);
            c2 = A + B + C;
        } else {
            c0 = (aut3 * fabs(aut1) - aut1 * fabs(aut3)) / (aut3 - aut1);
            c1 = (fabs(aut3) - fabs(aut1)) / (aut3 - aut1);
            c2 = 0.0;
        }
    } else {
        c0 = maxl;
        c1 = 0.0;
        c2 = 0.0;
    }

    // physical fluxes (src/GPU/Arista_kernel.cu:62-92)
    double FL0 = w01, FL1 = (w00 < LH_EPS) ? 0.0 : desing_u(w00, w01, eps) * w01;
    double FL2 = w03, FL3 = (w02 < LH_EPS) ? 0.0 : desing_u(w02, w03, eps) * w03;
    double FR0 = w11, FR1 = (w10 < LH_EPS) ? 0.0 : desing_u(w10, w11, eps) * w11;
    double FR2 = w13, FR3 = (w12 < LH_EPS) ? 0.0 : desing_u(w12, w13, eps) * w13;
    double dF0 = FR0 - FL0, dF1 = FR1 - FL1, dF2 = FR2 - FL2, dF3 = FR3 - FL3;
    t20 = t20 + dF0; t21 = t21 + dF1; t22 = t22 + dF2; t23 = t23 + dF3;
    double M0 = 0.5 * (dF0 + 0.0), M1 = 0.5 * (dF1 + tp1), M2 = 0.5 * (dF2 + 0.0), M3 = 0.5 * (dF3 + tp3);
    double ga = g * h1ij, gb = g * h2ij;
    double A0 = t21;
    double A1 = (ga - u1ij * u1ij) * t20 + 2 * u1ij * t21 + ga * t22;
    double A2 = t23;
    double A3 = r * gb * t20 + (gb - u2ij * u2ij) * t22 + 2 * u2ij * t23;
    // identity modification (src/GPU/Arista_kernel.cu:239-276)
    double I0 = deta1, I1 = w11 - w01, I2, I3 = w13 - w03;
    if (coul) {
        I2 = 0.0;
    } else {
        I0 -= deta2 * (((w00 > eps) && (w10 > eps)) ? 1.0 : 0.0);
        I2 = deta2;
    }
    double V0 = 0.5 * (c0 * I0 + c1 * t20 + c2 * A0);
    double V1 = 0.5 * (c0 * I1 + c1 * t21 + c2 * A1);
    double V2 = 0.5 * (c0 * I2 + c1 * t22 + c2 * A2);
    double V3 = 0.5 * (c0 * I3 + c1 * t23 + c2 * A3);
    double Fm0 = M0 + (FL0 - V0), Fm1 = M1 + (FL1 - V1), Fm2 = M2 + (FL2 - V2), Fm3 = M3 + (FL3 - V3);
    double Fp0 = M0 + (V0 - FR0), Fp1 = M1 + (V1 - FR1), Fp2 = M2 + (V2 - FR2), Fp3 = M3 + (V3 - FR3);

    // tangential upwinding (src/GPU/Arista_kernel.cu:506-532): the upwind
    // cell's tangential velocity, up to the normal's sign (see rotate_out)
    double ut1 = (fabs(Fm0) < LH_EPS) ? 0.0
                                      : ((Fm0 > 0) ? desing_u(w00, W0[2 - AX], eps) : desing_u(w10, W1[2 - AX], eps));
    double ut2 = (fabs(Fm2) < LH_EPS) ? 0.0
                                      : ((Fm2 > 0) ? desing_u(w02, W0[5 - AX], eps) : desing_u(w12, W1[5 - AX], eps));

    // back to 6 components times the edge length (src/GPU/Arista_kernel.cu:534-552)
    double* P = out.Fp;
    double* Mn = out.Fm;
    const double lsn = len * sn;
    P[0] = len * Fp0;
    Mn[0] = len * Fm0;
    P[3] = len * Fp2;
    Mn[3] = len * Fm2;
    rotate_out<AX>(Mn, P, Fm1, Fp1, Fm0 * ut1, len, lsn);
    rotate_out<AX>(Mn + 3, P + 3, Fm3, Fp3, Fm2 * ut2, len, lsn);

    // positivity with the partial accumulators (src/GPU/Arista_kernel.cu:554-609)
    const double factor = 1.0 * k.beta;
    double bq = dt / area;
    double hp00 = W0[0] + bq * a0h1, hp01 = W0[3] + bq * a0h2;
    double hp10 = 0.0, hp11 = 0.0;
    if (interior) {
        hp10 = W1[0] + bq * a1h1;
        hp11 = W1[3] + bq * a1h2;
    }
    double d1 = LH_BIG, d2 = LH_BIG;
    if (Mn[0] > 0.0) d1 = hp00 * area / (factor * Mn[0] + LH_EPS);
    if (interior && P[0] > 0.0) d2 = hp10 * area / (factor * P[0] + LH_EPS);
    double dt0 = fmin(d1, d2);
    d1 = LH_BIG;
    d2 = LH_BIG;
    if (Mn[3] > 0.0) d1 = hp01 * area / (factor * Mn[3] + LH_EPS);
    if (interior && P[3] > 0.0) d2 = hp11 * area / (factor * P[3] + LH_EPS);
    double dt1 = fmin(d1, d2);
    double al = (dt <= dt0) ? 1.0 : dt0 / (dt + LH_EPS);
    Mn[0] = al * Mn[0]; Mn[1] = al * Mn[1]; Mn[2] = al * Mn[2];
    P[0] = al * P[0]; P[1] = al * P[1]; P[2] = al * P[2];
    al = (dt <= dt1) ? 1.0 : dt1 / (dt + LH_EPS);
    Mn[3] = al * Mn[3]; Mn[4] = al * Mn[4]; Mn[5] = al * Mn[5];
    P[3] = al * P[3]; P[4] = al * P[4]; P[5] = al * P[5];

    if (maxl < eps) maxl += eps;
    out.lam = len * maxl / k.peso;
    if (k.peso != 1.0) {
        for (int i = 0; i < 6; i++) { Mn[i] = k.peso * Mn[i]; P[i] = k.peso * P[i]; }
    }
    return true;
}

// The compare-swap network of eig() when the internal pair is zero: on
// (Dx, 0, 0, Dw) with Dx <= Dw (Dx = u - c, Dw = u + c, c >= 0) it gives
// Dx' = min(Dx, 0), Dw' = max(Dw, 0) and an inner pair with
// max(|Dy'|, |Dz'|) = max(Dx, -Dw, 0) (Dx >= 0: (0, Dx, 0, Dw); Dw <= 0:
// (Dx, 0, Dw, 0); otherwise (Dx, 0, 0, Dw)); aut2 (if not null) gets it.
__device__ __forceinline__ void net_zero_pair(double& Dx, double& Dw, double* aut2)
{
    if (aut2) *aut2 = fmax(fmax(Dx, -Dw), 0.0);
    Dx = fmin(Dx, 0.0);
    Dw = fmax(Dw, 0.0);
}

// edge() for an edge whose two cells carry no sediment (h2 = 0, q2 = 0 on
// both sides).  Terms that such an edge makes zero are dropped and every
// other operation keeps edge()'s order, so without contraction the layer-1
// fluxes and the wave speed equal edge()'s (zero results may differ in
// sign); the layer-2 fluxes are zero and are not produced (Fm[3..5],
// Fp[3..5] unset).  With h2ij = 0 the Coulomb flag is false for either
// friction law, deta2 = 0, the internal eigenvalue pair is (0, 0) and the PVM
// takes its two-wave or Rusanov form.
template <int AX>
__device__ __forceinline__ bool edge1(const Consts& k, const double* W0, const double* W1, double H0, double H1,
                                      double sn, double len, double dt, double a0h1, double a1h1, bool interior,
                                      EdgeFlux& out)
{
    const double g = k.g, eps = k.eps_h, area = k.area;
    double w00 = W0[0], w01 = qn<AX>(W0, 1, sn);
    double w10 = W1[0], w11 = qn<AX>(W1, 1, sn);

    // dry/wet treatment (src/GPU/Arista_kernel.cu:301-321) with h2 = 0
    if ((w00 < eps) && (H1 - w10 > H0)) w11 = 0.0;
    if ((w10 < eps) && (H0 - w00 > H1)) w01 = 0.0;

    double h1ij = 0.5 * (w00 + w10);
    double u0n = desing_u(w00, w01, eps), u1n = desing_u(w10, w11, eps);
    double s0 = sqrt(w00), s1 = sqrt(w10);
    double u1ij = (s0 * u0n + s1 * u1n) / (s0 + s1 + LH_EPS);
    if (!(h1ij >= LH_EPS)) return false;

    double Hm = fmin(H0, H1);
    double e0 = fmax(w00 - H0 + Hm, 0.0), e1 = fmax(w10 - H1 + Hm, 0.0);
    double deta1 = e1 - e0;
    double tp1 = g * h1ij * deta1;

    // wave speeds (src/GPU/Arista_kernel.cu:427-444), internal pair (0, 0)
    double Dx, Dw, aut2, Lx, Lw, Rx, Rw;
    {
        double q = h1ij * u1ij, h = h1ij;
        double hd = sqrt(p4(h) + p4(fmax(h, eps)));
        double u = LH_SQRT2 * q * h / hd, cg = sqrt(g * h);
        Dx = u - cg;
        Dw = u + cg;
        net_zero_pair(Dx, Dw, &aut2);
        h = w00;
        hd = sqrt(p4(h) + p4(fmax(h, eps)));
        u = LH_SQRT2 * w01 * h / hd;
        cg = sqrt(g * h);
        Lx = u - cg;
        Lw = u + cg;
        net_zero_pair(Lx, Lw, nullptr);
        h = w10;
        hd = sqrt(p4(h) + p4(fmax(h, eps)));
        u = LH_SQRT2 * w11 * h / hd;
        cg = sqrt(g * h);
        Rx = u - cg;
        Rw = u + cg;
        net_zero_pair(Rx, Rw, nullptr);
    }
    double aut1 = fmin(Dx, fmin(Lx, Rx));
    double aut3 = fmax(Dw, fmax(Lw, Rw));
    double maxl = fmax(fabs(aut3), fabs(aut1));
    double a = fabs(u1ij);
    if (a > maxl) maxl = a;
    double sg = aut1 + aut3;
    if (!(fabs(sg) < LH_EPS) && !(sg > 0)) aut2 = -aut2;

    // PVM coefficients (src/GPU/Arista_kernel.cu:446-465), h2ij < eps
    double c0, c1;
    if ((fabs(aut1 - aut3) >= LH_EPS) && (fabs(aut1 - aut2) >= LH_EPS) && (fabs(aut2 - aut3) >= LH_EPS)) {
        c0 = (aut3 * fabs(aut1) - aut1 * fabs(aut3)) / (aut3 - aut1);
        c1 = (fabs(aut3) - fabs(aut1)) / (aut3 - aut1);
    } else {
        c0 = maxl;
        c1 = 0.0;
    }

    double FL0 = w01, FL1 = (w00 < LH_EPS) ? 0.0 : desing_u(w00, w01, eps) * w01;
    double FR0 = w11, FR1 = (w10 < LH_EPS) ? 0.0 : desing_u(w10, w11, eps) * w11;
    double dF0 = FR0 - FL0, dF1 = FR1 - FL1;
    double t20 = dF0, t21 = tp1 + dF1;
    double M0 = 0.5 * dF0, M1 = 0.5 * (dF1 + tp1);
    double I0 = deta1, I1 = w11 - w01;
    double V0 = 0.5 * (c0 * I0 + c1 * t20);
    double V1 = 0.5 * (c0 * I1 + c1 * t21);
    double Fm0 = M0 + (FL0 - V0), Fm1 = M1 + (FL1 - V1);
    double Fp0 = M0 + (V0 - FR0), Fp1 = M1 + (V1 - FR1);

    // tangential upwinding (src/GPU/Arista_kernel.cu:506-532)
    double ut1 = (fabs(Fm0) < LH_EPS) ? 0.0
                                      : ((Fm0 > 0) ? desing_u(w00, W0[2 - AX], eps) : desing_u(w10, W1[2 - AX], eps));
    double* P = out.Fp;
    double* Mn = out.Fm;
    P[0] = len * Fp0;
    Mn[0] = len * Fm0;
    rotate_out<AX>(Mn, P, Fm1, Fp1, Fm0 * ut1, len, len * sn);

    // positivity with the partial accumulators (src/GPU/Arista_kernel.cu:554-609)
    const double factor = 1.0 * k.beta;
    double bq = dt / area;
    double hp00 = W0[0] + bq * a0h1, hp10 = 0.0;
    if (interior) hp10 = W1[0] + bq * a1h1;
    double d1 = LH_BIG, d2 = LH_BIG;
    if (Mn[0] > 0.0) d1 = hp00 * area / (factor * Mn[0] + LH_EPS);
    if (interior && P[0] > 0.0) d2 = hp10 * area / (factor * P[0] + LH_EPS);
    double dt0 = fmin(d1, d2);
    double al = (dt <= dt0) ? 1.0 : dt0 / (dt + LH_EPS);
    Mn[0] = al * Mn[0]; Mn[1] = al * Mn[1]; Mn[2] = al * Mn[2];
    P[0] = al * P[0]; P[1] = al * P[1]; P[2] = al * P[2];

    if (maxl < eps) maxl += eps;
    out.lam = len * maxl / k.peso;
    if (k.peso != 1.0) {
        for (int i = 0; i < 3; i++) { Mn[i] = k.peso * Mn[i]; P[i] = k.peso * P[i]; }
    }
    return true;
}

// Initial wave-speed edge (src/GPU/Arista_kernel.cu:1040-1127).
__device__ __forceinline__ double edge_dt0(const Consts& k, const double* W0, const double* W1, double n1x,
                                           double n1y, double len)
{
    const double eps = k.eps_h;
    double w00 = W0[0], w01 = W0[1] * n1x + W0[2] * n1y, w02 = W0[3], w03 = W0[4] * n1x + W0[5] * n1y;
    double w10 = W1[0], w11 = W1[1] * n1x + W1[2] * n1y, w12 = W1[3], w13 = W1[4] * n1x + W1[5] * n1y;
    double h1ij = 0.5 * (w00 + w10);
    double u0n = desing_u(w00, w01, eps), u1n = desing_u(w10, w11, eps);
    double s0 = sqrt(w00), s1 = sqrt(w10);
    double u1ij = (s0 * u0n + s1 * u1n) / (s0 + s1 + LH_EPS);
    double h2ij = 0.5 * (w02 + w12);
    u0n = desing_u(w02, w03, eps);
    u1n = desing_u(w12, w13, eps);
    s0 = sqrt(w02);
    s1 = sqrt(w12);
    double u2ij = (s0 * u0n + s1 * u1n) / (s0 + s1 + LH_EPS);
    double Dx, Dy, Dz, Dw;
    eig<0>(h1ij, u1ij, h2ij, u2ij, k.r, k.g, eps, Dx, Dy, Dz, Dw);
    double maxl = fmax(Dx, Dw);
    double b = fmax(fabs(u1ij), fabs(u2ij));
    if (b > maxl) maxl = b;
    if (maxl < eps) maxl += eps;
    return len * maxl;
}

// Cell update: explicit step, filtroEstado, disImplicita, coulomb
// (src/GPU/Volumen_kernel.cu:7-180, 231-284).  o[6] old state, acc[6]
// flux sums; n[6] receives the new state.  hb (if given) collects the
// health bits: a depth below -eps_h before the clip, a non-finite new state.
__device__ __forceinline__ void cell_update(const Consts& k, const double* o, const double* acc, double dt,
                                            double* n, unsigned* hb = nullptr)
{
    const double eps = k.eps_h, r = k.r, g = k.g;
    double val = dt / k.area;
    double n10 = o[0] + val * acc[0], n11 = o[1] + val * acc[1], n12 = o[2] + val * acc[2];
    double n20 = o[3] + val * acc[3], n21 = o[4] + val * acc[4], n22 = o[5] + val * acc[5];
    if (hb && (n10 < -eps || n20 < -eps)) *hb |= HEALTH_NEGATIVE_DEPTH;

    // filtroEstado
    if (n10 < 0.0) n10 = 0.0;
    if (n20 < 0.0) n20 = 0.0;
    double aux0 = 1.0 / (p4(n10 / eps) + LH_EPS);
    double aux = dm::exp(-dt * aux0);
    n11 *= aux; n12 *= aux;
    aux0 = 1.0 / (p4(n20 / eps) + LH_EPS);
    aux = dm::exp(-dt * aux0);
    n21 *= aux; n22 *= aux;
    if (k.vmax1 > 0.0) {
        double h = n10, hm = sqrt(p4(h) + p4(fmax(h, eps)));
        double ux = LH_SQRT2 * n11 * h / hm, uy = LH_SQRT2 * n12 * h / hm;
        double u = sqrt(ux * ux + uy * uy);
        if (u > k.vmax1) { ux *= k.vmax1 / u; uy *= k.vmax1 / u; n11 = ux * h; n12 = uy * h; }
    }
    if (k.vmax2 > 0.0) {
        double h = n20, hm = sqrt(p4(h) + p4(fmax(h, eps)));
        double ux = LH_SQRT2 * n21 * h / hm, uy = LH_SQRT2 * n22 * h / hm;
        double u = sqrt(ux * ux + uy * uy);
        if (u > k.vmax2) { ux *= k.vmax2 / u; uy *= k.vmax2 / u; n21 = ux * h; n22 = uy * h; }
    }

    // disImplicita: three sequential, non-exclusive ifs
    {
        double h1 = n10, h2 = n20;
        double h1m = sqrt(p4(h1) + p4(fmax(h1, eps)));
        double h2m = sqrt(p4(h2) + p4(fmax(h2, eps)));
        double u1x = LH_SQRT2 * n11 * h1 / h1m, u1y = LH_SQRT2 * n12 * h1 / h1m;
        double u2x = LH_SQRT2 * n21 * h2 / h2m, u2y = LH_SQRT2 * n22 * h2 / h2m;
        double hmod = sqrt(p4(o[0]) + p4(fmax(o[0], eps)));
        double hmodm = sqrt(p4(o[3]) + p4(fmax(o[3], eps)));
        double uo1x = LH_SQRT2 * o[1] * o[0] / hmod, uo1y = LH_SQRT2 * o[2] * o[0] / hmod;
        double uo2x = LH_SQRT2 * o[4] * o[3] / hmodm, uo2y = LH_SQRT2 * o[5] * o[3] / hmodm;
        double ddx = uo1x - uo2x, ddy = uo1y - uo2y;
        double du = sqrt(ddx * ddx + ddy * ddy);
        double u1 = sqrt(uo1x * uo1x + uo1y * uo1y);
        double u2 = sqrt(uo2x * uo2x + uo2y * uo2y);
        hmod = h2 + r * h1;
        hmodm = sqrt(p4(hmod) + p4(fmax(hmod, eps * (1.0 + r))));
        if ((h1 > 0) && (h2 > 0)) {
            double c1 = dt * LH_SQRT2 * h2 * hmod / hmodm * k.mfc * du;
            double c2 = dt * r * LH_SQRT2 * h1 * hmod / hmodm * k.mfc * du;
            double c3 = dt * g * k.mfs * k.mfs * u2 / (dm::pow43(h2) + LH_EPS);
            double det = 1.0 / ((1.0 + c1) * (1.0 + c2 + c3) - c1 * c2);
            n11 = h1 * (u1x * (1.0 + c2 + c3) + c1 * u2x) * det;
            n21 = h2 * (u2x * (1.0 + c1) + c2 * u1x) * det;
            n12 = h1 * (u1y * (1.0 + c2 + c3) + c1 * u2y) * det;
            n22 = h2 * (u2y * (1.0 + c1) + c2 * u1y) * det;
        }
        if ((h1 > 0) && (h2 < eps)) {
            u1x = LH_SQRT2 * n11 * h1 / h1m;
            u1y = LH_SQRT2 * n12 * h1 / h1m;
            double c1 = dt * g * k.mf0 * k.mf0 * u1 / (dm::pow43(h1) + LH_EPS);
            n11 = h1 * u1x / (1.0 + c1);
            n12 = h1 * u1y / (1.0 + c1);
        }
        if ((h2 > 0) && (h1 < eps)) {
            u2x = LH_SQRT2 * n21 * h2 / h2m;
            u2y = LH_SQRT2 * n22 * h2 / h2m;
            double c1 = dt * g * k.mfs * k.mfs * u2 / (dm::pow43(h2) + LH_EPS);
            n21 = h2 * u2x / (1.0 + c1);
            n22 = h2 * u2y / (1.0 + c1);
        }
    }

    // coulomb
    {
        double x = 10.0 * n10 / eps;
        double fsc = 1.0 - r * (1.0 - dm::exp(-(x * x)));
        double muc;
        if (k.friction_law == 0) {
            muc = k.mu;
        } else {
            double u1 = LH_SQRT2 * sqrt(n11 * n11 + n12 * n12) * n10 / sqrt(p4(n10) + p4(fmax(n10, eps)));
            double u2 = LH_SQRT2 * sqrt(n21 * n21 + n22 * n22) * n20 / sqrt(p4(n20) + p4(fmax(n20, eps)));
            muc = friction(k, n10, u1, n20, u2);
        }
        double sc = fsc * muc * g * n20 * dt * 1.0;
        double normq = sqrt(n21 * n21 + n22 * n22);
        double f;
        if ((normq + sc >= LH_EPS) && (n20 > 0.0)) f = normq / (normq + sc) * ((normq >= sc) ? 1.0 : 0.0);
        else f = 0.0;
        n21 *= f;
        n22 *= f;
    }
    n[0] = n10; n[1] = n11; n[2] = n12;
    n[3] = n20; n[4] = n21; n[5] = n22;
    if (hb && !(fabs(((n10 + n11) + (n12 + n20)) + (n21 + n22)) <= 1.7976931348623157e308)) *hb |= HEALTH_NONFINITE;
}

// cell_update() for a cell without sediment: o[3..5] and acc[3..5] are zero.
// The layer-2 sums stay zero, their damping, speed cap and the Coulomb
// scaling multiply zeros, and disImplicita can only take its water-bottom
// case (for either friction law), so those steps are dropped; the layer-1
// steps are cell_update()'s in its order and give the same values.
__device__ __forceinline__ void cell_update1(const Consts& k, const double* o, const double* acc, double dt,
                                             double* n, unsigned* hb = nullptr)
{
    const double eps = k.eps_h, g = k.g;
    double val = dt / k.area;
    double n10 = o[0] + val * acc[0], n11 = o[1] + val * acc[1], n12 = o[2] + val * acc[2];
    if (hb && n10 < -eps) *hb |= HEALTH_NEGATIVE_DEPTH;

    // filtroEstado
    if (n10 < 0.0) n10 = 0.0;
    double aux0 = 1.0 / (p4(n10 / eps) + LH_EPS);
    double aux = dm::exp(-dt * aux0);
    n11 *= aux; n12 *= aux;
    if (k.vmax1 > 0.0) {
        double h = n10, hm = sqrt(p4(h) + p4(fmax(h, eps)));
        double ux = LH_SQRT2 * n11 * h / hm, uy = LH_SQRT2 * n12 * h / hm;
        double u = sqrt(ux * ux + uy * uy);
        if (u > k.vmax1) { ux *= k.vmax1 / u; uy *= k.vmax1 / u; n11 = ux * h; n12 = uy * h; }
    }

    // disImplicita, water-bottom case
    {
        double h1 = n10;
        double h1m = sqrt(p4(h1) + p4(fmax(h1, eps)));
        double hmod = sqrt(p4(o[0]) + p4(fmax(o[0], eps)));
        double uo1x = LH_SQRT2 * o[1] * o[0] / hmod, uo1y = LH_SQRT2 * o[2] * o[0] / hmod;
        double u1 = sqrt(uo1x * uo1x + uo1y * uo1y);
        if (h1 > 0) {
            double u1x = LH_SQRT2 * n11 * h1 / h1m;
            double u1y = LH_SQRT2 * n12 * h1 / h1m;
            double c1 = dt * g * k.mf0 * k.mf0 * u1 / (dm::pow43(h1) + LH_EPS);
            n11 = h1 * u1x / (1.0 + c1);
            n12 = h1 * u1y / (1.0 + c1);
        }
    }
    n[0] = n10; n[1] = n11; n[2] = n12;
    n[3] = 0.0; n[4] = 0.0; n[5] = 0.0;
    if (hb && !(fabs((n10 + n11) + n12) <= 1.7976931348623157e308)) *hb |= HEALTH_NONFINITE;
}

}  // namespace f64
}  // namespace lhysea
