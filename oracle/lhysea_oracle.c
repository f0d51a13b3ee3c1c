/*
 * lhysea_oracle.c -- FP64 C/OpenMP restatement of the LandSlide-HySEA solver step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: it is built
 * into oracle/liblhysea_oracle.so and may be loaded only by tests/, by
 * __graft_entry__.smoke() (to check the library's results) and by bench.py's
 * cpu_baseline leg.
 * The product path (landslidehysea_amd) never links, loads or calls it.
 *
 * It restates, in double precision and in the reference's own evaluation
 * order, the algorithm of the reference GPU kernels:
 *   - edge pipeline          src/GPU/Arista_kernel.cu:324-645   (lh_edge)
 *   - edge helpers           src/GPU/Arista_kernel.cu:11-16, 62-321
 *   - edge classes / ghosts  src/GPU/Arista_kernel.cu:651-949    (sweep_class)
 *   - initial-dt edge pass   src/GPU/Arista_kernel.cu:1040-1308  (lh_edge_dt0)
 *   - cell update            src/GPU/Volumen_kernel.cu:7-284     (lh_cell_update)
 *   - dt reduction           src/GPU/Reduccion_kernel.cu:130-172 (semantics: global min)
 *   - step loop              src/GPU/ShallowWater.cu:236-461     (lh_oracle_run)
 *   - eta max                src/GPU/Volumen_kernel.cu:182-213
 * Constants: EPSILON = FLT_EPSILON and SGN (src/GPU/Constantes.hxx:13-14) are
 * model constants and are kept even in FP64.  Coulomb friction law only
 * (src/GPU/Constantes.hxx:17 defines COULOMB); the Pouliquen law
 * (src/GPU/Arista_kernel.cu:20-58) is available as friction_law == 1.
 *
 * powf(x,4) and powf(x,2) of the reference are evaluated as (x*x)*(x*x) and
 * x*x; tan(delta) is a per-run constant evaluated once on the host.  expf and
 * powf are evaluated by lh_exp, lh_pow43 and lh_powpos below: FP64 functions
 * built from IEEE-exact operations only, so that a device restating the same
 * evaluation reproduces the oracle to the last bit (a platform libm's exp and
 * pow differ between machines in the last bit).
 * Build with -ffp-contract=off so that no fused multiply-add changes the
 * rounding of the restated expressions.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LH_EPS ((double)FLT_EPSILON)
#define LH_SQRT2 1.41421356237309504880168872420969808
#define LH_BIG 1e30

/* Flat parameter block mirrored by the ctypes structure in oracle/binding.py. */
typedef struct {
    int nx, ny;
    int friction_law;     /* 0 = Coulomb, 1 = Pouliquen */
    int pad_;
    double dx, dy;        /* ancho_vol, alto_vol (y-edge length dx, x-edge length dy) */
    double borde_sup, borde_inf, borde_izq, borde_der;
    double cfl, r;
    double angle1, angle2, angle3, angle4; /* radians */
    double mfc, mf0, mfs, vmax1, vmax2;
    double g, eps_h, L, H;
    double peso, beta;
} lh_params;

static inline double p4(double x) { double x2 = x * x; return x2 * x2; }

/*
 * Deterministic FP64 elementary functions (+, -, *, /, floor, frexp, ldexp only):
 *  lh_exp:  x = k ln2 + r, ln2 split into a 32-bit head and a tail (Cody and
 *           Waite, Software Manual for the Elementary Functions, 1980),
 *           |r| <= ln2/2, exp(r) by its Taylor series to r^13/13!, times 2^k;
 *           results below 2^-1020 are 0.
 *  lh_log:  x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1))
 *           by the odd series to s^23/23.
 *  lh_cbrt: x = m 2^(3q), m in [1/16, 1/2), Newton's iteration for y^3 = m
 *           from the chord of cbrt on that interval, six steps.
 * Checked on the host against the C library (a few ulp) in tests/.
 */
#define LH_LN2_HI 6.93147180369123816490e-01
#define LH_LN2_LO 1.90821492927058770002e-10

double lh_exp(double x)
{
    double k, r, p;
    if (x != x) return x;
    if (x > 709.782712893383973096) return HUGE_VAL;
    if (x < -707.0) return 0.0;
    k = floor(x * 1.44269504088896338700e+00 + 0.5);
    r = (x - k * LH_LN2_HI) - k * LH_LN2_LO;
    p = 1.6059043836821614599e-10;                 /* 1/13! */
    p = p * r + 2.0876756987868098979e-09;         /* 1/12! */
    p = p * r + 2.5052108385441718775e-08;
    p = p * r + 2.7557319223985890653e-07;
    p = p * r + 2.7557319223985890653e-06;
    p = p * r + 2.4801587301587301587e-05;
    p = p * r + 1.9841269841269841270e-04;
    p = p * r + 1.3888888888888888889e-03;
    p = p * r + 8.3333333333333333333e-03;
    p = p * r + 4.1666666666666666667e-02;
    p = p * r + 1.6666666666666666667e-01;         /* 1/3! */
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)k);
}

double lh_log(double x)
{
    int e, i;
    double m, s, s2, t, de;
    if (!(x > 0.0)) return (x == 0.0) ? -HUGE_VAL : (x - x) / (x - x);
    if (x == HUGE_VAL) return x;
    m = frexp(x, &e);
    if (m < 7.07106781186547524401e-01) { m = m * 2.0; e = e - 1; }
    s = (m - 1.0) / (m + 1.0);
    s2 = s * s;
    t = 1.0 / 23.0;
    for (i = 21; i >= 1; i -= 2) t = t * s2 + 1.0 / (double)i;
    de = (double)e;
    return de * LH_LN2_HI + (de * LH_LN2_LO + 2.0 * s * t);
}

double lh_cbrt(double x)
{
    int e, q, it, neg;
    double m, y;
    if (x == 0.0 || x != x || x == HUGE_VAL || x == -HUGE_VAL) return x;
    neg = x < 0.0;
    m = frexp(neg ? -x : x, &e);
    q = (e >= 0) ? e / 3 : -((2 - e) / 3);          /* floor(e / 3) */
    m = ldexp(m, e - 3 * q) * 0.125;                /* [1/16, 1/2) */
    q = q + 1;
    y = 0.34016 + 0.90710 * m;
    for (it = 0; it < 6; it++) y = y - (y * y * y - m) / (3.0 * y * y);
    y = ldexp(y, q);
    return neg ? -y : y;
}

static inline double lh_pow43(double h) { return h * lh_cbrt(h); }
static inline double lh_powpos(double x, double y) { return lh_exp(y * lh_log(x)); }
static inline double sgn(double x) { return (fabs(x) < LH_EPS) ? 0.0 : ((x > 0) ? 1.0 : -1.0); }

/* U(h,q) = sqrt2*h*q / sqrt(h^4 + max(h,eps_h)^4)   (Arista_kernel.cu:378-379) */
static inline double desing_u(double h, double q, double eps_h)
{
    return LH_SQRT2 * h * q / sqrt(p4(h) + p4(fmax(h, eps_h)));
}

/* Friction coefficient.  Coulomb: |tan a1| L/H (Arista_kernel.cu:11-16).
 * Pouliquen: Arista_kernel.cu:20-58. */
double lh_friction(const lh_params *p, double h1ij, double u1ij, double h2ij, double u2ij)
{
    if (p->friction_law == 0)
        return fabs(tan(p->angle1)) * (p->L / p->H);
    {
        double r = p->r, eps = p->eps_h, g = p->g;
        double x = 10.0 * h1ij / eps;
        double rr = 1.0 - r * (1.0 - lh_exp(-(x * x)));
        double gp = g * rr;
        double fr1 = LH_SQRT2 * (u1ij * u1ij) * h1ij / (gp * sqrt(p4(h1ij) + p4(fmax(h1ij, eps))));
        double fr2 = LH_SQRT2 * (u2ij * u2ij) * h2ij / (gp * sqrt(p4(h2ij) + p4(fmax(h2ij, eps))));
        double fr = sqrt(fr1 + fr2 + (1.0 - rr) * fr1 * fr2);
        double beta = 0.136, L2 = 8.0e-4, chi = 1.0e-3;
        double t1 = tan(p->angle1), t2 = tan(p->angle2), t3 = tan(p->angle3), t4 = tan(p->angle4);
        double mustop = t1 + (t2 - t1) / (1.0 + h2ij / L2);
        double mustart = t3 + (t4 - t3) / (1.0 + h2ij / L2);
        double muf;
        if (fr > beta)
            muf = t1 + (t2 - t1) / (1.0 + beta * h2ij / (fr * L2));
        else if (fabs(fr) < LH_EPS)
            muf = t3 + (t4 - t3) / (1.0 + h2ij / L2);
        else
            muf = mustart + lh_powpos(fr / beta, chi) * (mustop - mustart);
        return muf * p->L / p->H;
    }
}

/* Normal physical flux per layer (Arista_kernel.cu:62-92). */
void lh_flux1d(const double w[4], double eps_h, double F[4])
{
    F[0] = w[1];
    F[1] = (w[0] < LH_EPS) ? 0.0 : desing_u(w[0], w[1], eps_h) * w[1];
    F[2] = w[3];
    F[3] = (w[2] < LH_EPS) ? 0.0 : desing_u(w[2], w[3], eps_h) * w[3];
}

/* Eigenvalue estimate with the reference's compare-swap network
 * (Arista_kernel.cu:152-237).  flag 0: f = (h1,u1,h2,u2); flag 1: f = (h1,q1,h2,q2). */
void lh_eig(const double fin[4], double r, double g, double eps_h, int flag, double D[4])
{
    /* the reference returns a fresh vector; copy so that D may alias the input */
    const double f[4] = {fin[0], fin[1], fin[2], fin[3]};
    double gp = g * (1.0 - r);
    double q = (flag == 1) ? (f[1] + f[3]) : (f[0] * f[1] + f[2] * f[3]);
    double h = f[0] + f[2];
    double hd = sqrt(p4(h) + p4(fmax(h, eps_h)));
    double u = LH_SQRT2 * q * h / hd;
    double cg = sqrt(g * h);
    double uu, u1, u2, du, aux, t;
    D[0] = u - cg;
    D[3] = u + cg;
    if (flag == 0) {
        uu = LH_SQRT2 * (f[0] * f[3] + f[1] * f[2]) * h / hd;
        u1 = f[1];
        u2 = f[3];
    } else {
        u1 = desing_u(f[0], f[1], eps_h);
        u2 = desing_u(f[2], f[3], eps_h);
        uu = LH_SQRT2 * (u1 * f[2] + u2 * f[0]) * h / hd;
    }
    du = u1 - u2;
    aux = 1.0 - LH_SQRT2 * (du * du) * h / (gp * hd);
    cg = sqrt(gp * f[0] * f[2] * LH_SQRT2 * h / hd * fabs(aux));
    D[1] = uu - cg;
    D[2] = uu + cg;
    /* fixed sequence (0,1) (0,2) (0,3) (1,2) (1,3) then the inverted (2,1), then (2,3) */
#define CSWAP(a, b) if (D[a] > D[b]) { t = D[b]; D[b] = D[a]; D[a] = t; }
    CSWAP(0, 1) CSWAP(0, 2) CSWAP(0, 3) CSWAP(1, 2) CSWAP(1, 3)
    if (D[2] > D[1]) { t = D[1]; D[1] = D[2]; D[2] = t; }
    CSWAP(2, 3)
#undef CSWAP
}

/* PVM viscosity coefficients (Arista_kernel.cu:436-465). */
void lh_pvm_coeffs(double aut1, double aut2, double aut3, double h1ij, double h2ij,
                   double maxl, double eps_h, double a[3])
{
    if ((fabs(aut1 - aut3) >= LH_EPS) && (fabs(aut1 - aut2) >= LH_EPS) && (fabs(aut2 - aut3) >= LH_EPS)) {
        if ((h1ij >= eps_h) && (h2ij >= eps_h)) {
            double A = fabs(aut1) / ((aut1 - aut3) * (aut1 - aut2));
            double B = fabs(aut2) / ((aut2 - aut1) * (aut2 - aut3));
            double C = fabs(aut3) / ((aut3 - aut1) * (aut3 - aut2));
            a[0] = A * aut2 * aut3 + B * aut1 * aut3 + C * aut1 * aut2;
            a[1] = -A * (aut2 + aut3) - B * (aut1 + aut3) - C * (aut1 + aut2);
            a[2] = A + B + C;
        } else {
            a[0] = (aut3 * fabs(aut1) - aut1 * fabs(aut3)) / (aut3 - aut1);
            a[1] = (fabs(aut3) - fabs(aut1)) / (aut3 - aut1);
            a[2] = 0.0;
        }
    } else {
        a[0] = maxl;
        a[1] = 0.0;
        a[2] = 0.0;
    }
}

/* Dry/wet momentum treatment (Arista_kernel.cu:301-321). */
static void drywet(double w0[4], double w1[4], double H0, double H1, double eps_h)
{
    if ((w0[2] < eps_h) && (H1 - w1[2] > H0)) w1[3] = 0.0;
    if ((w1[2] < eps_h) && (H0 - w0[2] > H1)) w0[3] = 0.0;
    if ((w0[0] < eps_h) && (H1 - w1[0] - w1[2] > H0 - w0[2])) w1[1] = 0.0;
    if ((w1[0] < eps_h) && (H0 - w0[0] - w0[2] > H1 - w1[2])) w0[1] = 0.0;
    if ((w0[2] + w0[0] < eps_h) && (H1 - w1[0] - w1[2] > H0)) { w1[1] = 0.0; w1[3] = 0.0; }
    if ((w1[2] + w1[0] < eps_h) && (H0 - w0[0] - w0[2] > H1)) { w0[1] = 0.0; w0[3] = 0.0; }
}

/*
 * One edge (Arista_kernel.cu:324-645).
 * W0, W1: unrotated cell states (h1,q1x,q1y,h2,q2x,q2y); W1 is the ghost when
 * interior == 0.  (nxv,nyv) = normal * length.  acc0/acc1 point at the
 * 7-entry partial accumulators (dh1,dq1x,dq1y,dh2,dq2x,dq2y,A) of the two
 * cells; they are read (positivity) and updated.  Returns 1 if the edge is wet.
 */
int lh_edge(const lh_params *p, const double W0[6], const double W1[6], double H0, double H1,
            double nxv, double nyv, double len, double dt, double *acc0, double *acc1, int interior)
{
    const double r = p->r, g = p->g, eps = p->eps_h, area = p->dx * p->dy;
    const double peso = p->peso, factor = 1.0 * p->beta;
    double n1x = nxv / len, n1y = nyv / len;
    double w0[4], w1[4];
    double h0, h1, s0, s1, u0n, u1n, h1ij, h2ij, u1ij, u2ij;

    w0[0] = W0[0]; w0[1] = W0[1] * n1x + W0[2] * n1y; w0[2] = W0[3]; w0[3] = W0[4] * n1x + W0[5] * n1y;
    w1[0] = W1[0]; w1[1] = W1[1] * n1x + W1[2] * n1y; w1[2] = W1[3]; w1[3] = W1[4] * n1x + W1[5] * n1y;
    drywet(w0, w1, H0, H1, eps);

    h0 = w0[0]; h1 = w1[0];
    h1ij = 0.5 * (h0 + h1);
    u0n = desing_u(h0, w0[1], eps);
    u1n = desing_u(h1, w1[1], eps);
    s0 = sqrt(h0); s1 = sqrt(h1);
    u1ij = (s0 * u0n + s1 * u1n) / (s0 + s1 + LH_EPS);

    h0 = w0[2]; h1 = w1[2];
    h2ij = 0.5 * (h0 + h1);
    u0n = desing_u(h0, w0[3], eps);
    u1n = desing_u(h1, w1[3], eps);
    s0 = sqrt(h0); s1 = sqrt(h1);
    u2ij = (s0 * u0n + s1 * u1n) / (s0 + s1 + LH_EPS);

    if (!((h1ij >= LH_EPS) || (h2ij >= LH_EPS)))
        return 0;

    {
        double a0h1 = acc0[0], a0h2 = acc0[3];
        double a1h1 = interior ? acc1[0] : 0.0, a1h2 = interior ? acc1[3] : 0.0;
        double Hm, e0, e1, deta1, deta2, tp[4], tp2[4], fsc, muc, sc, x;
        int coul;
        double DES[4], EL[4], ER[4], aut1, aut2, aut3, maxl, a, b, cf[3];
        double FL[4], FR[4], dF[4], M[4], Atp[4], I2[4], Dv[4], Fm[4], Fp[4];
        double q0t, q1t, ut1, ut2, T1, T2, Fm6[6], Fp6[6];
        int k;

        /* pressure terms (Arista_kernel.cu:94-150) */
        Hm = fmin(H0, H1);
        e0 = fmax(w0[0] + w0[2] - H0 + Hm, 0.0);
        e1 = fmax(w1[0] + w1[2] - H1 + Hm, 0.0);
        deta1 = e1 - e0;
        e0 = fmax(w0[2] - H0 + Hm, 0.0);
        e1 = fmax(w1[2] - H1 + Hm, 0.0);
        deta2 = e1 - e0;
        tp[0] = 0.0;
        tp[1] = g * h1ij * deta1;
        tp[2] = 0.0;
        tp[3] = g * h2ij * ((1 - r) * deta2 + r * deta1);
        muc = lh_friction(p, h1ij, u1ij, h2ij, u2ij);
        x = 10.0 * h1ij / eps;
        fsc = 1.0 - r * (1.0 - lh_exp(-(x * x)));
        sc = fsc * muc * g * h2ij;
        coul = (fabs(h2ij * u2ij) < peso * sc * dt) ? 1 : 0;
        tp2[0] = tp[0]; tp2[1] = tp[1]; tp2[2] = tp[2];
        tp2[3] = coul ? (g * h2ij * r * deta1) : tp[3];

        /* eigenvalue estimates and wave speeds */
        DES[0] = h1ij; DES[1] = u1ij; DES[2] = h2ij; DES[3] = u2ij;
        lh_eig(DES, r, g, eps, 0, DES);
        lh_eig(w0, r, g, eps, 1, EL);
        lh_eig(w1, r, g, eps, 1, ER);
        aut1 = fmin(DES[0], fmin(EL[0], ER[0]));
        aut3 = fmax(DES[3], fmax(EL[3], ER[3]));
        maxl = fmax(fabs(aut3), fabs(aut1));
        a = fmax(fabs(u1ij), fabs(u2ij));
        if (a > maxl) maxl = a;
        aut2 = fmax(fabs(DES[1]), fabs(DES[2]));
        if (sgn(aut1 + aut3) < 0) aut2 = -aut2;
        lh_pvm_coeffs(aut1, aut2, aut3, h1ij, h2ij, maxl, eps, cf);

        /* fluxes and viscosity (Arista_kernel.cu:467-504) */
        lh_flux1d(w1, eps, FR);
        lh_flux1d(w0, eps, FL);
        for (k = 0; k < 4; k++) dF[k] = FR[k] - FL[k];
        for (k = 0; k < 4; k++) tp2[k] = tp2[k] + dF[k];
        for (k = 0; k < 4; k++) M[k] = 0.5 * (dF[k] + tp[k]);
        a = g * h1ij;
        b = g * h2ij;
        Atp[0] = tp2[1];
        Atp[1] = (a - u1ij * u1ij) * tp2[0] + 2 * u1ij * tp2[1] + a * tp2[2];
        Atp[2] = tp2[3];
        Atp[3] = r * b * tp2[0] + (b - u2ij * u2ij) * tp2[2] + 2 * u2ij * tp2[3];
        /* identity modification (Arista_kernel.cu:239-276) */
        I2[0] = deta1;
        I2[1] = w1[1] - w0[1];
        I2[3] = w1[3] - w0[3];
        if (coul) {
            I2[2] = 0.0;
        } else {
            I2[0] -= deta2 * (((w0[0] > eps) && (w1[0] > eps)) ? 1.0 : 0.0);
            I2[2] = deta2;
        }
        for (k = 0; k < 4; k++) Dv[k] = 0.5 * (cf[0] * I2[k] + cf[1] * tp2[k] + cf[2] * Atp[k]);
        for (k = 0; k < 4; k++) Fm[k] = M[k] + (FL[k] - Dv[k]);
        for (k = 0; k < 4; k++) Fp[k] = M[k] + (Dv[k] - FR[k]);

        /* tangential upwinding (Arista_kernel.cu:506-532) */
        q0t = W0[2] * n1x - W0[1] * n1y;
        q1t = W1[2] * n1x - W1[1] * n1y;
        u0n = desing_u(w0[0], q0t, eps);
        u1n = desing_u(w1[0], q1t, eps);
        ut1 = (fabs(Fm[0]) < LH_EPS) ? 0.0 : ((Fm[0] > 0) ? u0n : u1n);
        q0t = W0[5] * n1x - W0[4] * n1y;
        q1t = W1[5] * n1x - W1[4] * n1y;
        u0n = desing_u(w0[2], q0t, eps);
        u1n = desing_u(w1[2], q1t, eps);
        ut2 = (fabs(Fm[2]) < LH_EPS) ? 0.0 : ((Fm[2] > 0) ? u0n : u1n);

        /* back to 6 components, times the edge length (Arista_kernel.cu:534-552) */
        T1 = Fm[0] * ut1;
        Fp6[0] = Fp[0]; Fp6[1] = Fp[1] * n1x + T1 * n1y; Fp6[2] = Fp[1] * n1y - T1 * n1x;
        Fm6[0] = Fm[0]; Fm6[1] = Fm[1] * n1x - T1 * n1y; Fm6[2] = Fm[1] * n1y + T1 * n1x;
        T2 = Fm[2] * ut2;
        Fp6[3] = Fp[2]; Fp6[4] = Fp[3] * n1x + T2 * n1y; Fp6[5] = Fp[3] * n1y - T2 * n1x;
        Fm6[3] = Fm[2]; Fm6[4] = Fm[3] * n1x - T2 * n1y; Fm6[5] = Fm[3] * n1y + T2 * n1x;
        for (k = 0; k < 6; k++) { Fp6[k] = len * Fp6[k]; Fm6[k] = len * Fm6[k]; }

        /* positivity with the partial accumulators (Arista_kernel.cu:554-609) */
        {
            double bq = dt / area, hp00, hp01, hp10 = 0.0, hp11 = 0.0, d1, d2, dt0, dt1, alpha;
            hp00 = W0[0] + bq * a0h1;
            hp01 = W0[3] + bq * a0h2;
            if (interior) {
                hp10 = W1[0] + bq * a1h1;
                hp11 = W1[3] + bq * a1h2;
            }
            d1 = d2 = LH_BIG;
            if (Fm6[0] > 0.0) d1 = hp00 * area / (factor * Fm6[0] + LH_EPS);
            if (interior && Fp6[0] > 0.0) d2 = hp10 * area / (factor * Fp6[0] + LH_EPS);
            dt0 = fmin(d1, d2);
            d1 = d2 = LH_BIG;
            if (Fm6[3] > 0.0) d1 = hp01 * area / (factor * Fm6[3] + LH_EPS);
            if (interior && Fp6[3] > 0.0) d2 = hp11 * area / (factor * Fp6[3] + LH_EPS);
            dt1 = fmin(d1, d2);
            alpha = (dt <= dt0) ? 1.0 : dt0 / (dt + LH_EPS);
            for (k = 0; k < 3; k++) { Fm6[k] = alpha * Fm6[k]; Fp6[k] = alpha * Fp6[k]; }
            alpha = (dt <= dt1) ? 1.0 : dt1 / (dt + LH_EPS);
            for (k = 3; k < 6; k++) { Fm6[k] = alpha * Fm6[k]; Fp6[k] = alpha * Fp6[k]; }
        }

        if (maxl < eps) maxl += eps;
        {
            double c = len * maxl / peso;
            for (k = 0; k < 6; k++) acc0[k] -= peso * Fm6[k];
            acc0[6] += c;
            if (interior) {
                for (k = 0; k < 6; k++) acc1[k] -= peso * Fp6[k];
                acc1[6] += c;
            }
        }
    }
    return 1;
}

/* Initial wave-speed edge (Arista_kernel.cu:1040-1127): no dry/wet, no wet test. */
double lh_edge_dt0(const lh_params *p, const double W0[6], const double W1[6], double nxv, double nyv, double len)
{
    const double eps = p->eps_h;
    double n1x = nxv / len, n1y = nyv / len;
    double w0[4], w1[4], h0, h1, s0, s1, u0n, u1n, h1ij, h2ij, u1ij, u2ij, D[4], maxl, b;
    w0[0] = W0[0]; w0[1] = W0[1] * n1x + W0[2] * n1y; w0[2] = W0[3]; w0[3] = W0[4] * n1x + W0[5] * n1y;
    w1[0] = W1[0]; w1[1] = W1[1] * n1x + W1[2] * n1y; w1[2] = W1[3]; w1[3] = W1[4] * n1x + W1[5] * n1y;
    h0 = w0[0]; h1 = w1[0];
    h1ij = 0.5 * (h0 + h1);
    u0n = desing_u(h0, w0[1], eps); u1n = desing_u(h1, w1[1], eps);
    s0 = sqrt(h0); s1 = sqrt(h1);
    u1ij = (s0 * u0n + s1 * u1n) / (s0 + s1 + LH_EPS);
    h0 = w0[2]; h1 = w1[2];
    h2ij = 0.5 * (h0 + h1);
    u0n = desing_u(h0, w0[3], eps); u1n = desing_u(h1, w1[3], eps);
    s0 = sqrt(h0); s1 = sqrt(h1);
    u2ij = (s0 * u0n + s1 * u1n) / (s0 + s1 + LH_EPS);
    D[0] = h1ij; D[1] = u1ij; D[2] = h2ij; D[3] = u2ij;
    lh_eig(D, p->r, p->g, eps, 0, D);
    maxl = fmax(D[0], D[3]);
    b = fmax(fabs(u1ij), fabs(u2ij));
    if (b > maxl) maxl = b;
    if (maxl < eps) maxl += eps;
    return len * maxl;
}

/* filtroEstado (Volumen_kernel.cu:92-180); the cf block is dead (cf = 0). */
void lh_filtro(const lh_params *p, double n1[3], double n2[3], double dt)
{
    const double eps = p->eps_h;
    double aux0, aux;
    if (n1[0] < 0.0) n1[0] = 0.0;
    if (n2[0] < 0.0) n2[0] = 0.0;
    aux0 = 1.0 / (p4(n1[0] / eps) + LH_EPS);
    aux = lh_exp(-dt * aux0);
    n1[1] *= aux; n1[2] *= aux;
    aux0 = 1.0 / (p4(n2[0] / eps) + LH_EPS);
    aux = lh_exp(-dt * aux0);
    n2[1] *= aux; n2[2] *= aux;
    if (p->vmax1 > 0.0) {
        double h = n1[0], hm = sqrt(p4(h) + p4(fmax(h, eps)));
        double ux = LH_SQRT2 * n1[1] * h / hm, uy = LH_SQRT2 * n1[2] * h / hm;
        double u = sqrt(ux * ux + uy * uy);
        if (u > p->vmax1) { ux *= p->vmax1 / u; uy *= p->vmax1 / u; n1[1] = ux * h; n1[2] = uy * h; }
    }
    if (p->vmax2 > 0.0) {
        double h = n2[0], hm = sqrt(p4(h) + p4(fmax(h, eps)));
        double ux = LH_SQRT2 * n2[1] * h / hm, uy = LH_SQRT2 * n2[2] * h / hm;
        double u = sqrt(ux * ux + uy * uy);
        if (u > p->vmax2) { ux *= p->vmax2 / u; uy *= p->vmax2 / u; n2[1] = ux * h; n2[2] = uy * h; }
    }
}

/* disImplicita (Volumen_kernel.cu:31-90): three sequential, non-exclusive ifs. */
void lh_dis_implicita(const lh_params *p, const double o1[3], const double o2[3], double n1[3], double n2[3], double dt)
{
    const double eps = p->eps_h, r = p->r, g = p->g, mfc = p->mfc, mf0 = p->mf0, mfs = p->mfs;
    double h1, h1m, h2, h2m, u1x, u1y, u2x, u2y, hmod, hmodm, uo1x, uo1y, uo2x, uo2y, du, u1, u2, dx, dy;
    if (n1[0] < 0.0) n1[0] = 0.0;
    if (n2[0] < 0.0) n2[0] = 0.0;
    h1 = n1[0]; h1m = sqrt(p4(h1) + p4(fmax(h1, eps)));
    h2 = n2[0]; h2m = sqrt(p4(h2) + p4(fmax(h2, eps)));
    u1x = LH_SQRT2 * n1[1] * h1 / h1m; u1y = LH_SQRT2 * n1[2] * h1 / h1m;
    u2x = LH_SQRT2 * n2[1] * h2 / h2m; u2y = LH_SQRT2 * n2[2] * h2 / h2m;
    hmod = sqrt(p4(o1[0]) + p4(fmax(o1[0], eps)));
    hmodm = sqrt(p4(o2[0]) + p4(fmax(o2[0], eps)));
    uo1x = LH_SQRT2 * o1[1] * o1[0] / hmod; uo1y = LH_SQRT2 * o1[2] * o1[0] / hmod;
    uo2x = LH_SQRT2 * o2[1] * o2[0] / hmodm; uo2y = LH_SQRT2 * o2[2] * o2[0] / hmodm;
    dx = uo1x - uo2x; dy = uo1y - uo2y;
    du = sqrt(dx * dx + dy * dy);
    u1 = sqrt(uo1x * uo1x + uo1y * uo1y);
    u2 = sqrt(uo2x * uo2x + uo2y * uo2y);
    hmod = h2 + r * h1;
    hmodm = sqrt(p4(hmod) + p4(fmax(hmod, eps * (1.0 + r))));
    if ((h1 > 0) && (h2 > 0)) {
        double c1 = dt * LH_SQRT2 * h2 * hmod / hmodm * mfc * du;
        double c2 = dt * r * LH_SQRT2 * h1 * hmod / hmodm * mfc * du;
        double c3 = dt * g * mfs * mfs * u2 / (lh_pow43(h2) + LH_EPS);
        double det = 1.0 / ((1.0 + c1) * (1.0 + c2 + c3) - c1 * c2);
        n1[1] = h1 * (u1x * (1.0 + c2 + c3) + c1 * u2x) * det;
        n2[1] = h2 * (u2x * (1.0 + c1) + c2 * u1x) * det;
        n1[2] = h1 * (u1y * (1.0 + c2 + c3) + c1 * u2y) * det;
        n2[2] = h2 * (u2y * (1.0 + c1) + c2 * u1y) * det;
    }
    if ((h1 > 0) && (h2 < eps)) {
        double c1;
        u1x = LH_SQRT2 * n1[1] * h1 / h1m; u1y = LH_SQRT2 * n1[2] * h1 / h1m;
        c1 = dt * g * mf0 * mf0 * u1 / (lh_pow43(h1) + LH_EPS);
        n1[1] = h1 * u1x / (1.0 + c1);
        n1[2] = h1 * u1y / (1.0 + c1);
    }
    if ((h2 > 0) && (h1 < eps)) {
        double c1;
        u2x = LH_SQRT2 * n2[1] * h2 / h2m; u2y = LH_SQRT2 * n2[2] * h2 / h2m;
        c1 = dt * g * mfs * mfs * u2 / (lh_pow43(h2) + LH_EPS);
        n2[1] = h2 * u2x / (1.0 + c1);
        n2[2] = h2 * u2y / (1.0 + c1);
    }
}

/* coulomb (Volumen_kernel.cu:7-26) */
void lh_coulomb(const lh_params *p, double n1[3], double n2[3], double dt)
{
    const double eps = p->eps_h, r = p->r, g = p->g;
    double x = 10.0 * n1[0] / eps, fsc, muc, sc, normq, aux, u1 = 0.0, u2 = 0.0;
    if (p->friction_law != 0) {
        u1 = LH_SQRT2 * sqrt(n1[1] * n1[1] + n1[2] * n1[2]) * n1[0] / sqrt(p4(n1[0]) + p4(fmax(n1[0], eps)));
        u2 = LH_SQRT2 * sqrt(n2[1] * n2[1] + n2[2] * n2[2]) * n2[0] / sqrt(p4(n2[0]) + p4(fmax(n2[0], eps)));
    }
    fsc = 1.0 - r * (1.0 - lh_exp(-(x * x)));
    muc = lh_friction(p, n1[0], u1, n2[0], u2);
    sc = fsc * muc * g * n2[0] * dt * 1.0;
    normq = sqrt(n2[1] * n2[1] + n2[2] * n2[2]);
    if ((normq + sc >= LH_EPS) && (n2[0] > 0.0))
        aux = normq / (normq + sc) * ((normq >= sc) ? 1.0 : 0.0);
    else
        aux = 0.0;
    n2[1] *= aux;
    n2[2] *= aux;
}

/* Cell update (Volumen_kernel.cu:231-284).  Wold/Wnew: 6 values; acc: 7 entries.
 * Returns this cell's dt contribution for the next step. */
double lh_cell_update(const lh_params *p, const double Wold[6], const double acc[7], double dt, double Wnew[6])
{
    double area = p->dx * p->dy, val = dt / area;
    double dti = (acc[6] < LH_EPS) ? LH_BIG : (2.0 * p->cfl * area) / acc[6];
    double n1[3], n2[3];
    n1[0] = Wold[0] + val * acc[0]; n1[1] = Wold[1] + val * acc[1]; n1[2] = Wold[2] + val * acc[2];
    n2[0] = Wold[3] + val * acc[3]; n2[1] = Wold[4] + val * acc[4]; n2[2] = Wold[5] + val * acc[5];
    lh_filtro(p, n1, n2, dt);
    lh_dis_implicita(p, Wold, Wold + 3, n1, n2, dt);
    lh_coulomb(p, n1, n2, dt);
    Wnew[0] = n1[0]; Wnew[1] = n1[1]; Wnew[2] = n1[2];
    Wnew[3] = n2[0]; Wnew[4] = n2[1]; Wnew[5] = n2[2];
    return dti;
}

/* ------------------------------------------------------------------------- */
/* Grid-level sweeps.  State planes W[k*N + j*nx + i], k = 0..5; H[j*nx + i]. */
/* Row j = 0 is the "superior" row (borde_sup).                              */

static void load6(const double *W, size_t N, size_t c, double out[6])
{
    int k;
    for (k = 0; k < 6; k++) out[k] = W[k * N + c];
}

/* classes: 0 = even y-edges, 1 = odd y-edges, 2 = even x-edges, 3 = odd x-edges
 * (the order of ShallowWater.cu:391-427). Each class touches every cell at most once. */
static void sweep_class(const lh_params *p, const double *W, const double *H, double dt, double *acc, int cls)
{
    const int nx = p->nx, ny = p->ny;
    const size_t N = (size_t)nx * ny;
    if (cls < 2) {
        const int par = cls;
        const int nrow = (ny - par) / 2 + 1;  /* edges je = par, par+2, ..., <= ny */
        long e, ne = (long)nrow * nx;
        #pragma omp parallel for schedule(static)
        for (e = 0; e < ne; e++) {
            int je = par + 2 * (int)(e / nx), i = (int)(e % nx);
            double W0[6], W1[6], a0[7], a1[7];
            int k;
            if (je == 0) {
                size_t c0 = (size_t)i;
                load6(W, N, c0, W0);
                memcpy(W1, W0, sizeof W0); W1[2] *= p->borde_sup; W1[5] *= p->borde_sup;
                for (k = 0; k < 7; k++) a0[k] = acc[k * N + c0];
                if (lh_edge(p, W0, W1, H[c0], H[c0], 0.0, -p->dx, p->dx, dt, a0, NULL, 0))
                    for (k = 0; k < 7; k++) acc[k * N + c0] = a0[k];
            } else if (je == ny) {
                size_t c0 = (size_t)(ny - 1) * nx + i;
                load6(W, N, c0, W0);
                memcpy(W1, W0, sizeof W0); W1[2] *= p->borde_inf; W1[5] *= p->borde_inf;
                for (k = 0; k < 7; k++) a0[k] = acc[k * N + c0];
                if (lh_edge(p, W0, W1, H[c0], H[c0], 0.0, p->dx, p->dx, dt, a0, NULL, 0))
                    for (k = 0; k < 7; k++) acc[k * N + c0] = a0[k];
            } else {
                size_t c0 = (size_t)(je - 1) * nx + i, c1 = c0 + nx;
                load6(W, N, c0, W0); load6(W, N, c1, W1);
                for (k = 0; k < 7; k++) { a0[k] = acc[k * N + c0]; a1[k] = acc[k * N + c1]; }
                if (lh_edge(p, W0, W1, H[c0], H[c1], 0.0, p->dx, p->dx, dt, a0, a1, 1))
                    for (k = 0; k < 7; k++) { acc[k * N + c0] = a0[k]; acc[k * N + c1] = a1[k]; }
            }
        }
    } else {
        const int par = cls - 2;
        const int ncol = (nx - par) / 2 + 1;
        long e, ne = (long)ncol * ny;
        #pragma omp parallel for schedule(static)
        for (e = 0; e < ne; e++) {
            int j = (int)(e / ncol), ie = par + 2 * (int)(e % ncol);
            double W0[6], W1[6], a0[7], a1[7];
            int k;
            if (ie == 0) {
                size_t c0 = (size_t)j * nx;
                load6(W, N, c0, W0);
                memcpy(W1, W0, sizeof W0); W1[1] *= p->borde_izq; W1[4] *= p->borde_izq;
                for (k = 0; k < 7; k++) a0[k] = acc[k * N + c0];
                if (lh_edge(p, W0, W1, H[c0], H[c0], -p->dy, 0.0, p->dy, dt, a0, NULL, 0))
                    for (k = 0; k < 7; k++) acc[k * N + c0] = a0[k];
            } else if (ie == nx) {
                size_t c0 = (size_t)j * nx + nx - 1;
                load6(W, N, c0, W0);
                memcpy(W1, W0, sizeof W0); W1[1] *= p->borde_der; W1[4] *= p->borde_der;
                for (k = 0; k < 7; k++) a0[k] = acc[k * N + c0];
                if (lh_edge(p, W0, W1, H[c0], H[c0], p->dy, 0.0, p->dy, dt, a0, NULL, 0))
                    for (k = 0; k < 7; k++) acc[k * N + c0] = a0[k];
            } else {
                size_t c0 = (size_t)j * nx + ie - 1, c1 = c0 + 1;
                load6(W, N, c0, W0); load6(W, N, c1, W1);
                for (k = 0; k < 7; k++) { a0[k] = acc[k * N + c0]; a1[k] = acc[k * N + c1]; }
                if (lh_edge(p, W0, W1, H[c0], H[c1], p->dy, 0.0, p->dy, dt, a0, a1, 1))
                    for (k = 0; k < 7; k++) { acc[k * N + c0] = a0[k]; acc[k * N + c1] = a1[k]; }
            }
        }
    }
}

/* Initial dt (ShallowWater.cu:236-257): wave-speed sums over every edge, then
 * the global min of the per-cell dt (Volumen_kernel.cu:215-229). */
double lh_oracle_initial_dt(const lh_params *p, const double *W, const double *H)
{
    const int nx = p->nx, ny = p->ny;
    const size_t N = (size_t)nx * ny;
    double *A = (double *)calloc(N, sizeof(double));
    double area = p->dx * p->dy, dtmin = LH_BIG;
    int cls;
    long c;
    (void)H;
    if (!A) return -1.0;
    for (cls = 0; cls < 4; cls++) {
        if (cls < 2) {
            const int par = cls, nrow = (ny - par) / 2 + 1;
            long e, ne = (long)nrow * nx;
            #pragma omp parallel for schedule(static)
            for (e = 0; e < ne; e++) {
                int je = par + 2 * (int)(e / nx), i = (int)(e % nx);
                double W0[6], W1[6], v;
                if (je == 0 || je == ny) {
                    size_t c0 = (je == 0) ? (size_t)i : (size_t)(ny - 1) * nx + i;
                    double bd = (je == 0) ? p->borde_sup : p->borde_inf;
                    load6(W, N, c0, W0);
                    memcpy(W1, W0, sizeof W0); W1[2] *= bd; W1[5] *= bd;
                    v = lh_edge_dt0(p, W0, W1, 0.0, (je == 0) ? -p->dx : p->dx, p->dx);
                    A[c0] += v;
                } else {
                    size_t c0 = (size_t)(je - 1) * nx + i, c1 = c0 + nx;
                    load6(W, N, c0, W0); load6(W, N, c1, W1);
                    v = lh_edge_dt0(p, W0, W1, 0.0, p->dx, p->dx);
                    A[c0] += v; A[c1] += v;
                }
            }
        } else {
            const int par = cls - 2, ncol = (nx - par) / 2 + 1;
            long e, ne = (long)ncol * ny;
            #pragma omp parallel for schedule(static)
            for (e = 0; e < ne; e++) {
                int j = (int)(e / ncol), ie = par + 2 * (int)(e % ncol);
                double W0[6], W1[6], v;
                if (ie == 0 || ie == nx) {
                    size_t c0 = (ie == 0) ? (size_t)j * nx : (size_t)j * nx + nx - 1;
                    double bd = (ie == 0) ? p->borde_izq : p->borde_der;
                    load6(W, N, c0, W0);
                    memcpy(W1, W0, sizeof W0); W1[1] *= bd; W1[4] *= bd;
                    v = lh_edge_dt0(p, W0, W1, (ie == 0) ? -p->dy : p->dy, 0.0, p->dy);
                    A[c0] += v;
                } else {
                    size_t c0 = (size_t)j * nx + ie - 1, c1 = c0 + 1;
                    load6(W, N, c0, W0); load6(W, N, c1, W1);
                    v = lh_edge_dt0(p, W0, W1, p->dy, 0.0, p->dy);
                    A[c0] += v; A[c1] += v;
                }
            }
        }
    }
    for (c = 0; c < (long)N; c++) {
        double d = (A[c] < LH_EPS) ? LH_BIG : (2.0 * p->cfl * area) / A[c];
        if (d < dtmin) dtmin = d;
    }
    free(A);
    return dtmin;
}

/* One solver step on W (in place) with time step dt; returns dt for the next
 * step (computed from this step's wave speeds, i.e. on the pre-step state).
 * acc: 7N scratch (may be NULL); if acc_out is not NULL the final
 * accumulators are copied there (7N). */
double lh_oracle_step(const lh_params *p, double *W, const double *H, double dt, double *acc_out)
{
    const size_t N = (size_t)p->nx * p->ny;
    double *acc = (double *)calloc(7 * N, sizeof(double));
    double *Wn = (double *)malloc(6 * N * sizeof(double));
    double dtmin = LH_BIG;
    long c;
    int cls;
    if (!acc || !Wn) { free(acc); free(Wn); return -1.0; }
    for (cls = 0; cls < 4; cls++) sweep_class(p, W, H, dt, acc, cls);
    #pragma omp parallel for schedule(static) reduction(min:dtmin)
    for (c = 0; c < (long)N; c++) {
        double Wo[6], a[7], Wv[6], d;
        int k;
        for (k = 0; k < 6; k++) Wo[k] = W[k * N + c];
        for (k = 0; k < 7; k++) a[k] = acc[k * N + c];
        d = lh_cell_update(p, Wo, a, dt, Wv);
        for (k = 0; k < 6; k++) Wn[k * N + c] = Wv[k];
        if (d < dtmin) dtmin = d;
    }
    memcpy(W, Wn, 6 * N * sizeof(double));
    if (acc_out) memcpy(acc_out, acc, 7 * N * sizeof(double));
    free(acc);
    free(Wn);
    return dtmin;
}

/*
 * Fixed-step driver (ShallowWater.cu:236-461 with a step count in place of
 * the t < t_end test).  *dt_io: dt of the first step on entry (pass <= 0 to
 * compute the initial dt), dt of the next step on exit.  *t_io: time.
 * eta_max (2N: value, time; may be NULL) is updated before each step with
 * h1 - H (Volumen_kernel.cu:182-213).  gauge_series (nsteps*ng) receives
 * eta = h1 + h2 - H at the gauge cells after each step; dt_hist (nsteps)
 * receives the dt used by each step.  Both may be NULL.
 */
int lh_oracle_run(const lh_params *p, double *W, const double *H, int nsteps, double *dt_io, double *t_io,
                  double *eta_max, const int *gauges, int ng, double *gauge_series, double *dt_hist)
{
    const size_t N = (size_t)p->nx * p->ny;
    double dt = *dt_io, t = *t_io;
    int n, gI;
    long c;
    if (dt <= 0.0) dt = lh_oracle_initial_dt(p, W, H);
    if (dt < 0.0) return 2;
    for (n = 0; n < nsteps; n++) {
        double dtn;
        if (eta_max) {
            #pragma omp parallel for schedule(static)
            for (c = 0; c < (long)N; c++) {
                double v = W[c] - H[c];
                if (v > eta_max[2 * c]) { eta_max[2 * c] = v; eta_max[2 * c + 1] = t; }
            }
        }
        dtn = lh_oracle_step(p, W, H, dt, NULL);
        if (dtn < 0.0) return 2;
        if (dt_hist) dt_hist[n] = dt;
        t += dt;
        dt = dtn;
        if (gauge_series)
            for (gI = 0; gI < ng; gI++) {
                size_t gc = (size_t)gauges[gI];
                gauge_series[(size_t)n * ng + gI] = W[gc] + W[3 * N + gc] - H[gc];
            }
    }
    *dt_io = dt;
    *t_io = t;
    return 0;
}

/* Same as lh_oracle_run, but every step uses the prescribed dt_seq[n]
 * (frozen dt sequence, used for finite differences of the adjoint). */
int lh_oracle_run_fixed_dt(const lh_params *p, double *W, const double *H, int nsteps, const double *dt_seq,
                           const int *gauges, int ng, double *gauge_series)
{
    const size_t N = (size_t)p->nx * p->ny;
    int n, gI;
    for (n = 0; n < nsteps; n++) {
        if (lh_oracle_step(p, W, H, dt_seq[n], NULL) < 0.0) return 2;
        if (gauge_series)
            for (gI = 0; gI < ng; gI++) {
                size_t gc = (size_t)gauges[gI];
                gauge_series[(size_t)n * ng + gI] = W[gc] + W[3 * N + gc] - H[gc];
            }
    }
    return 0;
}

void lh_oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int lh_oracle_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

double lh_oracle_wtime(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}
