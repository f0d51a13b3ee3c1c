// CPU check of the FP32 no-sediment forms:
//  * edge1() must give the same layer-1 fluxes and wave speed as edge() (equal
//    values; zero results may differ in sign) and edge() zero layer-2 fluxes,
//    on random edges whose two cells carry water only: wet, thin, dry and
//    near-dry depths, steps in the bottom, interior and ghost (wall / open)
//    edges in x and y;
//  * an x-edge and the y-edge of the same states rotated by 90 degrees give
//    rotated fluxes of equal values (edge() with and without sediment,
//    edge1(), both normal signs);
//  * cell_update1() must give cell_update()'s layer-1 state and cell_update()
//    a zero layer-2 state, on random water-only cells with zero layer-2 sums
//    (both FP32 checks under the Coulomb and the Pouliquen law);
//  * the FP64 edge1() must give the FP64 edge()'s layer-1 fluxes and wave
//    speed, and the FP64 cell_update1() cell_update()'s state, for the
//    Coulomb and the Pouliquen law; the FP64 edges pass the rotation check.
// Build: g++ -std=c++17 -O2 -ffp-contract=off -I tests/native edge1_check.cpp
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <random>
#include "../../landslidehysea_amd/csrc/lhysea_fp32.hpp"
#include "../../landslidehysea_amd/csrc/lhysea_fp64.hpp"
using namespace lhysea::f32;

static bool same(float a, float b) { return a == b; }

int main(int argc, char** argv)
{
    const long iters = argc > 1 ? std::atol(argv[1]) : 500000;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0, 1);
    long wet = 0, dry = 0, bad = 0;
    for (long it = 0; it < iters; it++) {
        K32 k{};
        k.eps = 5e-3f;
        k.eps4 = k.eps * k.eps * k.eps * k.eps;
        k.g = 9.81f;
        k.r = (float)(0.45 + 0.15 * U(rng));
        k.gp = (float)(9.81 * (1.0 - k.r));
        k.igp = (float)(1.0 / (9.81 * (1.0 - k.r)));
        k.mu = 0.17f;
        k.peso = 1.0f;
        k.dtf = (float)(0.05 + 0.3 * U(rng));
        k.area = 3091.0f;
        k.len_x = k.len_y = 55.6f;
        k.ten_eps = 10.0f / k.eps;
        k.idt_eps = (float)(1.0 / ((double)k.dtf + (double)LH_EPSF));
        k.bq = (float)(k.dtf / (double)k.area);
        k.pouliquen = U(rng) < 0.3;
        k.tan1 = 0.15f; k.tan2 = 0.25f; k.tan3 = 0.17f; k.tan4 = 0.3f; k.LoverH = 1.0f;
        const double H0 = 400 * U(rng);
        double H1 = std::max(0.0, H0 + (U(rng) < 0.3 ? 0.0 : 20 * (U(rng) - 0.5)));
        auto depth = [&](double H) {
            const double r = U(rng);
            return r < 0.15 ? 0.0 : r < 0.3 ? 0.01 * U(rng) : r < 0.4 ? 1e-4 * U(rng) : H * U(rng);
        };
        double h10 = depth(H0), h11 = depth(H1);
        float W0[6] = {(float)h10, 0, 0, 0, 0, 0}, W1[6] = {(float)h11, 0, 0, 0, 0, 0};
        for (int f = 1; f < 3; f++) {
            W0[f] = (float)((U(rng) - 0.5) * 20 * (h10 > 0));
            W1[f] = (float)((U(rng) - 0.5) * 20 * (h11 > 0));
        }
        const bool interior = U(rng) < 0.85;
        float n1x = 1, n1y = 0, len = k.len_x;
        if (U(rng) < 0.5) {
            n1x = 0;
            n1y = 1;
            len = k.len_y;
        }
        if (!interior) {
            const float bd = U(rng) < 0.5 ? -1.0f : 1.0f;
            for (int f = 0; f < 6; f++) W1[f] = W0[f];
            if (n1x != 0) W1[1] *= bd;
            else W1[2] *= bd;
            h11 = h10;
            H1 = H0;
            if (U(rng) < 0.5) {
                n1x = -n1x;
                n1y = -n1y;
            }
        }
        const float a0 = (float)((U(rng) - 0.5) * 100), a1 = (float)((U(rng) - 0.5) * 100);
        EdgeFlux A, B;
        const float sn = n1x + n1y;
        bool wa, wb;
        if (n1x != 0) {
            wa = edge<0>(k, W0, W1, h10, 0.0, h11, 0.0, H0, H1, sn, len, a0, 0.0f, a1, 0.0f, interior, A);
            wb = edge1<0>(k, W0, W1, h10, h11, H0, H1, sn, len, a0, a1, interior, B);
        } else {
            wa = edge<1>(k, W0, W1, h10, 0.0, h11, 0.0, H0, H1, sn, len, a0, 0.0f, a1, 0.0f, interior, A);
            wb = edge1<1>(k, W0, W1, h10, h11, H0, H1, sn, len, a0, a1, interior, B);
        }
        bool ok = wa == wb;
        if (ok && wa) {
            ok = same(A.lam, B.lam);
            for (int f = 0; f < 3; f++) ok = ok && same(A.Fm[f], B.Fm[f]) && same(A.Fp[f], B.Fp[f]);
            for (int f = 3; f < 6; f++) ok = ok && A.Fm[f] == 0.0f && A.Fp[f] == 0.0f;
        }
        wet += wa;
        dry += !wa;
        if (!ok && ++bad <= 5)
            std::printf("mismatch at %ld: Fm %g %g %g / %g %g %g, lam %g / %g\n", it, A.Fm[0], A.Fm[1], A.Fm[2],
                        B.Fm[0], B.Fm[1], B.Fm[2], A.lam, B.lam);
    }
    std::printf("edges %ld wet %ld dry %ld mismatches %ld\n", iters, wet, dry, bad);

    // Rotation by 90 degrees, q -> (-qy, qx): the x-edge (normal (sn, 0)) on
    // two states and the y-edge (normal (0, sn)) on the rotated states give
    // the rotated fluxes, (Fx, Fy) -> (-Fy, Fx), with equal values; both edge
    // forms, both layers (edge()), interior and ghost edges, both signs.
    long rbad = 0, rwet = 0;
    for (long it = 0; it < iters; it++) {
        K32 k{};
        k.eps = 5e-3f;
        k.eps4 = k.eps * k.eps * k.eps * k.eps;
        k.g = 9.81f;
        k.r = (float)(0.45 + 0.15 * U(rng));
        k.gp = (float)(9.81 * (1.0 - k.r));
        k.igp = (float)(1.0 / (9.81 * (1.0 - k.r)));
        k.mu = 0.17f;
        k.peso = 1.0f;
        k.dtf = (float)(0.05 + 0.3 * U(rng));
        k.area = 3091.0f;
        k.len_x = k.len_y = 55.6f;
        k.ten_eps = 10.0f / k.eps;
        k.idt_eps = (float)(1.0 / ((double)k.dtf + (double)LH_EPSF));
        k.bq = (float)(k.dtf / (double)k.area);
        k.pouliquen = U(rng) < 0.3;
        k.tan1 = 0.15f; k.tan2 = 0.25f; k.tan3 = 0.17f; k.tan4 = 0.3f; k.LoverH = 1.0f;
        const double H0 = 400 * U(rng);
        double H1 = std::max(0.0, H0 + (U(rng) < 0.3 ? 0.0 : 20 * (U(rng) - 0.5)));
        auto depth = [&](double H) {
            const double r = U(rng);
            return r < 0.15 ? 0.0 : r < 0.3 ? 0.01 * U(rng) : r < 0.4 ? 1e-4 * U(rng) : H * U(rng);
        };
        const bool sed = U(rng) < 0.5;
        double h10 = depth(H0), h11 = depth(H1), h20 = sed ? depth(H0) : 0.0, h21 = sed ? depth(H1) : 0.0;
        float W0[6] = {(float)h10, 0, 0, (float)h20, 0, 0}, W1[6] = {(float)h11, 0, 0, (float)h21, 0, 0};
        for (int f : {1, 2}) {
            W0[f] = (float)((U(rng) - 0.5) * 20 * (h10 > 0));
            W1[f] = (float)((U(rng) - 0.5) * 20 * (h11 > 0));
            W0[f + 3] = (float)((U(rng) - 0.5) * 4 * (h20 > 0));
            W1[f + 3] = (float)((U(rng) - 0.5) * 4 * (h21 > 0));
        }
        const bool interior = U(rng) < 0.7;
        const float sn = U(rng) < 0.5 ? 1.0f : -1.0f;
        if (!interior) {
            const float bd = U(rng) < 0.5 ? -1.0f : 1.0f;
            for (int f = 0; f < 6; f++) W1[f] = W0[f];
            W1[1] *= bd;
            W1[4] *= bd;
            h11 = h10;
            h21 = h20;
            H1 = H0;
        }
        float R0[6], R1[6];   // rotated states
        for (int l = 0; l < 2; l++) {
            R0[3 * l] = W0[3 * l]; R0[3 * l + 1] = -W0[3 * l + 2]; R0[3 * l + 2] = W0[3 * l + 1];
            R1[3 * l] = W1[3 * l]; R1[3 * l + 1] = -W1[3 * l + 2]; R1[3 * l + 2] = W1[3 * l + 1];
        }
        const float a0 = (float)((U(rng) - 0.5) * 100), a1 = (float)((U(rng) - 0.5) * 100);
        const float a2 = (float)((U(rng) - 0.5) * 100), a3 = (float)((U(rng) - 0.5) * 100);
        auto rotated = [](const EdgeFlux& X, const EdgeFlux& Y, int nf) {
            bool ok = X.lam == Y.lam;
            for (int l = 0; l < nf / 3; l++) {
                const int b = 3 * l;
                ok = ok && X.Fm[b] == Y.Fm[b] && X.Fp[b] == Y.Fp[b];
                ok = ok && Y.Fm[b + 1] == -X.Fm[b + 2] && Y.Fm[b + 2] == X.Fm[b + 1];
                ok = ok && Y.Fp[b + 1] == -X.Fp[b + 2] && Y.Fp[b + 2] == X.Fp[b + 1];
            }
            return ok;
        };
        EdgeFlux X, Y;
        const bool wx = edge<0>(k, W0, W1, h10, h20, h11, h21, H0, H1, sn, 55.6f, a0, a2, a1, a3, interior, X);
        const bool wy = edge<1>(k, R0, R1, h10, h20, h11, h21, H0, H1, sn, 55.6f, a0, a2, a1, a3, interior, Y);
        bool ok = wx == wy && (!wx || rotated(X, Y, 6));
        if (!sed) {
            EdgeFlux X1, Y1;
            const bool vx = edge1<0>(k, W0, W1, h10, h11, H0, H1, sn, 55.6f, a0, a1, interior, X1);
            const bool vy = edge1<1>(k, R0, R1, h10, h11, H0, H1, sn, 55.6f, a0, a1, interior, Y1);
            ok = ok && vx == vy && (!vx || rotated(X1, Y1, 3));
        }
        rwet += wx;
        if (!ok && ++rbad <= 5)
            std::printf("rotation mismatch at %ld: x Fm %g %g %g / y Fm %g %g %g\n", it, X.Fm[0], X.Fm[1], X.Fm[2],
                        Y.Fm[0], Y.Fm[1], Y.Fm[2]);
    }
    std::printf("rotated edges %ld wet %ld mismatches %ld\n", iters, rwet, rbad);
    bad += rbad;

    long cbad = 0;
    for (long it = 0; it < iters; it++) {
        KCell k{};
        k.eps = 5e-3f;
        k.eps4 = k.eps * k.eps * k.eps * k.eps;
        k.r = (float)(0.45 + 0.15 * U(rng));
        k.g = 9.81f;
        k.mu = 0.17f;
        k.mfc = 0.08f;
        k.mf0 = (float)(0.03 * U(rng));
        k.mfs = 0.03f;
        k.vmax1 = U(rng) < 0.3 ? (float)(5 * U(rng)) : 0.0f;
        k.vmax2 = U(rng) < 0.3 ? (float)(5 * U(rng)) : 0.0f;
        k.dtf = (float)(0.05 + 0.3 * U(rng));
        k.valf = (float)(k.dtf / 3091.0);
        k.ieps = 1.0f / k.eps;
        k.ten_eps = 10.0f / k.eps;
        k.pouliquen = U(rng) < 0.3;
        k.tan1 = 0.15f; k.tan2 = 0.25f; k.tan3 = 0.17f; k.tan4 = 0.3f; k.LoverH = 1.0f;
        const double r = U(rng);
        const double h = r < 0.1 ? 0.0 : r < 0.3 ? 0.02 * U(rng) : 300 * U(rng);
        Pair o[6], p[6];
        o[0] = Pair{(float)h, (float)(h - (double)(float)h)};
        for (int f = 1; f < 3; f++) {
            const double q = (U(rng) - 0.5) * 40 * (h > 0);
            o[f] = Pair{(float)q, (float)(q - (double)(float)q)};
        }
        for (int f = 3; f < 6; f++) o[f] = Pair{0.0f, 0.0f};
        float acc[6] = {0, 0, 0, 0, 0, 0};
        for (int f = 0; f < 3; f++) acc[f] = (float)((U(rng) - 0.5) * (f == 0 ? 2e3 : 2e4));
        for (int f = 0; f < 6; f++) p[f] = o[f];
        cell_update(k, o, acc);
        cell_update1(k, p, acc);
        bool ok = true;
        for (int f = 0; f < 3; f++) ok = ok && same(o[f].hi, p[f].hi) && same(o[f].lo, p[f].lo);
        for (int f = 3; f < 6; f++) ok = ok && o[f].hi == 0.0f && o[f].lo == 0.0f && p[f].hi == 0.0f && p[f].lo == 0.0f;
        if (!ok && ++cbad <= 5)
            std::printf("cell mismatch at %ld: %g %g %g / %g %g %g\n", it, o[0].hi, o[1].hi, o[2].hi, p[0].hi, p[1].hi,
                        p[2].hi);
    }
    std::printf("cells %ld mismatches %ld\n", iters, cbad);

    long dbad = 0, dwet = 0;
    for (long it = 0; it < iters; it++) {
        lhysea::Consts k{};
        k.r = 0.45 + 0.15 * U(rng);
        k.g = 9.81;
        k.eps_h = 5e-3;
        k.area = 3091.0;
        k.mu = 0.17;
        k.peso = 1.0;
        k.beta = 1.0;
        k.friction_law = U(rng) < 0.3 ? 1 : 0;
        k.tan1 = 0.1; k.tan2 = 0.3; k.tan3 = 0.12; k.tan4 = 0.35; k.LoverH = 1.0;
        const double dt = 0.05 + 0.3 * U(rng);
        const double H0 = 400 * U(rng);
        double H1 = std::max(0.0, H0 + (U(rng) < 0.3 ? 0.0 : 20 * (U(rng) - 0.5)));
        auto depth = [&](double H) {
            const double r = U(rng);
            return r < 0.15 ? 0.0 : r < 0.3 ? 0.01 * U(rng) : r < 0.4 ? 1e-4 * U(rng) : H * U(rng);
        };
        double W0[6] = {depth(H0), 0, 0, 0, 0, 0}, W1[6] = {depth(H1), 0, 0, 0, 0, 0};
        for (int f = 1; f < 3; f++) {
            W0[f] = (U(rng) - 0.5) * 20 * (W0[0] > 0);
            W1[f] = (U(rng) - 0.5) * 20 * (W1[0] > 0);
        }
        const bool interior = U(rng) < 0.85;
        double n1x = 1, n1y = 0, len = 55.6;
        if (U(rng) < 0.5) {
            n1x = 0;
            n1y = 1;
        }
        if (!interior) {
            const double bd = U(rng) < 0.5 ? -1.0 : 1.0;
            for (int f = 0; f < 6; f++) W1[f] = W0[f];
            if (n1x != 0) W1[1] *= bd;
            else W1[2] *= bd;
            H1 = H0;
            if (U(rng) < 0.5) {
                n1x = -n1x;
                n1y = -n1y;
            }
        }
        const double a0 = (U(rng) - 0.5) * 100, a1 = (U(rng) - 0.5) * 100;
        lhysea::f64::EdgeFlux A, B;
        const double sn = n1x + n1y;
        bool wa, wb;
        if (n1x != 0) {
            wa = lhysea::f64::edge<0>(k, W0, W1, H0, H1, sn, len, dt, a0, 0.0, a1, 0.0, interior, A);
            wb = lhysea::f64::edge1<0>(k, W0, W1, H0, H1, sn, len, dt, a0, a1, interior, B);
        } else {
            wa = lhysea::f64::edge<1>(k, W0, W1, H0, H1, sn, len, dt, a0, 0.0, a1, 0.0, interior, A);
            wb = lhysea::f64::edge1<1>(k, W0, W1, H0, H1, sn, len, dt, a0, a1, interior, B);
        }
        bool ok = wa == wb;
        if (ok && wa) {
            ok = A.lam == B.lam;
            for (int f = 0; f < 3; f++) ok = ok && A.Fm[f] == B.Fm[f] && A.Fp[f] == B.Fp[f];
            for (int f = 3; f < 6; f++) ok = ok && A.Fm[f] == 0.0 && A.Fp[f] == 0.0;
        }
        dwet += wa;
        if (!ok && ++dbad <= 5)
            std::printf("fp64 mismatch at %ld: Fm %g %g %g / %g %g %g, lam %g / %g\n", it, A.Fm[0], A.Fm[1],
                        A.Fm[2], B.Fm[0], B.Fm[1], B.Fm[2], A.lam, B.lam);
    }
    std::printf("fp64 edges %ld wet %ld mismatches %ld\n", iters, dwet, dbad);

    // the FP64 edges under the 90-degree rotation, as the FP32 ones above
    long drbad = 0, drwet = 0;
    for (long it = 0; it < iters; it++) {
        lhysea::Consts k{};
        k.r = 0.45 + 0.15 * U(rng);
        k.g = 9.81;
        k.eps_h = 5e-3;
        k.area = 3091.0;
        k.mu = 0.17;
        k.peso = U(rng) < 0.5 ? 1.0 : 0.9;
        k.beta = 1.0;
        k.friction_law = U(rng) < 0.3 ? 1 : 0;
        k.tan1 = 0.1; k.tan2 = 0.3; k.tan3 = 0.12; k.tan4 = 0.35; k.LoverH = 1.0;
        const double dt = 0.05 + 0.3 * U(rng);
        const double H0 = 400 * U(rng);
        double H1 = std::max(0.0, H0 + (U(rng) < 0.3 ? 0.0 : 20 * (U(rng) - 0.5)));
        auto depth = [&](double H) {
            const double r = U(rng);
            return r < 0.15 ? 0.0 : r < 0.3 ? 0.01 * U(rng) : r < 0.4 ? 1e-4 * U(rng) : H * U(rng);
        };
        const bool sed = U(rng) < 0.5;
        double W0[6] = {depth(H0), 0, 0, sed ? depth(H0) : 0.0, 0, 0};
        double W1[6] = {depth(H1), 0, 0, sed ? depth(H1) : 0.0, 0, 0};
        for (int f : {1, 2}) {
            W0[f] = (U(rng) - 0.5) * 20 * (W0[0] > 0);
            W1[f] = (U(rng) - 0.5) * 20 * (W1[0] > 0);
            W0[f + 3] = (U(rng) - 0.5) * 4 * (W0[3] > 0);
            W1[f + 3] = (U(rng) - 0.5) * 4 * (W1[3] > 0);
        }
        const bool interior = U(rng) < 0.7;
        const double sn = U(rng) < 0.5 ? 1.0 : -1.0;
        if (!interior) {
            const double bd = U(rng) < 0.5 ? -1.0 : 1.0;
            for (int f = 0; f < 6; f++) W1[f] = W0[f];
            W1[1] *= bd;
            W1[4] *= bd;
            H1 = H0;
        }
        double R0[6], R1[6];
        for (int l = 0; l < 2; l++) {
            R0[3 * l] = W0[3 * l]; R0[3 * l + 1] = -W0[3 * l + 2]; R0[3 * l + 2] = W0[3 * l + 1];
            R1[3 * l] = W1[3 * l]; R1[3 * l + 1] = -W1[3 * l + 2]; R1[3 * l + 2] = W1[3 * l + 1];
        }
        const double a0 = (U(rng) - 0.5) * 100, a1 = (U(rng) - 0.5) * 100;
        const double a2 = (U(rng) - 0.5) * 100, a3 = (U(rng) - 0.5) * 100;
        auto rotated = [](const lhysea::f64::EdgeFlux& X, const lhysea::f64::EdgeFlux& Y, int nf) {
            bool ok = X.lam == Y.lam;
            for (int l = 0; l < nf / 3; l++) {
                const int b = 3 * l;
                ok = ok && X.Fm[b] == Y.Fm[b] && X.Fp[b] == Y.Fp[b];
                ok = ok && Y.Fm[b + 1] == -X.Fm[b + 2] && Y.Fm[b + 2] == X.Fm[b + 1];
                ok = ok && Y.Fp[b + 1] == -X.Fp[b + 2] && Y.Fp[b + 2] == X.Fp[b + 1];
            }
            return ok;
        };
        lhysea::f64::EdgeFlux X, Y;
        const bool wx = lhysea::f64::edge<0>(k, W0, W1, H0, H1, sn, 55.6, dt, a0, a2, a1, a3, interior, X);
        const bool wy = lhysea::f64::edge<1>(k, R0, R1, H0, H1, sn, 55.6, dt, a0, a2, a1, a3, interior, Y);
        bool ok = wx == wy && (!wx || rotated(X, Y, 6));
        if (!sed) {
            lhysea::f64::EdgeFlux X1, Y1;
            const bool vx = lhysea::f64::edge1<0>(k, W0, W1, H0, H1, sn, 55.6, dt, a0, a1, interior, X1);
            const bool vy = lhysea::f64::edge1<1>(k, R0, R1, H0, H1, sn, 55.6, dt, a0, a1, interior, Y1);
            ok = ok && vx == vy && (!vx || rotated(X1, Y1, 3));
        }
        drwet += wx;
        if (!ok && ++drbad <= 5)
            std::printf("fp64 rotation mismatch at %ld: x Fm %g %g %g / y Fm %g %g %g\n", it, X.Fm[0], X.Fm[1],
                        X.Fm[2], Y.Fm[0], Y.Fm[1], Y.Fm[2]);
    }
    std::printf("fp64 rotated edges %ld wet %ld mismatches %ld\n", iters, drwet, drbad);
    dbad += drbad;

    long dcbad = 0;
    for (long it = 0; it < iters; it++) {
        lhysea::Consts k{};
        k.r = 0.45 + 0.15 * U(rng);
        k.g = 9.81;
        k.eps_h = 5e-3;
        k.area = 3091.0;
        k.mu = 0.17;
        k.peso = 1.0;
        k.beta = 1.0;
        k.friction_law = U(rng) < 0.3 ? 1 : 0;
        k.tan1 = 0.1; k.tan2 = 0.3; k.tan3 = 0.12; k.tan4 = 0.35; k.LoverH = 1.0;
        k.mfc = 0.08; k.mf0 = 0.03 * U(rng); k.mfs = 0.03;
        k.vmax1 = U(rng) < 0.3 ? 5 * U(rng) : 0.0;
        k.vmax2 = U(rng) < 0.3 ? 5 * U(rng) : 0.0;
        const double dt = 0.05 + 0.3 * U(rng);
        const double q = U(rng);
        const double h = q < 0.1 ? 0.0 : q < 0.3 ? 0.02 * U(rng) : 300 * U(rng);
        const double o[6] = {h, (U(rng) - 0.5) * 40 * (h > 0), (U(rng) - 0.5) * 40 * (h > 0), 0.0, 0.0, 0.0};
        const double acc[6] = {(U(rng) - 0.5) * 2e3, (U(rng) - 0.5) * 2e4, (U(rng) - 0.5) * 2e4, 0.0, 0.0, 0.0};
        double na[6], nb[6];
        lhysea::f64::cell_update(k, o, acc, dt, na);
        lhysea::f64::cell_update1(k, o, acc, dt, nb);
        bool ok = true;
        for (int f = 0; f < 3; f++) ok = ok && na[f] == nb[f];
        for (int f = 3; f < 6; f++) ok = ok && na[f] == 0.0 && nb[f] == 0.0;
        if (!ok && ++dcbad <= 5)
            std::printf("fp64 cell mismatch at %ld: %g %g %g / %g %g %g\n", it, na[0], na[1], na[2], nb[0], nb[1], nb[2]);
    }
    std::printf("fp64 cells %ld mismatches %ld\n", iters, dcbad);
    return (bad + cbad + dbad + dcbad) != 0;
}
