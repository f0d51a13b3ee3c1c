// CPU check of the generated no-sediment adjoint functions: adj_edge_ns()
// and adj_cell_ns() must return the same values and the same adjoints of
// every input (the zero-valued layer-2 inputs included) as adj_edge() and
// adj_cell(), on random sediment-free edges and cells with random output
// adjoints, in FP32 and FP64.
// Build: g++ -std=c++17 -O2 -ffp-contract=off -I tests/native adjoint_ns_check.cpp
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <random>
#include "../../landslidehysea_amd/csrc/lhysea_adjoint.hpp"
using namespace lhysea::adj;

template <typename F>
static bool same(F a, F b) { return a == b; }

template <typename F>
static long check_edges(long iters, std::mt19937_64& rng, long& wet)
{
    std::uniform_real_distribution<double> U(0, 1);
    long bad = 0;
    for (long it = 0; it < iters; it++) {
        const F eps = F(5e-3), eps4 = eps * eps * eps * eps, g = F(9.81);
        const F r = F(0.45 + 0.15 * U(rng)), gp = F(9.81 * (1.0 - (double)r)), mu = F(0.17 * U(rng));
        const F dt = F(0.05 + 0.3 * U(rng)), area = F(3091.0), len = F(55.6);
        const double H0 = 400 * U(rng);
        double H1 = std::max(0.0, H0 + (U(rng) < 0.3 ? 0.0 : 20 * (U(rng) - 0.5)));
        auto depth = [&](double H) {
            const double q = U(rng);
            return q < 0.15 ? 0.0 : q < 0.3 ? 0.01 * U(rng) : q < 0.4 ? 1e-4 * U(rng) : H * U(rng);
        };
        double h10 = depth(H0), h11 = depth(H1);
        F q0[4] = {0, 0, 0, 0}, q1[4] = {0, 0, 0, 0};
        for (int k = 0; k < 2; k++) {
            q0[k] = F((U(rng) - 0.5) * 20 * (h10 > 0));
            q1[k] = F((U(rng) - 0.5) * 20 * (h11 > 0));
        }
        F n1x = 1, n1y = 0;
        if (U(rng) < 0.5) {
            n1x = 0;
            n1y = 1;
        }
        const F intr = U(rng) < 0.85 ? F(1) : F(0);
        if (intr == F(0)) {
            const F bd = U(rng) < 0.5 ? F(-1) : F(1);
            h11 = h10;
            H1 = H0;
            q1[0] = q0[0] * (n1x != 0 ? bd : F(1));
            q1[1] = q0[1] * (n1y != 0 ? bd : F(1));
        }
        const F a0h1 = F((U(rng) - 0.5) * 100), a1h1 = F((U(rng) - 0.5) * 100);
        F go[12];
        for (int k = 0; k < 12; k++) go[k] = F(U(rng) - 0.5);
        F o[2][12];
        double gh[2][4];
        F gq[2][8], ga[2][4];
        for (int v = 0; v < 2; v++) {
            auto fn = v == 0 ? adj_edge<F> : adj_edge_ns<F>;
            fn(eps, eps4, g, gp, r, mu, dt, area, len, n1x, n1y, intr, H0, H1, h10, 0.0, h11, 0.0, q0[0], q0[1], q0[2],
               q0[3], q1[0], q1[1], q1[2], q1[3], a0h1, F(0), a1h1, F(0), o[v][0], o[v][1], o[v][2], o[v][3], o[v][4],
               o[v][5], o[v][6], o[v][7], o[v][8], o[v][9], o[v][10], o[v][11], go[0], go[1], go[2], go[3], go[4],
               go[5], go[6], go[7], go[8], go[9], go[10], go[11], gh[v][0], gh[v][1], gh[v][2], gh[v][3], gq[v][0],
               gq[v][1], gq[v][2], gq[v][3], gq[v][4], gq[v][5], gq[v][6], gq[v][7], ga[v][0], ga[v][1], ga[v][2],
               ga[v][3]);
        }
        bool ok = true;
        for (int k = 0; k < 12; k++) ok = ok && same(o[0][k], o[1][k]);
        for (int k = 0; k < 4; k++) ok = ok && same(gh[0][k], gh[1][k]) && same(ga[0][k], ga[1][k]);
        for (int k = 0; k < 8; k++) ok = ok && same(gq[0][k], gq[1][k]);
        wet += (o[0][0] != F(0) || o[0][6] != F(0));
        if (!ok && ++bad <= 5) {
            std::printf("edge mismatch at %ld (F size %zu):\n", it, sizeof(F));
            for (int k = 0; k < 4; k++) std::printf("  gh%d %.17g %.17g\n", k, gh[0][k], gh[1][k]);
            for (int k = 0; k < 8; k++) std::printf("  gq%d %.9g %.9g\n", k, (double)gq[0][k], (double)gq[1][k]);
            for (int k = 0; k < 4; k++) std::printf("  ga%d %.9g %.9g\n", k, (double)ga[0][k], (double)ga[1][k]);
        }
    }
    return bad;
}

template <typename F>
static long check_cells(long iters, std::mt19937_64& rng)
{
    std::uniform_real_distribution<double> U(0, 1);
    long bad = 0;
    for (long it = 0; it < iters; it++) {
        const F eps = F(5e-3), eps4 = eps * eps * eps * eps, r = F(0.45 + 0.15 * U(rng)), g = F(9.81);
        const F mfc = F(0.08), mf0 = F(0.03 * U(rng)), mfs = F(0.03), dt = F(0.05 + 0.3 * U(rng));
        const F val = F((double)dt / 3091.0);
        const F vmax1 = U(rng) < 0.3 ? F(5 * U(rng)) : F(0), vmax2 = U(rng) < 0.3 ? F(5 * U(rng)) : F(0);
        const double q = U(rng);
        const F h = F(q < 0.1 ? 0.0 : q < 0.3 ? 0.02 * U(rng) : 300 * U(rng));
        F in[6] = {h, F((U(rng) - 0.5) * 40 * (h > 0)), F((U(rng) - 0.5) * 40 * (h > 0)), 0, 0, 0};
        F a[6] = {F((U(rng) - 0.5) * 2e3), F((U(rng) - 0.5) * 2e4), F((U(rng) - 0.5) * 2e4), 0, 0, 0};
        const F mu = F(0.17);
        F gn[6];
        for (int k = 0; k < 6; k++) gn[k] = F(U(rng) - 0.5);
        F n[2][6], go[2][6], gac[2][6], gmu[2];
        for (int v = 0; v < 2; v++) {
            auto fn = v == 0 ? adj_cell<F> : adj_cell_ns<F>;
            fn(eps, eps4, r, g, mfc, mf0, mfs, dt, val, vmax1, vmax2, in[0], in[1], in[2], in[3], in[4], in[5], a[0],
               a[1], a[2], a[3], a[4], a[5], mu, n[v][0], n[v][1], n[v][2], n[v][3], n[v][4], n[v][5], gn[0], gn[1],
               gn[2], gn[3], gn[4], gn[5], go[v][0], go[v][1], go[v][2], go[v][3], go[v][4], go[v][5], gac[v][0],
               gac[v][1], gac[v][2], gac[v][3], gac[v][4], gac[v][5], gmu[v]);
        }
        bool ok = same(gmu[0], gmu[1]);
        for (int k = 0; k < 6; k++) ok = ok && same(n[0][k], n[1][k]) && same(go[0][k], go[1][k]) && same(gac[0][k], gac[1][k]);
        if (!ok && ++bad <= 5) {
            std::printf("cell mismatch at %ld (F size %zu):\n", it, sizeof(F));
            for (int k = 0; k < 6; k++)
                std::printf("  n%d %.9g %.9g  go %.9g %.9g  ga %.9g %.9g\n", k, (double)n[0][k], (double)n[1][k],
                            (double)go[0][k], (double)go[1][k], (double)gac[0][k], (double)gac[1][k]);
            std::printf("  gmu %.9g %.9g\n", (double)gmu[0], (double)gmu[1]);
        }
    }
    return bad;
}

int main(int argc, char** argv)
{
    const long iters = argc > 1 ? std::atol(argv[1]) : 200000;
    std::mt19937_64 rng(11);
    long wet32 = 0, wet64 = 0;
    const long e32 = check_edges<float>(iters, rng, wet32), e64 = check_edges<double>(iters, rng, wet64);
    const long c32 = check_cells<float>(iters, rng), c64 = check_cells<double>(iters, rng);
    std::printf("edges fp32 %ld (wet %ld) mismatches %ld\n", iters, wet32, e32);
    std::printf("edges fp64 %ld (wet %ld) mismatches %ld\n", iters, wet64, e64);
    std::printf("cells fp32 %ld mismatches %ld\n", iters, c32);
    std::printf("cells fp64 %ld mismatches %ld\n", iters, c64);
    return (e32 + e64 + c32 + c64) != 0;
}
