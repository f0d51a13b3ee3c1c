// CPU check of the generated discrete-adjoint functions against finite
// differences of their own forward: for random edges and cells (wet, thin,
// dry and near-dry depths, active positivity limiter through random partial
// sums, both layers, speed caps on and off) and random output adjoints g,
// every input adjoint of adj_edge<double> / adj_cell<double> must equal the
// central difference of g . forward(x) in that input, wherever the two
// one-sided differences agree (a branch point otherwise) and the derivative
// exceeds the round-off floor of the differences.  Prints one line per
// function: samples checked, the worst relative mismatch, the number beyond
// the bar.  The Pouliquen variants adj_edge_pq / adj_cell_pq are checked the
// same way, the cell's four angle tangents included, at rest, slow and fast
// states (the three regimes of the law).
// Build: g++ -std=c++17 -O1 -ffp-contract=off -I tests/native/emu adjoint_vjp_check.cpp
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include "../../landslidehysea_amd/csrc/lhysea_adjoint.hpp"
using namespace lhysea::adj;
typedef double D;

static const D kBar = 2e-3, kFloor = 1e-5;

struct Edge {
    D eps, eps4, g, gp, r, mu, dt, area, len, n1x, n1y, intr, H0, H1;
    void fwd(const D* x, D* o) const
    {
        adj_edge_fwd<D>(eps, eps4, g, gp, r, mu, dt, area, len, n1x, n1y, intr, H0, H1, x[0], x[1], x[2], x[3], x[4],
                        x[5], x[6], x[7], x[8], x[9], x[10], x[11], x[12], x[13], x[14], x[15], o[0], o[1], o[2], o[3],
                        o[4], o[5], o[6], o[7], o[8], o[9], o[10], o[11]);
    }
    void vjp(const D* x, const D* g, D* gr) const
    {
        D o[12];
        adj_edge<D>(eps, eps4, this->g, gp, r, mu, dt, area, len, n1x, n1y, intr, H0, H1, x[0], x[1], x[2], x[3], x[4],
                    x[5], x[6], x[7], x[8], x[9], x[10], x[11], x[12], x[13], x[14], x[15], o[0], o[1], o[2], o[3],
                    o[4], o[5], o[6], o[7], o[8], o[9], o[10], o[11], g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7],
                    g[8], g[9], g[10], g[11], gr[0], gr[1], gr[2], gr[3], gr[4], gr[5], gr[6], gr[7], gr[8], gr[9],
                    gr[10], gr[11], gr[12], gr[13], gr[14], gr[15]);
    }
};

struct Cell {
    D eps, eps4, r, g, mfc, mf0, mfs, dt, val, v1, v2;
    void fwd(const D* x, D* n) const
    {
        D z[6] = {0, 0, 0, 0, 0, 0}, go[6], ga[6], gm;
        adj_cell<D>(eps, eps4, r, g, mfc, mf0, mfs, dt, val, v1, v2, x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7],
                    x[8], x[9], x[10], x[11], x[12], n[0], n[1], n[2], n[3], n[4], n[5], z[0], z[1], z[2], z[3], z[4],
                    z[5], go[0], go[1], go[2], go[3], go[4], go[5], ga[0], ga[1], ga[2], ga[3], ga[4], ga[5], gm);
    }
    void vjp(const D* x, const D* gn, D* gr) const
    {
        D n[6];
        adj_cell<D>(eps, eps4, r, g, mfc, mf0, mfs, dt, val, v1, v2, x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7],
                    x[8], x[9], x[10], x[11], x[12], n[0], n[1], n[2], n[3], n[4], n[5], gn[0], gn[1], gn[2], gn[3],
                    gn[4], gn[5], gr[0], gr[1], gr[2], gr[3], gr[4], gr[5], gr[6], gr[7], gr[8], gr[9], gr[10], gr[11],
                    gr[12]);
    }
};

// the Pouliquen law: the edge takes the four tangents as parameters (they only
// decide the Coulomb stop flag), the cell as inputs (x[12..15])
struct EdgePQ {
    D eps, eps4, g, gp, r, mu, dt, area, len, n1x, n1y, intr, H0, H1, t1, t2, t3, t4, LoH;
    void fwd(const D* x, D* o) const
    {
        adj_edge_pq_fwd<D>(eps, eps4, g, gp, r, mu, dt, area, len, n1x, n1y, intr, H0, H1, t1, t2, t3, t4, LoH, x[0],
                           x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], x[10], x[11], x[12], x[13], x[14],
                           x[15], o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], o[8], o[9], o[10], o[11]);
    }
    void vjp(const D* x, const D* g, D* gr) const
    {
        D o[12];
        adj_edge_pq<D>(eps, eps4, this->g, gp, r, mu, dt, area, len, n1x, n1y, intr, H0, H1, t1, t2, t3, t4, LoH,
                       x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], x[10], x[11], x[12], x[13], x[14],
                       x[15], o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], o[8], o[9], o[10], o[11], g[0], g[1],
                       g[2], g[3], g[4], g[5], g[6], g[7], g[8], g[9], g[10], g[11], gr[0], gr[1], gr[2], gr[3], gr[4],
                       gr[5], gr[6], gr[7], gr[8], gr[9], gr[10], gr[11], gr[12], gr[13], gr[14], gr[15]);
    }
};

struct CellPQ {
    D eps, eps4, r, g, mfc, mf0, mfs, dt, val, v1, v2, LoH;
    void fwd(const D* x, D* n) const
    {
        D z[6] = {0, 0, 0, 0, 0, 0}, go[6], ga[6], gt[4];
        adj_cell_pq<D>(eps, eps4, r, g, mfc, mf0, mfs, dt, val, v1, v2, LoH, x[0], x[1], x[2], x[3], x[4], x[5], x[6],
                       x[7], x[8], x[9], x[10], x[11], x[12], x[13], x[14], x[15], n[0], n[1], n[2], n[3], n[4], n[5],
                       z[0], z[1], z[2], z[3], z[4], z[5], go[0], go[1], go[2], go[3], go[4], go[5], ga[0], ga[1],
                       ga[2], ga[3], ga[4], ga[5], gt[0], gt[1], gt[2], gt[3]);
    }
    void vjp(const D* x, const D* gn, D* gr) const
    {
        D n[6];
        adj_cell_pq<D>(eps, eps4, r, g, mfc, mf0, mfs, dt, val, v1, v2, LoH, x[0], x[1], x[2], x[3], x[4], x[5], x[6],
                       x[7], x[8], x[9], x[10], x[11], x[12], x[13], x[14], x[15], n[0], n[1], n[2], n[3], n[4], n[5],
                       gn[0], gn[1], gn[2], gn[3], gn[4], gn[5], gr[0], gr[1], gr[2], gr[3], gr[4], gr[5], gr[6], gr[7],
                       gr[8], gr[9], gr[10], gr[11], gr[12], gr[13], gr[14], gr[15]);
    }
};

// one input of one sample: returns 1 when checked, and updates worst / bad
template <class Fn, int NI, int NO>
static int check_input(const Fn& f, const D* x, const D* g, const D* grad, int v, D& worst, long& bad)
{
    const D h = 1e-7 * std::max(1e-2, std::fabs(x[v]));   // smaller steps drown in round-off
    D xp[NI], xm[NI], op[NO], om[NO], oo[NO];
    std::copy(x, x + NI, xp);
    std::copy(x, x + NI, xm);
    xp[v] += h;
    xm[v] -= h;
    f.fwd(xp, op);
    f.fwd(xm, om);
    f.fwd(x, oo);
    D fc = 0, fp = 0, fm = 0;
    for (int k = 0; k < NO; k++) {
        fc += g[k] * (op[k] - om[k]) / (2 * h);
        fp += g[k] * (op[k] - oo[k]) / h;
        fm += g[k] * (oo[k] - om[k]) / h;
    }
    const D sc = std::max({std::fabs(fc), std::fabs(fp), std::fabs(fm)});
    if (sc < kFloor || std::fabs(fp - fm) > 1e-4 * sc) return 0;
    const D rel = std::fabs(grad[v] - fc) / sc;
    worst = std::max(worst, rel);
    if (rel > kBar) bad++;
    return 1;
}

int main(int argc, char** argv)
{
    const long iters = argc > 1 ? std::atol(argv[1]) : 20000;
    std::mt19937_64 rng(3);
    std::uniform_real_distribution<D> U(0, 1);
    auto depth = [&]() {
        const D q = U(rng);
        return q < 0.25 ? 0.0 : q < 0.5 ? 0.02 * U(rng) : q < 0.6 ? 0.004 * U(rng) : U(rng);
    };
    long ne = 0, nc = 0, bade = 0, badc = 0, nep = 0, ncp = 0, badep = 0, badcp = 0, ntan = 0;
    D worste = 0, worstc = 0, worstep = 0, worstcp = 0;
    for (long it = 0; it < iters; it++) {
        Edge e{5e-3, 6.25e-10, 9.81, 9.81 * 0.5, 0.5, 0.17 * U(rng), 0.02 + 0.2 * U(rng), 0.25, 0.5, 1, 0, 1,
               1.0 + 0.1 * U(rng), 0};
        if (U(rng) < 0.5) { e.n1x = 0; e.n1y = 1; }
        e.H1 = U(rng) < 0.5 ? e.H0 : 1.0 + 0.1 * U(rng);
        D x[16];   // h10 h20 h11 h21, q00..q03, q10..q13, a0h1 a0h2 a1h1 a1h2
        x[0] = depth(); x[2] = depth();
        x[1] = U(rng) < 0.5 ? 0.0 : depth(); x[3] = U(rng) < 0.5 ? 0.0 : depth();
        for (int k = 0; k < 4; k++) {
            const D hc0 = x[k < 2 ? 0 : 1], hc1 = x[k < 2 ? 2 : 3];
            x[4 + k] = hc0 > 0 ? (U(rng) - 0.5) * 2 * hc0 : 0;
            x[8 + k] = hc1 > 0 ? (U(rng) - 0.5) * 2 * hc1 : 0;
        }
        for (int k = 12; k < 16; k++) x[k] = (U(rng) - 0.5) * 2.0;   // partial sums that drive the limiter
        D g[12], gr[16];
        for (int k = 0; k < 12; k++) g[k] = U(rng) - 0.5;
        e.vjp(x, g, gr);
        for (int v = 0; v < 16; v++) {
            if (v < 4 && x[v] < 1e-3) continue;   // sqrt(h) is not differentiable at h = 0
            ne += check_input<Edge, 16, 12>(e, x, g, gr, v, worste, bade);
        }

        Cell c{5e-3, 6.25e-10, 0.5, 9.81, 0.08 * U(rng), 0.03 * U(rng), 0.03 * U(rng), 0.02 + 0.2 * U(rng), 0, 0, 0};
        c.val = c.dt / 0.25;
        if (U(rng) < 0.3) { c.v1 = 0.3 * U(rng); c.v2 = 0.3 * U(rng); }
        D y[13];   // o0..o5 a0..a5 mu
        y[0] = depth(); y[3] = U(rng) < 0.4 ? 0.0 : depth();
        for (int k : {1, 2}) y[k] = y[0] > 0 ? (U(rng) - 0.5) * y[0] : 0;
        for (int k : {4, 5}) y[k] = y[3] > 0 ? (U(rng) - 0.5) * y[3] : 0;
        for (int k = 6; k < 12; k++) y[k] = (U(rng) - 0.5) * 0.05;
        y[12] = 0.2 * U(rng);
        D gn[6], gc[13];
        for (int k = 0; k < 6; k++) gn[k] = U(rng) - 0.5;
        c.vjp(y, gn, gc);
        for (int v = 0; v < 13; v++) {
            if ((v == 0 || v == 3) && y[v] < 1e-3) continue;
            nc += check_input<Cell, 13, 6>(c, y, gn, gc, v, worstc, badc);
        }

        // Pouliquen: angles 5-30 deg for stop / start, the same edge state
        const D ta = std::tan((5 + 10 * U(rng)) * M_PI / 180), tb = std::tan((15 + 15 * U(rng)) * M_PI / 180);
        const D tc = std::tan((5 + 10 * U(rng)) * M_PI / 180), td = std::tan((15 + 15 * U(rng)) * M_PI / 180);
        EdgePQ ep{e.eps, e.eps4, e.g, e.gp, e.r, e.mu, e.dt, e.area, e.len, e.n1x, e.n1y, e.intr, e.H0, e.H1,
                  ta, tb, tc, td, 1.0};
        D grp[16];
        ep.vjp(x, g, grp);
        for (int v = 0; v < 16; v++) {
            if (v < 4 && x[v] < 1e-3) continue;
            nep += check_input<EdgePQ, 16, 12>(ep, x, g, grp, v, worstep, badep);
        }
        // cells at rest (q = 0), slow and fast: the three regimes of the Froude number
        CellPQ cp{c.eps, c.eps4, c.r, c.g, c.mfc, c.mf0, c.mfs, c.dt, c.val, c.v1, c.v2, 1.0};
        D yp[16];
        std::copy(y, y + 12, yp);
        const D sp = U(rng) < 0.2 ? 0.0 : U(rng) < 0.5 ? 0.1 : 2.0;
        for (int k : {1, 2}) yp[k] *= sp;
        for (int k : {4, 5}) yp[k] *= sp;
        yp[12] = ta; yp[13] = tb; yp[14] = tc; yp[15] = td;
        D gcp[16];
        cp.vjp(yp, gn, gcp);
        for (int v = 0; v < 16; v++) {
            if ((v == 0 || v == 3) && yp[v] < 1e-3) continue;
            const int k = check_input<CellPQ, 16, 6>(cp, yp, gn, gcp, v, worstcp, badcp);
            ncp += k;
            if (v >= 12) ntan += k;
        }
    }
    std::printf("adj_edge checked %ld worst %.3e bad %ld\n", ne, worste, bade);
    std::printf("adj_cell checked %ld worst %.3e bad %ld\n", nc, worstc, badc);
    std::printf("adj_edge_pq checked %ld worst %.3e bad %ld\n", nep, worstep, badep);
    std::printf("adj_cell_pq checked %ld (angle tangents %ld) worst %.3e bad %ld\n", ncp, ntan, worstcp, badcp);
    return (bade || badc || badep || badcp) ? 1 : 0;
}
