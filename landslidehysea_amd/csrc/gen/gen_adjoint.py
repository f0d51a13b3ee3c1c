"""Generate the forward+reverse device functions of the discrete adjoint.

Writes landslidehysea_amd/csrc/lhysea_adjoint_gen.inc from two programs that
restate, in the same arithmetic as the FP32 step (lhysea_fp32.hpp), the edge
of the two-layer path-conservative solver (src/GPU/Arista_kernel.cu:324-645)
and the cell update (src/GPU/Volumen_kernel.cu:7-180, 231-284).  Branches
(wet/dry, Coulomb stop, wave-speed argmax, sort swaps, positivity, clipping)
are frozen at the forward choices.  Denominators of branches the forward may
not take are made safe so that every forward value stays finite.

Run: python landslidehysea_amd/csrc/gen/gen_adjoint.py
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from adgen import Prog, emit, select, sqrt, rsq, rcp, exp, fabs, fmin, fmax, powc, tof  # noqa: E402

EPS = 1.1920928955078125e-07
SQRT2 = 1.4142135623730951
OUT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lhysea_adjoint_gen.inc")


def sdes(h, eps, eps4):
    """U(h,q) = q * sdes(h): 1/h for h >= eps, else sqrt2 h / sqrt(h^4 + eps^4)."""
    big = h >= eps
    h2 = h * h
    return select(big, rcp(select(big, h, 1.0)), SQRT2 * h * rsq(h2 * h2 + eps4))


def hsdes(h, eps, eps4):
    """h * sdes(h): 1 for h >= eps, else sqrt2 h^2 / sqrt(h^4 + eps^4)."""
    big = h >= eps
    h2 = h * h
    return select(big, 1.0, SQRT2 * h2 * rsq(h2 * h2 + eps4))


def sort_net(D):
    x, y, z, w = D

    def cs(a, b):
        return fmin(a, b), fmax(a, b)

    x, y = cs(x, y)
    x, z = cs(x, z)
    x, w = cs(x, w)
    y, z = cs(y, z)
    y, w = cs(y, w)
    lo, hi = cs(y, z)
    y, z = hi, lo                      # the reference's inverted (2,1) step
    z, w = cs(z, w)
    return x, y, z, w


def eig(f, flag, g, gp, eps, eps4):
    f0, f1, f2, f3 = f
    h = f0 + f2
    sh = sdes(h, eps, eps4)
    q = (f1 + f3) if flag else (f0 * f1 + f2 * f3)
    u = q * sh
    cg = sqrt(g * h)
    if flag == 0:
        uu = (f0 * f3 + f1 * f2) * sh
        u1, u2 = f1, f3
    else:
        u1 = f1 * sdes(f0, eps, eps4)
        u2 = f3 * sdes(f2, eps, eps4)
        uu = (u1 * f2 + u2 * f0) * sh
    du = u1 - u2
    aux = 1.0 - du * du * sh * rcp(gp)
    ci = sqrt(gp * f0 * f2 * sh * fabs(aux))
    return sort_net([u - cg, uu - ci, uu + ci, u + cg])


def friction(h1, u1, h2, u2, tans, LoverH, eps, eps4, r, g):
    """Pouliquen friction coefficient (src/GPU/Arista_kernel.cu:20-58) in the
    arithmetic of the FP32 step (friction32): the reduced-gravity Froude
    number of the two layers, then the stop / start coefficients of the four
    angles' tangents, in the three regimes of the Froude number."""
    tan1, tan2, tan3, tan4 = tans
    one = h1.prog.const(1.0)
    x = 10.0 * h1 * rcp(eps)
    rr = 1.0 - r * (1.0 - exp(-(x * x)))
    igp = rcp(g * rr)
    fr1 = u1 * u1 * sdes(h1, eps, eps4) * igp
    fr2 = u2 * u2 * sdes(h2, eps, eps4) * igp
    fr = sqrt(fr1 + fr2 + (1.0 - rr) * fr1 * fr2)
    beta, L2, chi = 0.136, 8.0e-4, 1.0e-3
    dstop, dstart = tan2 - tan1, tan4 - tan3
    den = rcp(1.0 + h2 * (1.0 / L2))
    fast = fr > beta
    still = fabs(fr) < EPS
    mid = (~fast) & (~still)
    mf_fast = tan1 + dstop * rcp(1.0 + beta * h2 * rcp(select(fast, fr, one) * L2))
    mf_still = tan3 + dstart * den
    mustop, mustart = tan1 + dstop * den, tan3 + dstart * den
    mf_mid = mustart + powc(select(mid, fr, one) * (1.0 / beta), chi) * (mustop - mustart)
    return select(fast, mf_fast, select(still, mf_still, mf_mid)) * LoverH


def build_edge(pouliquen=False):
    P = Prog("adj_edge")
    eps, eps4, g, gp, r, mu, dt, area, ln, n1x, n1y, intr = [
        P.param(n) for n in ("eps", "eps4", "g", "gp", "r", "mu", "dt", "area", "len", "n1x", "n1y", "intr")]
    H0, H1 = P.param("H0", "D"), P.param("H1", "D")
    if pouliquen:
        # the coefficient only decides the Coulomb stop flag (a frozen branch)
        tans = [P.param(n) for n in ("tan1", "tan2", "tan3", "tan4")]
        LoverH = P.param("LoverH")
    h10, h20, h11, h21 = [P.inp(n, "D") for n in ("h10", "h20", "h11", "h21")]
    q0 = [P.inp(f"q0{k}") for k in range(4)]
    q1 = [P.inp(f"q1{k}") for k in range(4)]
    a0h1, a0h2, a1h1, a1h2 = [P.inp(n) for n in ("a0h1", "a0h2", "a1h1", "a1h2")]
    interior = intr > 0.5
    w00, w02, w10, w12 = tof(h10), tof(h20), tof(h11), tof(h21)
    w01 = q0[0] * n1x + q0[1] * n1y
    w03 = q0[2] * n1x + q0[3] * n1y
    w11 = q1[0] * n1x + q1[1] * n1y
    w13 = q1[2] * n1x + q1[3] * n1y
    z = P.const(0.0)
    # dry/wet momentum treatment (FP64 depth comparisons)
    c = (w02 < eps) & (H1 - h21 > H0)
    w13 = select(c, z, w13)
    c = (w12 < eps) & (H0 - h20 > H1)
    w03 = select(c, z, w03)
    c = (w00 < eps) & (H1 - h11 - h21 > H0 - h20)
    w11 = select(c, z, w11)
    c = (w10 < eps) & (H0 - h10 - h20 > H1 - h21)
    w01 = select(c, z, w01)
    c = (w02 + w00 < eps) & (H1 - h11 - h21 > H0)
    w11 = select(c, z, w11)
    w13 = select(c, z, w13)
    c = (w12 + w10 < eps) & (H0 - h10 - h20 > H1)
    w01 = select(c, z, w01)
    w03 = select(c, z, w03)

    h1ij = 0.5 * (w00 + w10)
    h2ij = 0.5 * (w02 + w12)
    s00, s10, s02, s12 = [sdes(h, eps, eps4) for h in (w00, w10, w02, w12)]
    r00, r10, r02, r12 = [sqrt(h) for h in (w00, w10, w02, w12)]
    u1ij = (r00 * (w01 * s00) + r10 * (w11 * s10)) * rcp(r00 + r10 + EPS)
    u2ij = (r02 * (w03 * s02) + r12 * (w13 * s12)) * rcp(r02 + r12 + EPS)
    wet = (h1ij >= EPS) | (h2ij >= EPS)

    # hydrostatic reconstruction in FP64
    Hm = fmin(H0, H1)
    e10 = fmax(h10 + h20 - H0 + Hm, 0.0)
    e11 = fmax(h11 + h21 - H1 + Hm, 0.0)
    e20 = fmax(h20 - H0 + Hm, 0.0)
    e21 = fmax(h21 - H1 + Hm, 0.0)
    deta1 = tof(e11 - e10)
    deta2 = tof(e21 - e20)
    tp1 = g * h1ij * deta1
    tp3 = g * h2ij * ((1.0 - r) * deta2 + r * deta1)
    x = 10.0 * h1ij * rcp(eps)
    fsc = 1.0 - r * (1.0 - exp(-(x * x)))
    muc = friction(h1ij, u1ij, h2ij, u2ij, tans, LoverH, eps, eps4, r, g) if pouliquen else mu
    sc = fsc * muc * g * h2ij
    coul = fabs(h2ij * u2ij) < sc * dt
    t23 = select(coul, g * h2ij * r * deta1, tp3)

    Dx, Dy, Dz, Dw = eig([h1ij, u1ij, h2ij, u2ij], 0, g, gp, eps, eps4)
    Lx, _, _, Lw = eig([w00, w01, w02, w03], 1, g, gp, eps, eps4)
    Rx, _, _, Rw = eig([w10, w11, w12, w13], 1, g, gp, eps, eps4)
    aut1 = fmin(Dx, fmin(Lx, Rx))
    aut3 = fmax(Dw, fmax(Lw, Rw))
    maxl = fmax(fmax(fabs(aut3), fabs(aut1)), fmax(fabs(u1ij), fabs(u2ij)))
    aut2 = fmax(fabs(Dy), fabs(Dz))
    sg = aut1 + aut3
    flip = (~(fabs(sg) < EPS)) & (~(sg > 0.0))
    aut2 = select(flip, -aut2, aut2)
    d13, d12, d23 = aut1 - aut3, aut1 - aut2, aut2 - aut3
    distinct = (fabs(d13) >= EPS) & (fabs(d12) >= EPS) & (fabs(d23) >= EPS)
    thick = (h1ij >= eps) & (h2ij >= eps)
    quad = distinct & thick
    lin = distinct & (~thick)
    one = P.const(1.0)
    A = fabs(aut1) * rcp(select(quad, d13 * d12, one))
    B = -fabs(aut2) * rcp(select(quad, d12 * d23, one))
    C = fabs(aut3) * rcp(select(quad, d13 * d23, one))
    idl = rcp(select(lin, -d13, one))
    c0 = select(quad, A * aut2 * aut3 + B * aut1 * aut3 + C * aut1 * aut2,
                select(lin, (aut3 * fabs(aut1) - aut1 * fabs(aut3)) * idl, maxl))
    c1 = select(quad, -A * (aut2 + aut3) - B * (aut1 + aut3) - C * (aut1 + aut2),
                select(lin, (fabs(aut3) - fabs(aut1)) * idl, z))
    c2 = select(quad, A + B + C, z)

    FL1 = select(w00 < EPS, z, w01 * w01 * s00)
    FL3 = select(w02 < EPS, z, w03 * w03 * s02)
    FR1 = select(w10 < EPS, z, w11 * w11 * s10)
    FR3 = select(w12 < EPS, z, w13 * w13 * s12)
    dF0, dF1, dF2, dF3 = w11 - w01, FR1 - FL1, w13 - w03, FR3 - FL3
    t20, t21, t22 = dF0, tp1 + dF1, dF2
    t23 = t23 + dF3
    ga, gb = g * h1ij, g * h2ij
    A0 = t21
    A1 = (ga - u1ij * u1ij) * t20 + 2.0 * u1ij * t21 + ga * t22
    A2 = t23
    A3 = r * gb * t20 + (gb - u2ij * u2ij) * t22 + 2.0 * u2ij * t23
    I0 = select(coul, deta1, select((w00 > eps) & (w10 > eps), deta1 - deta2, deta1))
    I2 = select(coul, z, deta2)
    V0 = 0.5 * (c0 * I0 + c1 * t20 + c2 * A0)
    V1 = 0.5 * (c0 * dF0 + c1 * t21 + c2 * A1)
    V2 = 0.5 * (c0 * I2 + c1 * t22 + c2 * A2)
    V3 = 0.5 * (c0 * dF2 + c1 * t23 + c2 * A3)
    M0, M1, M2, M3 = 0.5 * dF0, 0.5 * (dF1 + tp1), 0.5 * dF2, 0.5 * (dF3 + tp3)
    Fm0, Fm1, Fm2, Fm3 = M0 + (w01 - V0), M1 + (FL1 - V1), M2 + (w03 - V2), M3 + (FL3 - V3)
    Fp1, Fp3 = M1 + (V1 - FR1), M3 + (V3 - FR3)

    ut1 = select(fabs(Fm0) < EPS, z, select(Fm0 > 0.0, (q0[1] * n1x - q0[0] * n1y) * s00,
                                            (q1[1] * n1x - q1[0] * n1y) * s10))
    ut2 = select(fabs(Fm2) < EPS, z, select(Fm2 > 0.0, (q0[3] * n1x - q0[2] * n1y) * s02,
                                            (q1[3] * n1x - q1[2] * n1y) * s12))
    T1, T2 = Fm0 * ut1, Fm2 * ut2
    Mn = [ln * Fm0, ln * (Fm1 * n1x - T1 * n1y), ln * (Fm1 * n1y + T1 * n1x),
          ln * Fm2, ln * (Fm3 * n1x - T2 * n1y), ln * (Fm3 * n1y + T2 * n1x)]
    Pp = [-Mn[0], ln * (Fp1 * n1x + T1 * n1y), ln * (Fp1 * n1y - T1 * n1x),
          -Mn[3], ln * (Fp3 * n1x + T2 * n1y), ln * (Fp3 * n1y - T2 * n1x)]
    # positivity with the partial accumulators (frozen branches)
    bq = dt * rcp(area)
    for lay, (hc0, hc1, p0, p1) in enumerate(((w00, w10, a0h1, a1h1), (w02, w12, a0h2, a1h2))):
        k = 3 * lay
        cm = Mn[k] > 0.0
        cp = interior & (Pp[k] > 0.0)
        da = (hc0 + bq * p0) * area * rcp(select(cm, Mn[k] + EPS, one))
        db = (hc1 + bq * p1) * area * rcp(select(cp, Pp[k] + EPS, one))
        dtl = select(cm, da, P.const(1e30))
        dtl = select(cp, fmin(dtl, db), dtl)
        lim = ~(dt <= dtl)
        al = select(lim, dtl * rcp(dt + EPS), one)
        for j in range(k, k + 3):
            Mn[j] = Mn[j] * al
            Pp[j] = Pp[j] * al
    for j in range(6):
        P.out(f"fm{j}", select(wet, Mn[j], z))
    for j in range(6):
        P.out(f"fp{j}", select(wet, Pp[j], z))
    return P


def build_cell(pouliquen=False):
    P = Prog("adj_cell")
    eps, eps4, r, g, mfc, mf0, mfs, dt, val, vmax1, vmax2 = [
        P.param(n) for n in ("eps", "eps4", "r", "g", "mfc", "mf0", "mfs", "dt", "val", "vmax1", "vmax2")]
    if pouliquen:
        LoverH = P.param("LoverH")
    o = [P.inp(f"o{k}") for k in range(6)]
    a = [P.inp(f"a{k}") for k in range(6)]
    if pouliquen:
        # the four angles' tangents: inputs, so that their adjoints come out
        tans = [P.inp(n) for n in ("tan1", "tan2", "tan3", "tan4")]
    else:
        mu = P.inp("mu")
    z = P.const(0.0)
    one = P.const(1.0)
    so1, so2 = sdes(o[0], eps, eps4), sdes(o[3], eps, eps4)
    uo1x, uo1y, uo2x, uo2y = o[1] * so1, o[2] * so1, o[4] * so2, o[5] * so2
    ddx, ddy = uo1x - uo2x, uo1y - uo2y
    du = sqrt(ddx * ddx + ddy * ddy)
    uo1 = sqrt(uo1x * uo1x + uo1y * uo1y)
    uo2 = sqrt(uo2x * uo2x + uo2y * uo2y)
    n = [o[k] + val * a[k] for k in range(6)]
    n[0] = select(n[0] < 0.0, z, n[0])
    n[3] = select(n[3] < 0.0, z, n[3])
    # filtroEstado
    for hk, qk in ((0, 1), (3, 4)):
        zz = n[hk] * rcp(eps)
        z2 = zz * zz
        aux = exp(-dt * rcp(z2 * z2 + EPS))
        n[qk] = n[qk] * aux
        n[qk + 1] = n[qk + 1] * aux
    # speed caps: q <- (q s(h)) (vmax / u) h (src/GPU/Volumen_kernel.cu:150-179)
    for hk, qk, vm in ((0, 1, vmax1), (3, 4, vmax2)):
        s = sdes(n[hk], eps, eps4)
        ux, uy = n[qk] * s, n[qk + 1] * s
        u = sqrt(ux * ux + uy * uy)
        cap = (vm > 0.0) & (u > vm)
        f = vm * rcp(select(cap, u, one)) * hsdes(n[hk], eps, eps4)
        n[qk] = select(cap, n[qk] * f, n[qk])
        n[qk + 1] = select(cap, n[qk + 1] * f, n[qk + 1])
    # disImplicita
    h1, h2 = n[0], n[3]
    s1, s2 = sdes(h1, eps, eps4), sdes(h2, eps, eps4)
    hmod = h2 + r * h1
    em = eps * (1.0 + r)
    bigm = hmod >= em
    m2 = hmod * hmod
    phi = select(bigm, rcp(select(bigm, hmod, one)), SQRT2 * hmod * rsq(m2 * m2 + (em * em) * (em * em)))
    case1 = (h1 > 0.0) & (h2 > 0.0)
    c1 = dt * h2 * phi * mfc * du
    c2 = dt * r * h1 * phi * mfc * du
    c3 = dt * g * mfs * mfs * uo2 * rcp(powc(h2, 4.0 / 3.0) + EPS)
    det = rcp((1.0 + c1) * (1.0 + c2 + c3) - c1 * c2)
    a11, a12 = h1 * s1 * (1.0 + c2 + c3) * det, h1 * c1 * s2 * det
    a21, a22 = h2 * c2 * s1 * det, h2 * s2 * (1.0 + c1) * det
    q1x, q2x = a11 * n[1] + a12 * n[4], a22 * n[4] + a21 * n[1]
    q1y, q2y = a11 * n[2] + a12 * n[5], a22 * n[5] + a21 * n[2]
    n[1], n[4], n[2], n[5] = [select(case1, v, w) for v, w in ((q1x, n[1]), (q2x, n[4]), (q1y, n[2]), (q2y, n[5]))]
    case2 = (h1 > 0.0) & (h2 < eps)
    cc = dt * g * mf0 * mf0 * uo1 * rcp(powc(h1, 4.0 / 3.0) + EPS)
    f = h1 * s1 * rcp(1.0 + cc)
    n[1], n[2] = select(case2, n[1] * f, n[1]), select(case2, n[2] * f, n[2])
    case3 = (h2 > 0.0) & (h1 < eps)
    cc = dt * g * mfs * mfs * uo2 * rcp(powc(h2, 4.0 / 3.0) + EPS)
    f = h2 * s2 * rcp(1.0 + cc)
    n[4], n[5] = select(case3, n[4] * f, n[4]), select(case3, n[5] * f, n[5])
    # coulomb
    xx = 10.0 * n[0] * rcp(eps)
    fsc = 1.0 - r * (1.0 - exp(-(xx * xx)))
    if pouliquen:
        # the coefficient of the state after the implicit friction (cell_update())
        u1 = sqrt(n[1] * n[1] + n[2] * n[2]) * sdes(n[0], eps, eps4)
        u2 = sqrt(n[4] * n[4] + n[5] * n[5]) * sdes(n[3], eps, eps4)
        muc = friction(n[0], u1, n[3], u2, tans, LoverH, eps, eps4, r, g)
    else:
        muc = mu
    sc = fsc * muc * g * n[3] * dt
    nq = sqrt(n[4] * n[4] + n[5] * n[5])
    go = (nq + sc >= EPS) & (n[3] > 0.0) & (nq >= sc)
    fco = select(go, nq * rcp(select(go, nq + sc, one)), z)
    n[4], n[5] = n[4] * fco, n[5] * fco
    for k in range(6):
        P.out(f"n{k}", n[k])
    return P


# inputs that are exactly zero at a sediment-free edge / cell, and the signs
# the parameters always have (eps_h > 0, g > 0, 0 < r < 1, dt > 0, ...)
EDGE_ZEROS = ("h20", "h21", "q02", "q03", "q12", "q13", "a0h2", "a1h2")
EDGE_POS = ("eps", "eps4", "g", "gp", "r", "dt", "area", "len")
EDGE_NONNEG = ("mu", "h10", "h11")
CELL_ZEROS = ("o3", "o4", "o5", "a3", "a4", "a5")
CELL_POS = ("eps", "eps4", "r", "g", "dt", "val")
CELL_NONNEG = ("mfc", "mf0", "mfs", "vmax1", "vmax2", "o0")


def main():
    text = ["// Generated by landslidehysea_amd/csrc/gen/gen_adjoint.py -- do not edit.",
            "// Forward recomputation + reverse sweep of the edge and the cell update",
            "// (branches frozen at the forward choices).", ""]
    text.append(emit(build_edge(), forward_only=True))
    text.append(emit(build_edge()))
    text.append(emit(build_cell()))
    # the same programs at states without sediment: every layer-2 input is 0
    # (the adjoints of those inputs are still produced)
    text.append("// Specialisations for edges / cells without sediment (h2 = q2 = 0, zero layer-2")
    text.append("// partial and flux sums): forward terms and reverse contributions that these")
    text.append("// zeros remove are dropped; results equal the general functions' up to the")
    text.append("// sign of zero results.")
    text.append(emit(build_edge(), name="adj_edge_ns", zeros=EDGE_ZEROS, positive=EDGE_POS, nonneg=EDGE_NONNEG))
    text.append(emit(build_cell(), name="adj_cell_ns", zeros=CELL_ZEROS, positive=CELL_POS, nonneg=CELL_NONNEG))
    # the Pouliquen law (friction_law 1): the edge's Coulomb flag and the cell's
    # Coulomb scaling take the Pouliquen coefficient; the cell returns the
    # adjoints of the four angles' tangents.  Edges and cells without sediment
    # do not depend on the coefficient and keep the specialisations above.
    text.append("// The Pouliquen law (friction_law 1).")
    text.append(emit(build_edge(pouliquen=True), forward_only=True, name="adj_edge_pq"))
    text.append(emit(build_edge(pouliquen=True), name="adj_edge_pq"))
    text.append(emit(build_cell(pouliquen=True), name="adj_cell_pq"))
    with open(OUT, "w") as f:
        f.write("\n".join(text))
    print(OUT)


if __name__ == "__main__":
    main()
