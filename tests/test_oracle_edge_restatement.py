"""Two-layer edge of the oracle vs an independent NumPy restatement (CPU).

The restatement below is written from the model's equations in vector form
(rotate, dry/wet, Roe-type averages, hydrostatic jumps, the two-layer
eigenvalue estimates with the reference's sort network, PVM coefficients,
flux + viscosity, tangential upwinding, positivity) without sharing code with
the C oracle.  Single-layer closed forms cannot see the internal eigenvalue
pair (it vanishes when h2 = 0); this comparison on random two-layer states
does, and it is what caught the oracle writing its eigenvalue output over its
own input.
"""
import math

import numpy as np
import pytest

EPS = float(np.finfo(np.float32).eps)
S2 = math.sqrt(2.0)
G = 9.81


def U(h, q, e):
    return S2 * h * q / math.sqrt(h ** 4 + max(h, e) ** 4)


def eig(f, r, flag, e):
    h1, a, h2, b = f
    gp = G * (1 - r)
    h = h1 + h2
    hd = math.sqrt(h ** 4 + max(h, e) ** 4)
    q = (a + b) if flag else (h1 * a + h2 * b)
    u = S2 * q * h / hd
    c = math.sqrt(G * h)
    if flag:
        u1, u2 = U(h1, a, e), U(h2, b, e)
        ub = S2 * (u1 * h2 + u2 * h1) * h / hd
    else:
        u1, u2 = a, b
        ub = S2 * (h1 * b + a * h2) * h / hd
    ci = math.sqrt(gp * h1 * h2 * S2 * h / hd * abs(1 - S2 * (u1 - u2) ** 2 * h / (gp * hd)))
    D = [u - c, ub - ci, ub + ci, u + c]
    for (i, j) in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3)):
        if D[i] > D[j]:
            D[i], D[j] = D[j], D[i]
    if D[2] > D[1]:
        D[1], D[2] = D[2], D[1]
    if D[2] > D[3]:
        D[2], D[3] = D[3], D[2]
    return D


def edge_np(W0, W1, H0, H1, n, ln, dt, acc0, acc1, r, mu, area, e, interior=True):
    n = np.array(n, dtype=float)
    rot = lambda W: np.array([W[0], W[1:3] @ n, W[3], W[4:6] @ n])
    w0, w1 = rot(W0), rot(W1)
    if w0[2] < e and H1 - w1[2] > H0: w1[3] = 0
    if w1[2] < e and H0 - w0[2] > H1: w0[3] = 0
    if w0[0] < e and H1 - w1[0] - w1[2] > H0 - w0[2]: w1[1] = 0
    if w1[0] < e and H0 - w0[0] - w0[2] > H1 - w1[2]: w0[1] = 0
    if w0[2] + w0[0] < e and H1 - w1[0] - w1[2] > H0: w1[1] = w1[3] = 0
    if w1[2] + w1[0] < e and H0 - w0[0] - w0[2] > H1: w0[1] = w0[3] = 0
    hij = 0.5 * (w0[[0, 2]] + w1[[0, 2]])
    uij = np.array([(math.sqrt(w0[k]) * U(w0[k], w0[k + 1], e) + math.sqrt(w1[k]) * U(w1[k], w1[k + 1], e))
                    / (math.sqrt(w0[k]) + math.sqrt(w1[k]) + EPS) for k in (0, 2)])
    a0, a1 = np.array(acc0, float), np.array(acc1, float)
    if not (hij[0] >= EPS or hij[1] >= EPS):
        return a0, a1
    Hm = min(H0, H1)
    d1 = max(w1[0] + w1[2] - H1 + Hm, 0) - max(w0[0] + w0[2] - H0 + Hm, 0)
    d2 = max(w1[2] - H1 + Hm, 0) - max(w0[2] - H0 + Hm, 0)
    tp = np.array([0, G * hij[0] * d1, 0, G * hij[1] * ((1 - r) * d2 + r * d1)])
    sc = (1 - r * (1 - math.exp(-(10 * hij[0] / e) ** 2))) * mu * G * hij[1]
    coul = abs(hij[1] * uij[1]) < sc * dt
    tp2 = tp.copy()
    if coul:
        tp2[3] = G * hij[1] * r * d1
    D0 = eig([hij[0], uij[0], hij[1], uij[1]], r, 0, e)
    DL, DR = eig(w0, r, 1, e), eig(w1, r, 1, e)
    l1, l3 = min(D0[0], DL[0], DR[0]), max(D0[3], DL[3], DR[3])
    lmax = max(abs(l1), abs(l3), abs(uij[0]), abs(uij[1]))
    l2 = max(abs(D0[1]), abs(D0[2]))
    if not abs(l1 + l3) < EPS and l1 + l3 < 0:
        l2 = -l2
    if abs(l1 - l3) >= EPS and abs(l1 - l2) >= EPS and abs(l2 - l3) >= EPS:
        if hij[0] >= e and hij[1] >= e:
            lam = np.array([l1, l2, l3])
            c = np.linalg.solve(np.vander(lam, 3, increasing=True), np.abs(lam))
        else:
            c = np.array([(l3 * abs(l1) - l1 * abs(l3)) / (l3 - l1), (abs(l3) - abs(l1)) / (l3 - l1), 0.0])
    else:
        c = np.array([lmax, 0.0, 0.0])
    F = lambda w: np.array([w[1], 0 if w[0] < EPS else U(w[0], w[1], e) * w[1],
                            w[3], 0 if w[2] < EPS else U(w[2], w[3], e) * w[3]])
    FL, FR = F(w0), F(w1)
    dF = FR - FL
    tp2 = tp2 + dF
    u1, u2 = uij
    ga, gb = G * hij[0], G * hij[1]
    A = np.array([[0, 1, 0, 0], [ga - u1 ** 2, 2 * u1, ga, 0], [0, 0, 0, 1], [r * gb, 0, gb - u2 ** 2, 2 * u2]])
    I2 = np.array([d1, w1[1] - w0[1], d2, w1[3] - w0[3]])
    if coul:
        I2[2] = 0
    elif w0[0] > e and w1[0] > e:
        I2[0] -= d2
    Dv = 0.5 * (c[0] * I2 + c[1] * tp2 + c[2] * (A @ tp2))
    M = 0.5 * (dF + tp)
    Fm, Fp = M + FL - Dv, M + Dv - FR
    tang = lambda W, k: W[k + 2] * n[0] - W[k + 1] * n[1]
    ut = []
    for lay, k in ((0, 0), (2, 3)):
        if abs(Fm[lay]) < EPS:
            ut.append(0.0)
        elif Fm[lay] > 0:
            ut.append(U(w0[lay], tang(W0, k), e))
        else:
            ut.append(U(w1[lay], tang(W1, k), e))
    T = Fm[[0, 2]] * np.array(ut)
    back = lambda F4, s: np.array([F4[0], F4[1] * n[0] - s * T[0] * n[1], F4[1] * n[1] + s * T[0] * n[0],
                                   F4[2], F4[3] * n[0] - s * T[1] * n[1], F4[3] * n[1] + s * T[1] * n[0]]) * ln
    Fm6, Fp6 = back(Fm, 1), back(Fp, -1)
    for lay, sl in ((0, slice(0, 3)), (3, slice(3, 6))):
        cand = []
        if Fm6[lay] > 0:
            cand.append((W0[lay] + dt / area * acc0[lay]) * area / (Fm6[lay] + EPS))
        if interior and Fp6[lay] > 0:
            cand.append((W1[lay] + dt / area * acc1[lay]) * area / (Fp6[lay] + EPS))
        dtl = min(cand) if cand else 1e30
        if not dt <= dtl:
            Fm6[sl] *= dtl / (dt + EPS)
            Fp6[sl] *= dtl / (dt + EPS)
    lmax = lmax + e if lmax < e else lmax
    a0[:6] -= Fm6
    a0[6] += ln * lmax
    if interior:
        a1[:6] -= Fp6
        a1[6] += ln * lmax
    return a0, a1


@pytest.mark.parametrize("seed", range(6))
def test_two_layer_edge_matches_restatement(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    p = oracle.make_params(8, 8, 50.0, 40.0, angle_deg=3.0, r=0.45)
    mu = math.tan(math.radians(3.0))
    for trial in range(300):
        W0 = np.array([rng.uniform(0, 40) * (rng.random() > 0.15), rng.normal(0, 4), rng.normal(0, 4),
                       rng.uniform(0, 12) * (rng.random() > 0.25), rng.normal(0, 2), rng.normal(0, 2)])
        W1 = np.array([rng.uniform(0, 40) * (rng.random() > 0.15), rng.normal(0, 4), rng.normal(0, 4),
                       rng.uniform(0, 12) * (rng.random() > 0.25), rng.normal(0, 2), rng.normal(0, 2)])
        H0, H1 = rng.uniform(0, 60, 2)
        n = [(1, 0), (0, 1), (-1, 0), (0, -1)][trial % 4]
        ln = 40.0 if n[1] == 0 else 50.0
        interior = trial % 5 != 0
        acc0, acc1 = rng.normal(0, 50, 7), rng.normal(0, 50, 7)
        dt = rng.uniform(0.05, 3.0)
        wet, b0, b1 = oracle.edge(p, W0, W1, H0, H1, n[0] * ln, n[1] * ln, ln, dt, acc0, acc1, interior)
        a0, a1 = edge_np(W0, W1, H0, H1, n, ln, dt, acc0, acc1, 0.45, mu, 2000.0, 5e-3, interior)
        scale = 1.0 + np.abs(a0).max() + np.abs(a1).max()
        assert np.abs(a0 - b0).max() <= 1e-9 * scale, (trial, a0 - b0)
        if interior:
            assert np.abs(a1 - b1).max() <= 1e-9 * scale, (trial, a1 - b1)
