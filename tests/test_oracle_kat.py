"""Known-answer tests that pin the FP64 oracle (CPU only).

The reference ships no tests, golden vectors or sample outputs, so every
function of the oracle is pinned here by what the mathematics fixes: closed
forms (single-layer eigenvalues, uniform-flow flux, the PVM interpolation
property, the implicit-friction 2x2 system, Coulomb thresholds, the CFL dt of
a uniform state), invariants (lake at rest, mass conservation with walls,
x-mirror symmetry) and a textbook solution (Stoker's wet dam break).  Each
check is chosen so that a plausible slip (a dropped pressure term, a sign or
normal error, an un-inverted sort step, a transposed friction coefficient)
fails it.
"""
import math

import numpy as np
import pytest

EPS = float(np.float32(np.finfo(np.float32).eps))
G = 9.81


def prm(oracle, nx=4, ny=4, dx=50.0, dy=40.0, **kw):
    return oracle.make_params(nx, ny, dx, dy, **kw)


# ---------------------------------------------------------------- eigenvalues
def test_eig_single_layer_closed_form(oracle):
    h, u = 12.0, 0.7
    D = oracle.eig([h, u, 0.0, 0.0], 0.5, G, 5e-3, 0)          # averaged form (h, u, h2, u2)
    c = math.sqrt(G * h)
    assert D[0] == pytest.approx(u - c, rel=1e-14)
    assert D[3] == pytest.approx(u + c, rel=1e-14)
    D1 = oracle.eig([h, u * h, 0.0, 0.0], 0.5, G, 5e-3, 1)     # conserved form (h, q, h2, q2)
    assert np.allclose(D1, D, rtol=1e-14, atol=1e-14)


def _unsorted(f, r, flag, eps=5e-3):
    """External pair u -+ sqrt(g h) and internal pair ubar -+ ci of the
    two-layer system written from the model equations (desingularised
    velocities equal q/h for h >= eps)."""
    h1, a, h2, b = f
    if flag == 0:
        u1, u2 = a, b
    else:
        u1, u2 = a / h1, b / h2
    h = h1 + h2
    u = (h1 * u1 + h2 * u2) / h
    ubar = (h1 * u2 + h2 * u1) / h
    gp = G * (1 - r)
    ci = math.sqrt(gp * h1 * h2 / h * abs(1 - (u1 - u2) ** 2 / (gp * h)))
    c = math.sqrt(G * h)
    return [u - c, ubar - ci, ubar + ci, u + c]


@pytest.mark.parametrize("f,flag", [([30.0, 0.4, 10.0, -0.2], 0), ([30.0, 12.0, 10.0, -2.0], 1),
                                    ([5.0, 0.0, 2.0, 0.0], 0)])
def test_eig_two_layer_order_with_inverted_step(oracle, f, flag):
    r = 0.5
    lm, im, ip, lp = _unsorted(f, r, flag)
    assert lm < im < ip < lp
    D = oracle.eig(f, r, G, 5e-3, flag)
    # the reference's network ends with an inverted (2,1) compare-swap, which
    # leaves the internal pair as (ubar + ci, ubar - ci)
    assert np.allclose(D, [lm, ip, im, lp], rtol=1e-12, atol=1e-12)


def test_eig_min_always_first(oracle):
    rng = np.random.default_rng(3)
    for _ in range(200):
        f = [rng.uniform(0, 50), rng.uniform(-5, 5), rng.uniform(0, 50), rng.uniform(-5, 5)]
        D = oracle.eig(f, 0.5, G, 5e-3, 0)
        assert D[0] == min(D)


# --------------------------------------------------------------- PVM and flux
def test_pvm_interpolates_abs_at_three_speeds(oracle):
    a1, a2, a3 = -11.0, 2.5, 9.0
    c = oracle.pvm_coeffs(a1, a2, a3, 10.0, 10.0, 11.0, 5e-3)
    for lam in (a1, a2, a3):
        assert c[0] + c[1] * lam + c[2] * lam ** 2 == pytest.approx(abs(lam), rel=1e-13)
    # thin layer: linear interpolation through (a1, |a1|), (a3, |a3|)
    c = oracle.pvm_coeffs(a1, a2, a3, 1e-3, 10.0, 11.0, 5e-3)
    assert c[2] == 0.0
    for lam in (a1, a3):
        assert c[0] + c[1] * lam == pytest.approx(abs(lam), rel=1e-13)
    # degenerate speeds: Rusanov fallback a0 = max speed
    c = oracle.pvm_coeffs(1.0, 1.0, 3.0, 10.0, 10.0, 7.0, 5e-3)
    assert list(c) == [7.0, 0.0, 0.0]


def test_flux1d_closed_form(oracle):
    F = oracle.flux1d([4.0, 3.0, 2.0, -1.0], 5e-3)
    assert np.allclose(F, [3.0, 9.0 / 4.0, -1.0, 0.5], rtol=1e-14)
    F = oracle.flux1d([0.0, 3.0, 1e-9, 2.0], 5e-3)   # dry layers: no momentum flux
    assert F[1] == 0.0 and F[3] == 0.0 and F[0] == 3.0 and F[2] == 2.0


# ------------------------------------------------------------------- edge
def test_dry_edge_does_nothing(oracle):
    p = prm(oracle)
    wet, a0, a1 = oracle.edge(p, np.zeros(6), np.zeros(6), 10.0, 12.0, 40.0, 0.0, 40.0, 0.1)
    assert wet == 0 and not a0.any() and not a1.any()


@pytest.mark.parametrize("normal", [(1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0)])
def test_uniform_flow_edge_is_the_physical_flux(oracle, normal):
    """Same state on both sides, flat bottom: cell 0 loses and cell 1 gains len*F(W).n."""
    p = prm(oracle, angle_deg=0.0)
    h1, qx, qy, h2, px, py = 8.0, 1.2, -0.4, 3.0, 0.3, 0.2
    W = np.array([h1, qx, qy, h2, px, py])
    length = 40.0
    nx, ny = normal
    wet, a0, a1 = oracle.edge(p, W, W, 20.0, 20.0, nx * length, ny * length, length, 1e-3)
    qn1, qn2 = qx * nx + qy * ny, px * nx + py * ny
    # normal flux of (h, q) is (q.n, q (q.n) / h) in x/y components
    F = np.array([qn1, qx * qn1 / h1, qy * qn1 / h1, qn2, px * qn2 / h2, py * qn2 / h2]) * length
    assert wet == 1
    assert np.allclose(a0[:6], -F, rtol=1e-12, atol=1e-12)
    assert np.allclose(a1[:6], F, rtol=1e-12, atol=1e-12)
    # wave-speed term: len * max(|u +- c|) of the two-layer system
    lam = max(abs(v) for v in _unsorted([h1, qn1, h2, qn2], 0.5, 1))
    assert a0[6] == pytest.approx(length * lam, rel=1e-12)


def test_edge_mass_flux_conservative(oracle):
    p = prm(oracle)
    W0 = np.array([10.0, 2.0, 0.5, 4.0, 0.6, 0.1])
    W1 = np.array([7.0, -1.0, 0.2, 1.0, 0.0, 0.3])
    wet, a0, a1 = oracle.edge(p, W0, W1, 20.0, 16.0, 40.0, 0.0, 40.0, 1e-3)
    assert a0[0] == pytest.approx(-a1[0], rel=1e-13)
    assert a0[3] == pytest.approx(-a1[3], rel=1e-13)
    assert a0[6] == a1[6] > 0


def test_edge_lake_at_rest_step_in_bathymetry(oracle):
    """Flat free surface over a bottom step (one side emerged): zero flux."""
    p = prm(oracle)
    # shifted depths: surface at level -S (S = 10); right cell is dry land above the sea
    S = 10.0
    W0 = np.array([30.0 - S, 0, 0, 0, 0, 0])
    W1 = np.zeros(6)
    wet, a0, a1 = oracle.edge(p, W0, W1, 30.0, 4.0, 40.0, 0.0, 40.0, 0.5)
    assert wet == 1
    assert not a0[:6].any() and not a1[:6].any()


def test_edge_pressure_drives_towards_lower_surface(oracle):
    p = prm(oracle, angle_deg=0.0)
    W0 = np.array([11.0, 0, 0, 0, 0, 0])
    W1 = np.array([10.0, 0, 0, 0, 0, 0])
    wet, a0, a1 = oracle.edge(p, W0, W1, 10.0, 10.0, 40.0, 0.0, 40.0, 1e-3)
    # mass leaves the higher cell 0 through +x and momentum points to +x in both cells
    assert a0[0] < 0 < a1[0]
    assert a0[1] > 0 and a1[1] > 0


def test_positivity_limits_outflow(oracle):
    """An outflow that would empty cell 0 within dt is scaled by alpha = dtA / (dt + EPS)."""
    p = prm(oracle, angle_deg=0.0)
    W0 = np.array([0.01, 1.0, 0.0, 0, 0, 0])
    W1 = np.array([0.01, 1.0, 0.0, 0, 0, 0])
    dt = 10.0
    area = p.dx * p.dy
    wet, a0, a1 = oracle.edge(p, W0, W1, 10.0, 10.0, 40.0, 0.0, 40.0, dt)
    free = 1.0 * 40.0          # unlimited outflow len * q
    dtA = 0.01 * area / (free + EPS)
    assert -a0[0] == pytest.approx(free * dtA / (dt + EPS), rel=1e-12)


# ----------------------------------------------------------------- cell ops
def test_coulomb_stop_and_slide(oracle):
    p = prm(oracle, angle_deg=30.0, r=0.5)
    mu = math.tan(math.radians(30.0))
    h1, h2, dt = 10.0, 2.0, 0.1
    sc = (1 - 0.5 * (1 - math.exp(-(10 * h1 / 5e-3) ** 2))) * mu * G * h2 * dt
    n1, n2 = oracle.coulomb(p, [h1, 0, 0], [h2, 0.5 * sc, 0.0], dt)       # below threshold: stop
    assert n2[1] == 0.0 and n2[2] == 0.0
    q = np.array([3.0 * sc, -1.0 * sc])
    n1, n2 = oracle.coulomb(p, [h1, 0, 0], [h2, q[0], q[1]], dt)
    N = np.hypot(*q)
    assert np.allclose(n2[1:], q * N / (N + sc), rtol=1e-13)


def test_dis_implicita_solves_the_coupled_system(oracle):
    p = prm(oracle, mfc=0.08, mfs=0.05, mf0=0.02, r=0.45)
    o1, o2 = [12.0, 3.0, -1.0], [4.0, 0.4, 0.2]
    n1, n2 = [11.5, 2.5, -0.8], [4.2, 0.7, 0.1]
    dt, r = 0.3, 0.45
    out1, out2 = oracle.dis_implicita(p, o1, o2, n1, n2, dt)
    h1, h2 = n1[0], n2[0]
    uo1, uo2 = np.array(o1[1:]) / o1[0], np.array(o2[1:]) / o2[0]
    du = np.linalg.norm(uo1 - uo2)
    hmod = h2 + r * h1
    hmodm = math.sqrt(hmod ** 4 + max(hmod, 5e-3 * (1 + r)) ** 4)
    c1 = dt * math.sqrt(2) * h2 * hmod / hmodm * 0.08 * du
    c2 = dt * r * math.sqrt(2) * h1 * hmod / hmodm * 0.08 * du
    c3 = dt * G * 0.05 ** 2 * np.linalg.norm(uo2) / (h2 ** (4 / 3) + EPS)
    v1, v2 = np.array(out1[1:]) / h1, np.array(out2[1:]) / h2
    u1, u2 = np.array(n1[1:]) / h1, np.array(n2[1:]) / h2
    # (1 + c1) v1 - c1 v2 = u1 ;  -c2 v1 + (1 + c2 + c3) v2 = u2
    assert np.allclose((1 + c1) * v1 - c1 * v2, u1, rtol=1e-12)
    assert np.allclose(-c2 * v1 + (1 + c2 + c3) * v2, u2, rtol=1e-12)


def test_dis_implicita_water_only_manning(oracle):
    p = prm(oracle, mf0=0.025)
    o1, o2 = [5.0, 2.0, 0.0], [0.0, 0.0, 0.0]
    n1, n2 = [5.0, 2.0, 1.0], [0.0, 0.0, 0.0]
    dt = 0.5
    out1, _ = oracle.dis_implicita(p, o1, o2, n1, n2, dt)
    c = dt * G * 0.025 ** 2 * (2.0 / 5.0) / (5.0 ** (4 / 3) + EPS)
    assert np.allclose(out1[1:], np.array(n1[1:]) / (1 + c), rtol=1e-13)


def test_filtro_damps_thin_layers_only(oracle):
    p = prm(oracle)
    n1, n2 = oracle.filtro(p, [1e-3, 1.0, 1.0], [10.0, 1.0, -1.0], 0.1)
    assert abs(n1[1]) < 1e-10                     # h << eps_h: momentum damped away
    assert n2[1] == pytest.approx(math.exp(-0.1 / ((10.0 / 5e-3) ** 4 + EPS)), rel=1e-15)
    n1, n2 = oracle.filtro(p, [-1.0, 1.0, 1.0], [1.0, 0, 0], 0.1)
    assert n1[0] == 0.0                           # negative depth clipped


def test_vmax_cap(oracle):
    p = prm(oracle, vmax1=2.0)
    n1, _ = oracle.filtro(p, [4.0, 12.0, 16.0], [0.0, 0, 0], 1e-9)
    assert math.hypot(n1[1], n1[2]) / 4.0 == pytest.approx(2.0, rel=1e-9)


@pytest.mark.parametrize("h", [4.0, 5e-3, 2e-3, 4e-4])
def test_vmax_cap_sets_discharge_to_vmax_times_depth(oracle, h):
    """The cap rescales the desingularised velocity to vmax and stores q = u h
    (src/GPU/Volumen_kernel.cu:150-179), so |q| = vmax h whatever h; below
    eps_h the velocity of the capped state, |q| s(h) = vmax h s(h), is then
    under vmax.  Scaling q itself by vmax/u (dropping the h s(h) factor)
    gives |q| = vmax / s(h), three times vmax h at h = eps_h / 2."""
    vmax1, vmax2, eps = 1.5, 0.8, 5e-3
    p = prm(oracle, vmax1=vmax1, vmax2=vmax2)
    s = math.sqrt(2) * h / math.sqrt(h ** 4 + max(h, eps) ** 4)     # U(h, q) = q s(h)
    q1, q2 = np.array([30.0, -40.0]) / s, np.array([-6.0, 8.0]) / s  # speeds 50 and 10
    assert 10.0 > max(vmax1, vmax2)
    dt = 1e-12                                    # no thin-layer damping
    n1, n2 = oracle.filtro(p, [h, *q1], [h, *q2], dt)
    assert math.hypot(n1[1], n1[2]) == pytest.approx(vmax1 * h, rel=1e-12)
    assert math.hypot(n2[1], n2[2]) == pytest.approx(vmax2 * h, rel=1e-12)
    assert n1[1] / n1[2] == pytest.approx(q1[0] / q1[1], rel=1e-12)    # direction kept
    assert n2[1] / n2[2] == pytest.approx(q2[0] / q2[1], rel=1e-12)


def test_dis_implicita_interlayer_then_water_manning(oracle):
    """h1 >= eps_h > h2 > 0 takes the coupled case (i) and then the water-bottom
    case (ii) (src/GPU/Volumen_kernel.cu:64-82, sequential ifs): (ii) divides
    (i)'s layer-1 output by 1 + c, c = dt g mf0^2 |u1_old| / (h1^(4/3) + EPS),
    and leaves layer 2 as (i) left it.  mf0 = 0 isolates (i) (its (ii) factor
    is h1 s(h1) = 1), whose output solves the 2x2 friction system."""
    eps, r, dt, mfc, mfs, mf0 = 5e-3, 0.45, 0.3, 0.08, 0.05, 0.03
    o1, o2 = [12.0, 3.0, -1.0], [2e-3, 1e-4, 5e-5]
    n1, n2 = [11.5, 2.5, -0.8], [3e-3, 2e-4, -1e-4]
    pa = prm(oracle, mfc=mfc, mfs=mfs, mf0=0.0, r=r)
    pb = prm(oracle, mfc=mfc, mfs=mfs, mf0=mf0, r=r)
    a1, a2 = oracle.dis_implicita(pa, o1, o2, n1, n2, dt)
    b1, b2 = oracle.dis_implicita(pb, o1, o2, n1, n2, dt)

    def U(h, q):                                  # desingularised velocity
        return math.sqrt(2) * h * np.asarray(q) / math.sqrt(h ** 4 + max(h, eps) ** 4)

    h1, h2 = n1[0], n2[0]
    uo1, uo2 = U(o1[0], o1[1:]), U(o2[0], o2[1:])
    du = np.linalg.norm(uo1 - uo2)
    hmod = h2 + r * h1
    phi = math.sqrt(2) * hmod / math.sqrt(hmod ** 4 + max(hmod, eps * (1 + r)) ** 4)
    c1, c2 = dt * h2 * phi * mfc * du, dt * r * h1 * phi * mfc * du
    c3 = dt * G * mfs ** 2 * np.linalg.norm(uo2) / (h2 ** (4 / 3) + EPS)
    v1, v2 = np.array(a1[1:]) / h1, np.array(a2[1:]) / h2
    u1, u2 = U(h1, n1[1:]), U(h2, n2[1:])
    assert np.allclose((1 + c1) * v1 - c1 * v2, u1, rtol=1e-12)
    assert np.allclose(-c2 * v1 + (1 + c2 + c3) * v2, u2, rtol=1e-12)
    c = dt * G * mf0 ** 2 * np.linalg.norm(uo1) / (h1 ** (4 / 3) + EPS)
    assert np.allclose(b1[1:], np.array(a1[1:]) / (1 + c), rtol=1e-13)
    assert np.array_equal(b2, a2)


def test_dis_implicita_sediment_only_manning(oracle):
    """h1 = 0 < h2: only the sediment-bottom case (iii) applies
    (src/GPU/Volumen_kernel.cu:83-89): q2 / (1 + c), c built from mfs (not
    mf0) and the old layer-2 speed; layer 1 untouched."""
    dt, mfs, mf0 = 0.4, 0.05, 0.02
    p = prm(oracle, mfc=0.08, mfs=mfs, mf0=mf0, r=0.5)
    o2, n2 = [4.0, 0.4, 0.2], [4.2, 0.7, 0.1]
    out1, out2 = oracle.dis_implicita(p, [0.0, 0.0, 0.0], o2, [0.0, 0.0, 0.0], n2, dt)
    c = dt * G * mfs ** 2 * math.hypot(0.4 / 4.0, 0.2 / 4.0) / (4.2 ** (4 / 3) + EPS)
    assert np.allclose(out2[1:], np.array(n2[1:]) / (1 + c), rtol=1e-13)
    assert list(out1) == [0.0, 0.0, 0.0]


def test_dis_implicita_interlayer_then_sediment_manning(oracle):
    """0 < h1 < eps_h <= h2 takes the coupled case (i) and then (iii).  With
    mfc = 0 the coupled case reduces to v2 = u2 / (1 + c3), and (iii) applies
    the same sediment-bottom factor again: q2 / (1 + c3)^2."""
    dt, mfs = 0.4, 0.05
    p = prm(oracle, mfc=0.0, mfs=mfs, mf0=0.02, r=0.5)
    o1, o2 = [1e-3, 1e-4, 0.0], [4.0, 0.4, 0.2]
    n1, n2 = [1e-3, 1e-4, 2e-5], [4.2, 0.7, 0.1]
    _, out2 = oracle.dis_implicita(p, o1, o2, n1, n2, dt)
    c3 = dt * G * mfs ** 2 * math.hypot(0.4 / 4.0, 0.2 / 4.0) / (4.2 ** (4 / 3) + EPS)
    assert np.allclose(out2[1:], np.array(n2[1:]) / (1 + c3) ** 2, rtol=1e-13)


def test_cell_update_dt_contribution(oracle):
    p = prm(oracle, cfl=0.5)
    _, d = oracle.cell_update(p, np.array([5.0, 0, 0, 0, 0, 0]), np.array([0, 0, 0, 0, 0, 0, 80.0]), 0.1)
    assert d == pytest.approx(2 * 0.5 * p.dx * p.dy / 80.0, rel=1e-15)
    _, d = oracle.cell_update(p, np.zeros(6), np.zeros(7), 0.1)
    assert d == 1e30


# ------------------------------------------------------------ grid-level
def _grid(nx, ny, seed=3, semi=(6.0, 4.0)):
    from landslidehysea_amd import scenario as S
    sc = S.make_scenario(nx, ny, seed=seed, semi_axes=semi)
    H, _ = S.hmin_shift(sc.H)
    return sc, np.ascontiguousarray(sc.state), np.ascontiguousarray(H)


def test_initial_dt_uniform_state(oracle):
    nx, ny, dx, dy = 6, 5, 30.0, 20.0
    p = oracle.make_params(nx, ny, dx, dy, cfl=0.5)
    W = np.zeros((6, ny, nx))
    W[0] = 9.0
    W[1] = 1.8                      # u = 0.2
    H = np.full((ny, nx), 9.0)
    dt = oracle.initial_dt(p, W, H)
    c = math.sqrt(G * 9.0)
    # interior cell: two x-edges (length dy) at u_ij + c, two y-edges (length dx) at c,
    # with the Roe-type average u_ij = (3 u + 3 u) / (3 + 3 + EPS); boundary x-edges
    # see the reflected ghost (u_ij = 0 there) and give smaller sums
    uij = (3.0 * 0.2 + 3.0 * 0.2) / (6.0 + EPS)
    A = 2 * dy * (uij + c) + 2 * dx * c
    assert dt == pytest.approx(2 * 0.5 * dx * dy / A, rel=1e-13)


@pytest.mark.parametrize("ufrac", [-0.05, -0.7])
def test_initial_dt_signed_speed(oracle, ufrac):
    """The initial edge speed is max(D_min, D_max, |u1ij|, |u2ij|) with the
    signed eigenvalues (src/GPU/Arista_kernel.cu:1107-1110): for a single layer
    flowing in -x, an interior x-edge gives max(u + c, |u|), not c + |u|.
    u = -0.05 c: u + c wins; u = -0.7 c: |u| wins.  Interior x-edges are then
    slower than c, so the largest sum belongs to a cell of the first or last
    column: one wall edge (reflected ghost, u_ij = 0: speed c), one interior
    x-edge and two y-edges at c."""
    nx, ny, dx, dy, h = 6, 5, 30.0, 20.0, 9.0
    c = math.sqrt(G * h)
    u = ufrac * c
    p = oracle.make_params(nx, ny, dx, dy, cfl=0.5)
    W = np.zeros((6, ny, nx))
    W[0] = h
    W[1] = u * h
    H = np.full((ny, nx), h)
    dt = oracle.initial_dt(p, W, H)
    uij = (3.0 * u + 3.0 * u) / (6.0 + EPS)
    lam = max(uij + c, abs(uij))
    A = dy * (c + lam) + 2 * dx * c
    assert lam < c
    assert dt == pytest.approx(2 * 0.5 * dx * dy / A, rel=1e-13)


def test_lake_at_rest_exact(oracle):
    sc, W, H = _grid(40, 30)
    H = np.round(H * 64) / 64
    sea = np.round(-float(sc.H.min()) * 64) / 64
    W[:] = 0.0
    W[0] = np.maximum(0.0, H - sea)
    p = oracle.make_params(40, 30, sc.dx, sc.dy)
    W0 = W.copy()
    oracle.run(p, W, H, 30)
    assert np.array_equal(W, W0)


def test_mass_conserved_with_walls(oracle):
    sc, W, H = _grid(40, 30)
    p = oracle.make_params(40, 30, sc.dx, sc.dy, angle_deg=2.0)
    m1, m2 = W[0].sum(), W[3].sum()
    oracle.run(p, W, H, 60)
    assert abs(W[0].sum() - m1) <= 1e-12 * m1
    assert abs(W[3].sum() - m2) <= 1e-12 * m2
    assert np.abs(W[4]).max() > 0          # the 2-degree slide moves


def test_x_mirror_symmetry(oracle):
    nx, ny = 40, 20
    sc, W, H = _grid(nx, ny)
    # make the case symmetric about the vertical centre line
    H = 0.5 * (H + H[:, ::-1])
    W[0] = 0.5 * (W[0] + W[0, :, ::-1])
    W[3] = 0.5 * (W[3] + W[3, :, ::-1])
    W[[1, 2, 4, 5]] = 0.0
    W[0] = W[0] + 0.05 * np.exp(-((np.arange(nx) - nx / 2 + 0.5) ** 2) / 20.0)[None, :]
    p = oracle.make_params(nx, ny, sc.dx, sc.dy, angle_deg=2.0)
    H = np.ascontiguousarray(H)
    W = np.ascontiguousarray(W)
    oracle.run(p, W, H, 40)
    M = W[:, :, ::-1].copy()
    M[[1, 4]] *= -1.0
    scale = np.abs(W).reshape(6, -1).max(axis=1)
    for k in range(6):
        assert np.abs(W[k] - M[k]).max() <= 1e-12 * max(scale[k], 1e-300) + 1e-300


def _stoker(hl, hr, x, t, g=G):
    """Stoker's wet-bed dam-break solution (dam at x = 0)."""
    cl, cr = math.sqrt(g * hl), math.sqrt(g * hr)
    # middle depth hm from the Rankine-Hugoniot shock + rarefaction matching
    def f(hm):
        cm = math.sqrt(g * hm)
        us = 2 * (cl - cm)                            # velocity behind rarefaction
        return us - (hm - hr) * math.sqrt(0.5 * g * (hm + hr) / (hm * hr))
    a, b = hr, hl
    for _ in range(200):
        m = 0.5 * (a + b)
        if f(a) * f(m) <= 0:
            b = m
        else:
            a = m
    hm = 0.5 * (a + b)
    cm = math.sqrt(g * hm)
    um = 2 * (cl - cm)
    s = hm * um / (hm - hr)                            # shock speed
    h = np.empty_like(x)
    for i, xi in enumerate(x):
        if xi <= -cl * t:
            h[i] = hl
        elif xi <= (um - cm) * t:
            h[i] = (2 * cl - xi / t) ** 2 / (9 * g)
        elif xi <= s * t:
            h[i] = hm
        else:
            h[i] = hr
    return h


def _dam_break_error(oracle, nx, width=400.0, hl=2.0, hr=1.0, t_end=20.0):
    ny = 2
    dx = width / nx
    x = (np.arange(nx) + 0.5) * dx - width / 2
    W = np.zeros((6, ny, nx))
    W[0] = np.where(x < 0, hl, hr)[None, :]
    H = np.full((ny, nx), 2.0)
    p = oracle.make_params(nx, ny, dx, dx, cfl=0.5, mfc=0.0, mf0=0.0, mfs=0.0, angle_deg=0.0)
    W = np.ascontiguousarray(W)
    t, dt = 0.0, 0.0
    while t < t_end:
        res = oracle.run(p, W, H, 1, dt=dt, t=t)
        t, dt = res["t"], res["dt_next"]
    ex = _stoker(hl, hr, x, t)
    assert np.array_equal(W[:, 0], W[:, 1])          # y-invariant problem stays y-invariant
    return np.abs(W[0, 0] - ex).sum() / np.abs(ex - hr).sum()


def test_single_layer_dam_break_converges_to_stoker(oracle):
    """h2 = 0 reduces the scheme to single-layer shallow water: wet dam break vs Stoker."""
    e1 = _dam_break_error(oracle, 200)
    e2 = _dam_break_error(oracle, 800)
    assert e2 < 0.05, e2
    assert e2 < 0.6 * e1, (e1, e2)


def test_pouliquen_friction_closed_forms(oracle):
    """Pouliquen law (src/GPU/Arista_kernel.cu:20-58) in its three regimes.
    With h1, h2 >> eps_h the reduced-gravity Froude numbers are
    Fr_i^2 = u_i^2 / (g (1 - r) h_i) and Fr^2 = Fr1^2 + Fr2^2 + r Fr1^2 Fr2^2:
    at rest mu = tan3 + (tan4 - tan3) / (1 + h2/L2); for Fr > beta
    mu = tan1 + (tan2 - tan1) / (1 + beta h2 / (Fr L2)); in between
    mu_start + (Fr/beta)^chi (mu_stop - mu_start); all times L/H."""
    import ctypes as C
    ang = (20.0, 32.0, 22.0, 35.0)
    r, Lr, Hr = 0.5, 2.0, 4.0
    p = prm(oracle, friction_law=1, angles_deg=ang, r=r, L=Lr, H=Hr)
    t1, t2, t3, t4 = [math.tan(math.radians(a)) for a in ang]
    beta, L2, chi = 0.136, 8.0e-4, 1.0e-3
    h1, h2 = 3.0, 0.02

    def mu(u1, u2):
        return oracle.lib().lh_friction(C.byref(p), h1, u1, h2, u2)

    def froude(u1, u2):
        a, b = u1 * u1 / (G * (1 - r) * h1), u2 * u2 / (G * (1 - r) * h2)
        return math.sqrt(a + b + r * a * b)

    assert mu(0.0, 0.0) == pytest.approx((t3 + (t4 - t3) / (1 + h2 / L2)) * Lr / Hr, rel=1e-14)
    F = froude(2.0, 0.5)
    assert F > beta
    assert mu(2.0, 0.5) == pytest.approx((t1 + (t2 - t1) / (1 + beta * h2 / (F * L2))) * Lr / Hr, rel=1e-12)
    F = froude(0.05, 0.002)
    assert 1e-7 < F < beta
    stop, start = t1 + (t2 - t1) / (1 + h2 / L2), t3 + (t4 - t3) / (1 + h2 / L2)
    assert mu(0.05, 0.002) == pytest.approx((start + (F / beta) ** chi * (stop - start)) * Lr / Hr, rel=1e-12)
    # Coulomb: |tan a1| L/H whatever the flow
    q = prm(oracle, friction_law=0, angles_deg=ang, L=Lr, H=Hr)
    assert oracle.lib().lh_friction(C.byref(q), h1, 2.0, h2, 0.5) == pytest.approx(t1 * Lr / Hr, rel=1e-15)
