"""The oracle's pins catch the slips they are meant to catch (CPU only).

Each case copies oracle/lhysea_oracle.c into a temporary directory, plants
one plausible mistake by a textual substitution (exactly one match), builds
the copy with the oracle's own flags and runs the known-answer tests that pin
that part of the oracle against it: at least one of them must fail.  The
unmodified oracle passes the same tests in test_oracle_kat.py.

Planted mistakes:
  * disImplicita (src/GPU/Volumen_kernel.cu:64-89) read as exclusive cases,
    the water-bottom case taking the state from before the coupled case, the
    sediment-only case under the wrong coefficient or condition;
  * the initial-dt edge speed (src/GPU/Arista_kernel.cu:1107-1113) with |.| on
    the signed eigenvalues, or without the |u_ij| term;
  * the speed cap (src/GPU/Volumen_kernel.cu:150-179) scaling q by vmax/u
    instead of storing q = u h.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "lhysea_oracle.c")
CFLAGS = ["-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared"]

# name: (text in the oracle, planted text, KAT tests (name, kwargs) that must catch it)
PLANTS = {
    "dis_water_case_exclusive": (
        "    if ((h1 > 0) && (h2 < eps)) {",
        "    else if ((h1 > 0) && (h2 < eps)) {",
        [("test_dis_implicita_interlayer_then_water_manning", {})]),
    "dis_water_case_stale_discharge": (
        "        u1x = LH_SQRT2 * n1[1] * h1 / h1m; u1y = LH_SQRT2 * n1[2] * h1 / h1m;\n"
        "        c1 = dt * g * mf0",
        "        c1 = dt * g * mf0",
        [("test_dis_implicita_interlayer_then_water_manning", {})]),
    "dis_sediment_case_needs_dry_water": (
        "    if ((h2 > 0) && (h1 < eps)) {",
        "    if ((h2 > 0) && (h1 <= 0)) {",
        [("test_dis_implicita_interlayer_then_sediment_manning", {})]),
    "dis_sediment_case_mf0": (
        "c1 = dt * g * mfs * mfs * u2",
        "c1 = dt * g * mf0 * mf0 * u2",
        [("test_dis_implicita_sediment_only_manning", {})]),
    "dt0_abs_eigenvalues": (
        "    maxl = fmax(D[0], D[3]);",
        "    maxl = fmax(fabs(D[0]), fabs(D[3]));",
        [("test_initial_dt_signed_speed", {"ufrac": -0.05}), ("test_initial_dt_signed_speed", {"ufrac": -0.7})]),
    "dt0_no_velocity_term": (
        "    if (b > maxl) maxl = b;",
        "    (void)b;",
        [("test_initial_dt_signed_speed", {"ufrac": -0.05}), ("test_initial_dt_signed_speed", {"ufrac": -0.7})]),
    "vmax_scales_discharge": (
        "ux *= p->vmax1 / u; uy *= p->vmax1 / u; n1[1] = ux * h; n1[2] = uy * h;",
        "n1[1] *= p->vmax1 / u; n1[2] *= p->vmax1 / u;",
        [("test_vmax_cap_sets_discharge_to_vmax_times_depth", {"h": 2e-3}),
         ("test_vmax_cap_sets_discharge_to_vmax_times_depth", {"h": 4e-4})]),
}


def _build_variant(tmp_path, old, new):
    text = open(SRC).read()
    assert text.count(old) == 1, f"plant target not unique: {old!r}"
    src = tmp_path / "oracle_planted.c"
    src.write_text(text.replace(old, new))
    so = tmp_path / "liboracle_planted.so"
    subprocess.run(["gcc"] + CFLAGS + ["-o", str(so), str(src), "-lm"], check=True)
    return str(so)


@pytest.mark.parametrize("name", sorted(PLANTS))
def test_planted_slip_is_caught(oracle, tmp_path, monkeypatch, name):
    import test_oracle_kat as K
    old, new, checks = PLANTS[name]
    variant = oracle.open_library(_build_variant(tmp_path, old, new))
    monkeypatch.setattr(oracle, "_lib", variant)
    caught = []
    for test_name, kwargs in checks:
        try:
            getattr(K, test_name)(oracle, **kwargs)
        except AssertionError:
            caught.append((test_name, kwargs))
    assert caught, f"planted slip {name} passed every pin: {checks}"


@pytest.mark.parametrize("layer", [1, 2])
def test_thin_front_case_separates_the_cap_forms(oracle, tmp_path, monkeypatch, layer):
    """The GPU cap test (test_gpu_reference_params.py) holds FP32C to 1e-5 on
    q and 1e-6 in the thin cells; the cap that scales q by vmax/u (no h s(h)
    factor) lands more than 3x those bars away from the oracle on that case."""
    import numpy as np
    from tests.cases import oracle_params, rel_l2, thin_front_case
    W, H, prm, steps = thin_front_case(layer)
    k = 0 if layer == 1 else 3
    Wo = W.copy()
    oracle.run(oracle_params(oracle, prm), Wo, H, steps)
    src = open(SRC).read()
    for v, n in (("vmax1", "n1"), ("vmax2", "n2")):
        old = f"ux *= p->{v} / u; uy *= p->{v} / u; {n}[1] = ux * h; {n}[2] = uy * h;"
        assert src.count(old) == 1
        src = src.replace(old, f"{n}[1] *= p->{v} / u; {n}[2] *= p->{v} / u;")
    (tmp_path / "o.c").write_text(src)
    so = tmp_path / "o.so"
    subprocess.run(["gcc"] + CFLAGS + ["-o", str(so), str(tmp_path / "o.c"), "-lm"], check=True)
    monkeypatch.setattr(oracle, "_lib", oracle.open_library(str(so)))
    Wb = W.copy()
    oracle.run(oracle_params(oracle, prm), Wb, H, steps)
    thin = (Wo[k] > 0) & (Wo[k] < prm.eps_h)
    assert rel_l2(Wb[k + 1], Wo[k + 1]) > 3e-5
    assert np.abs(Wb[k + 1] - Wo[k + 1])[thin].max() / np.abs(Wo[k + 1]).max() > 3e-6
