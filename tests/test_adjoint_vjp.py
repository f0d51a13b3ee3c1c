"""The generated discrete-adjoint functions against finite differences of
their own forward (CPU only; tests/native/adjoint_vjp_check.cpp).

adj_edge<double> and adj_cell<double> (landslidehysea_amd/csrc/
lhysea_adjoint_gen.inc, from csrc/gen/gen_adjoint.py) on 20000 random edges
and cells each -- wet, thin and dry depths, both layers, speed caps on and
off, partial sums that drive the positivity limiter -- with random output
adjoints: every input adjoint within 2e-3 of the central difference wherever
the one-sided differences agree; the same for the Pouliquen law's
adj_edge_pq / adj_cell_pq, the cell's four angle tangents included (states
at rest, slow and fast: the law's three regimes).  A wrong reverse rule (the
square root's 1/2 planted as 0.45 in a copy of the generated code) fails it.
"""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "adjoint_vjp_check.cpp")
EMU = os.path.join(ROOT, "tests", "native", "emu")


def build_and_run(root, out):
    exe = os.path.join(out, "vjp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-w", "-I", EMU,
                    os.path.join(root, "tests", "native", "adjoint_vjp_check.cpp"), "-o", exe], check=True)
    return subprocess.run([exe, "20000"], capture_output=True, text=True, timeout=300)


def test_generated_adjoint_matches_finite_differences(tmp_path):
    r = build_and_run(ROOT, str(tmp_path))
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    for line in r.stdout.splitlines():
        assert int(line.split("checked")[1].split()[0]) > 100000, line
    tangents = [line for line in r.stdout.splitlines() if "angle tangents" in line]
    assert tangents and int(tangents[0].split("angle tangents")[1].split(")")[0]) > 1000


def test_planted_reverse_rule_is_caught(tmp_path):
    root = tmp_path / "copy"
    (root / "tests" / "native").mkdir(parents=True)
    shutil.copy(SRC, root / "tests" / "native")
    shutil.copytree(os.path.join(ROOT, "landslidehysea_amd", "csrc"), root / "landslidehysea_amd" / "csrc")
    inc = root / "landslidehysea_amd" / "csrc" / "lhysea_adjoint_gen.inc"
    text = inc.read_text()
    assert text.count("* F(0.5), ") > 10
    inc.write_text(text.replace("* F(0.5), ", "* F(0.45), "))
    r = build_and_run(str(root), str(tmp_path))
    print(r.stdout)
    assert r.returncode == 1
    assert any(not line.endswith(" bad 0") for line in r.stdout.splitlines())
