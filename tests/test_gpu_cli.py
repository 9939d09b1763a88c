"""GPU tests of the thermal3d CLI (the reference's main3d.cu modes): the
verify suite passes, `case` writes tmap_layer<k>.csv files that match the
oracle's field, `transient` writes transient.csv in the reference's format."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import thermal_amd as T
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "cuda-thermal_amd", "bin", "thermal3d")


def run(args, cwd, timeout=200):
    p = subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    print(p.stdout[-3000:], p.stderr[-2000:])
    return p


def test_verify_suite_passes(tmp_path):
    p = run([], tmp_path)
    assert p.returncode == 0
    assert "FAIL" not in p.stdout and p.stdout.count("PASS") == 21


def test_case_writes_temperature_maps(tmp_path):
    case = os.path.join(GOLDEN, "cases", "example")
    n = 48
    p = run(["case", case, str(n)], tmp_path)
    assert p.returncode == 0 and "hottest point" in p.stdout
    m = T.Model.load_case(case, n)
    a = m.arrays()
    ref = O.steady(a, tol=1e-10)["T"].reshape(a["nz"], n, n)
    plane_layer = m.plane_layer()
    for L in (0, 2):  # the active (die) layers of the example stack
        f = tmp_path / f"tmap_layer{L}.csv"
        assert f.exists()
        got = np.loadtxt(f, delimiter=",")
        want = ref[np.array(plane_layer) == L].max(axis=0)[::-1] + 25.0  # top row = largest y
        assert got.shape == (n, n)
        assert np.abs(got - want).max() < 2e-3


def test_transient_writes_csv(tmp_path):
    p = run(["transient", "32"], tmp_path)
    assert p.returncode == 0
    lines = (tmp_path / "transient.csv").read_text().splitlines()
    assert lines[0] == "t_s,dt_s,pcg_iters,T_max_C,heat_out_W,heat_stored_W"
    rows = np.array([[float(v) for v in l.split(",")] for l in lines[1:]])
    assert np.isclose(rows[-1, 0], 2.0) and np.all(np.diff(rows[:, 3]) >= -1e-9)
    P = 50.0
    assert np.abs(rows[:, 4] + rows[:, 5] - P).max() / P < 1e-5  # 6 printed decimals


def write_ice_results(m, x, folder, bump=None):
    """3D-ICE result files for the field x of case model m, as 3D-ICE writes
    them (K to 3 decimals): tmap_L<k>.txt with the die's top plane, row y = 0
    first, and tflp_max / tflp_avg with each block's max and mean. bump:
    (layer, i, j, K) added to one Tmap cell."""
    pl = m.plane_layer()
    x3 = x.reshape(m.nz, m.ny, m.nx)
    for L, layer in enumerate(m.layers()):
        if not layer[5]:
            continue
        top = x3[np.nonzero(pl == L)[0].max()] + 298.15
        if bump and bump[0] == L:
            top = top.copy()
            top[bump[2], bump[1]] += bump[3]
        with open(os.path.join(folder, f"tmap_L{L}.txt"), "w") as f:
            f.write("% hand-made Tmap\n")
            for j in range(m.ny):
                f.write("  ".join(f"{v:.3f}" for v in top[j]) + "\n")
        blocks = [(k, b) for k, b in enumerate(m.blocks()) if b[1] == L]
        for kind, pick in (("max", 0), ("avg", 1)):
            with open(os.path.join(folder, f"tflp_{kind}_L{L}.txt"), "w") as f:
                f.write("% Time(s) " + " ".join(f"{b[0]}(K)" for _, b in blocks) + "\n")
                f.write("0.000 " + " ".join(f"{m.block_temps(k, x)[pick] + 298.15:.3f}" for k, _ in blocks) + "\n")


def test_compare3dice(tmp_path):
    """compare3dice against 3D-ICE-format result files made from this
    solver's own FP64 field (so the differences are 3D-ICE's 3-decimal
    rounding) with one Tmap cell of the bottom die raised by 0.25 K: the
    report names that cell and that difference, and every block agrees to
    the rounding."""
    case = os.path.join(GOLDEN, "cases", "validate_3dice")
    out = tmp_path / "ice"
    out.mkdir()
    assert run(["export3dice", case, "100", str(out)], tmp_path).returncode == 0
    m = T.Model.load_case(case, 100)
    s = T.Solver(m, precision="fp64", true_residual=True)
    s.set_x_zero()
    s.set_rhs_steady()
    assert s.solve(1e-10, 300)["converged"]
    x = s.get_x()
    s.close()
    write_ice_results(m, x, out, bump=(1, 37, 12, 0.25))
    p = run(["compare3dice", case, "100", str(out)], tmp_path)
    assert p.returncode == 0
    lines = p.stdout.splitlines()
    diffs = [l for l in lines if "max |difference|" in l]
    assert len(diffs) == 2
    assert "C at x = 3.75 mm, y = 1.25 mm" in diffs[0]
    assert abs(float(diffs[0].split("=")[1].split()[0]) - 0.25) <= 0.0005
    assert float(diffs[1].split("=")[1].split()[0]) <= 0.0005
    rows = [l.split() for l in lines if l.strip().startswith(("core", "l2_cache", "noc_io", "dram"))]
    assert len(rows) == 11
    for r in rows:
        assert abs(float(r[3])) <= 0.0005 and abs(float(r[6])) <= 0.0005, r
    (out / "tmap_L3.txt").unlink()
    p = run(["compare3dice", case, "100", str(out)], tmp_path)
    assert p.returncode == 1 and "tmap_L3.txt: missing, or not 100 rows x 100 columns" in p.stdout
