"""Host-side product code on the CPU: the model builder and case reader of
libthermal_hip.so (no GPU calls) against the reference's fixtures, the case
errors the reference reports (case3d.h:86-155), and the synthetic workloads.
"""
import os

import numpy as np
import pytest

import oracle as O
import thermal_amd as T
from conftest import GOLDEN
from thermal_amd import workloads as W

CASES = os.path.join(GOLDEN, "cases")


def fx(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def check(a, f):
    for k in ("gE", "gN", "gU", "gB", "cap", "P", "tsv_frac", "dz"):
        assert np.array_equal(a[k], f[k]), k


def test_model_builder_matches_reference():
    m = T.Model.default(13, 9)
    m.add_uniform_power(0, 22.0)
    m.add_hotspot(0, 3e-3, 3e-3, 1e-3, 3.0)
    m.add_uniform_power(2, 25.0)
    check(m.finalize().arrays(), fx("model_hotspot_13x9"))
    check(T.Model.example_case(20, 14).arrays(), fx("model_example_20x14"))


def test_case_loader_matches_reference():
    f = fx("model_case_validate_tsv_30x25")
    m = T.Model.load_case(os.path.join(CASES, "validate_tsv"), 30, 25)
    check(m.arrays(), f)
    assert m.total_power() == float(f["total_power"])
    assert [l[0] for l in m.layers()][:2] == ["base", "Si tier 2 (bottom)"]
    assert m.settings() == dict(Lx=0.01, Ly=0.01, h_top=2e4, h_bot=0.0, T_amb=25.0)


def test_example_folder_equals_builtin_example():
    a = T.Model.load_case(os.path.join(CASES, "example"), 24, 18).arrays()
    b = T.Model.example_case(24, 18).arrays()
    for k in ("gE", "gN", "gU", "gB", "cap"):
        assert np.array_equal(a[k], b[k]), k
    assert np.allclose(a["P"], b["P"], rtol=1e-12, atol=1e-18)


def write(folder, name, text):
    with open(os.path.join(folder, name), "w") as f:
        f.write(text)


STACK = ("name, thickness_um, k, rho_c, planes, active\n"
         "Si, 50, 130, 1.66e6, 2, 1\nbond, 10, 2, 1.7e6, 1, 0\n")


@pytest.mark.parametrize("stack,floor,tsv,msg", [
    (None, None, None, "cannot open"),
    ("name\nSi, 50, 130\n", "n\n", None, "stack.csv: expected 6 columns"),
    ("name\n", "n\n", None, "stack.csv has no layers"),
    (STACK, None, None, "cannot open"),
    (STACK, "h\nb, Si, 1, 1, 1\n", None, "floorplan.csv: expected 7 columns"),
    (STACK, "h\nb, Cu, 1, 1, 1, 1, 1\n", None, "unknown layer 'Cu'"),
    (STACK, "h\nb, bond, 1, 1, 1, 1, 1\n", None, "layer 'bond' is not active"),
    (STACK, "h\nb, Si, 9.5, 1, 1, 1, 1\n", None, "block 'b' extends outside the die"),
    (STACK, "h\nb, Si, x1, 1, 1, 1, 1\n", None, "could not parse a number"),
    (STACK, "h\n", "h\nbond, 1, 1, 1\n", "tsv.csv: expected 6 columns"),
    (STACK, "h\n", "h\nmetal, 1, 1, 1, 1, 0.5\n", "tsv.csv: unknown layer 'metal'"),
    (STACK, "h\n", "h\nbond, 1, 1, 1, 1, 1.5\n", "cu_fraction must be between 0 and 1"),
])
def test_case_errors(tmp_path, stack, floor, tsv, msg):
    for name, text in (("stack.csv", stack), ("floorplan.csv", floor), ("tsv.csv", tsv)):
        if text is not None:
            write(tmp_path, name, text)
    with pytest.raises(T.ThermalError) as e:
        T.Model.load_case(str(tmp_path), 8)
    assert e.value.code == 4 and msg in str(e.value)


def test_case_settings_lines(tmp_path):
    write(tmp_path, "stack.csv", "# die_mm, 8, 6\n# h_top, 1e4\n# h_bottom, 50\n# ambient_C, 40\n" + STACK)
    write(tmp_path, "floorplan.csv", "h\n\"b\", \"Si\", 1, 1, 2, 2, 3\n")
    m = T.Model.load_case(str(tmp_path), 8)
    assert m.settings() == dict(Lx=8e-3, Ly=6e-3, h_top=1e4, h_bot=50.0, T_amb=40.0)
    assert np.isclose(m.total_power(), 3.0)
    assert m.get("gB").min() > 0  # h_bottom > 0: the bottom is cooled


def test_model_errors():
    with pytest.raises(T.ThermalError) as e:
        T.Model.from_layers([("thick", 1e-3, 1.0, 1e6, 65, 0)], 4)
    assert e.value.code == 5
    m = T.Model.default(8)
    with pytest.raises(T.ThermalError) as e:
        m.add_hotspot(0, 20e-3, 20e-3, 1e-4, 1.0)
    assert e.value.code == 6
    with pytest.raises(T.ThermalError) as e:
        m.desc()
    assert e.value.code == 3
    with pytest.raises(T.ThermalError) as e:
        m.add_uniform_power(9, 1.0)
    assert e.value.code == 1


def test_splitmix64_known_answers():
    """Published test vector of splitmix64 (seed 1234567)."""
    r = W.SplitMix64(1234567)
    assert [r.next() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                           4593380528125082431, 16408922859458223821]


def test_config2_spec():
    spec = W.config2_spec(1000)
    for d, layer in ((0, 0), (1, 2)):
        blocks = [b for b in spec["blocks"] if b[1] == layer]
        assert len(blocks) == 12
        assert np.isclose(sum(b[6] for b in blocks), 22.5)
        for b in blocks:
            assert 0.5e-3 <= b[4] <= 3e-3 and 0.5e-3 <= b[5] <= 3e-3 and b[2] + b[4] <= 10e-3
        for p in range(12):
            for q in range(p):
                a, c = blocks[p], blocks[q]
                assert not (a[2] < c[2] + c[4] and c[2] < a[2] + a[4] and a[3] < c[3] + c[5] and c[3] < a[3] + a[5])
    m = W.build(spec, 64)
    assert m.nz == 16 and np.isclose(m.total_power(), 50.0, rtol=1e-12)
    frac = m.get("tsv_frac").reshape(16, -1)
    # die 0's Si: 205 of 4096 lattice squares; the bond layer takes both dies'
    # clusters, which may coincide (the copper fraction is capped at 1)
    assert np.isclose(frac[0].mean(), 205 / 4096) and 0.09 < frac[4].mean() <= 410 / 4096


def test_written_case_round_trips(tmp_path):
    spec = W.config2_spec(7)
    folder = W.write_case(spec, str(tmp_path / "c"))
    a = T.Model.load_case(folder, 48).arrays()
    b = W.build(spec, 48).arrays()
    for k in ("gE", "gN", "gU", "cap"):
        assert np.allclose(a[k], b[k], rtol=1e-6), k
    assert np.allclose(a["P"], b["P"], rtol=1e-9, atol=1e-15)
    if O.ref_available():  # the reference's loader reads the identical inputs
        r = O.RefModel.case(folder, 48).arrays()
        for k in ("gE", "gN", "gU", "gB", "cap", "P"):
            assert np.array_equal(a[k], r[k]), k
