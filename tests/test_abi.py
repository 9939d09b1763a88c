"""The C ABI on the CPU: the library loads, exports every symbol that
include/thermal_hip.h declares (and the binding declares each with a
signature), and rejects bad arguments without touching the GPU."""
import ctypes as C
import os
import re

import pytest

import thermal_amd as T
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "thermal_hip.h")


def declared():
    text = open(HEADER).read()
    return re.findall(r"TH_API\s+[\w\s\*]+?\b(th_\w+)\s*\(", text)


def test_header_declares_the_boundary():
    names = declared()
    for must in ("th_solver_create", "th_solver_set_dt", "th_solver_set_x_zero", "th_solver_set_rhs_steady",
                 "th_solver_set_rhs_transient", "th_solver_solve", "th_solver_get_x", "th_solver_set_power",
                 "th_solver_device_x", "th_solver_destroy", "th_run_transient", "th_model_load_case"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(T.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
        assert name in T.SIGNATURES, name


def test_errors_without_gpu():
    L = T.lib()
    assert b"gfx950" in L.th_version()
    h = C.c_void_p()
    assert L.th_solver_create(None, None, C.byref(h)) == 1
    assert b"null model descriptor" in L.th_last_error()
    d = T.ModelDesc()
    d.nx, d.ny, d.nz = 4, 4, 70
    assert L.th_solver_create(C.byref(d), None, C.byref(h)) == 1
    d.nz = 4
    assert L.th_solver_create(C.byref(d), None, C.byref(h)) == 3  # arrays missing: not finalized
    assert L.th_solver_solve(None, 1e-6, 10, None) == 1
    L.th_solver_destroy(None)  # never fails
    with pytest.raises(T.ThermalError):
        T.Model.load_case("/nonexistent/case", 8)


def test_cli_reports_a_bad_case_folder(tmp_path):
    """thermal3d exits 1 with the loader's message before touching the GPU."""
    import subprocess

    cli = os.path.join(ROOT, "cuda-thermal_amd", "bin", "thermal3d")
    p = subprocess.run([cli, "case", str(tmp_path / "missing"), "16"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "Could not load case" in p.stdout and "stack.csv" in p.stdout
    p = subprocess.run([cli, "export3dice", str(tmp_path / "missing"), "100", str(tmp_path)], capture_output=True,
                       text=True, timeout=60)
    assert p.returncode == 1 and p.stdout.startswith("Export failed: cannot open") and "stack.csv" in p.stdout


def test_case_blocks_exported():
    m = T.Model.load_case(os.path.join(ROOT, "tests", "golden", "cases", "example"), 16)
    L = T.lib()
    assert L.th_model_block_count(m.h) == 8 and L.th_model_tsv_count(m.h) == 4
    name = C.c_char_p()
    layer = C.c_int()
    v = [C.c_double() for _ in range(5)]
    assert L.th_model_block_info(m.h, 0, C.byref(name), C.byref(layer), *[C.byref(x) for x in v]) == 0
    assert name.value and 0 <= layer.value < 5 and v[4].value > 0
    assert L.th_model_block_info(m.h, 99, None, None, None, None, None, None, None) == 1
