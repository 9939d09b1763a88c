"""The die map (th_solver_colorize) against the oracle of the reference
viewer's k_colorize and colormap (viewer.cu:53-87): every RGBA byte equal,
the hottest cell and its rise equal, on solved fields of odd and even grids,
for colour ranges below, at and above the hottest rise."""
import ctypes as C

import numpy as np
import pytest

import thermal_amd as T
import viewer as V

pytestmark = pytest.mark.gpu


def solved(m, precision="mixed"):
    s = T.Solver(m, precision=precision)
    s.set_x_zero()
    s.set_rhs_steady()
    assert s.solve(1e-10, 200)["converged"]
    return s


@pytest.mark.parametrize("n,planes", [((37, 23), (0, 1)), ((130, 97), (2, 3)), ((256, 256), (0, 9))])
@pytest.mark.parametrize("frac", [0.5, 1.0, 2.0])
def test_colorize_matches_oracle(n, planes, frac):
    m = T.Model.example_case(*n)
    s = solved(m)
    x = s.get_x()
    l0, l1 = planes
    vmax = float(x.reshape(m.nz, m.ny, m.nx)[l0:l1 + 1].max())
    rise_hi = np.float32(frac * vmax)
    img, hot, (i, j) = s.colorize(l0, l1, rise_hi)
    oimg, ohot, ocell = V.colorize(x, m.nx, m.ny, m.nz, l0, l1, rise_hi)
    assert np.array_equal(img, oimg), int((img != oimg).any(axis=2).sum())
    assert hot == ohot and j * m.nx + i == ocell
    _, hot2, ij2 = s.colorize(l0, l1, rise_hi, image=False)
    assert (hot2, ij2) == (hot, (i, j))


def test_colorize_into_device_memory():
    m = T.Model.example_case(64, 48)
    s = solved(m, "fp64")
    img, hot, (i, j) = s.colorize(0, m.nz - 1, 40.0)
    hip = C.CDLL("libamdhip64.so.7")
    p = C.c_void_p()
    nbytes = m.nx * m.ny * 4
    assert hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
    try:
        v, cell = C.c_float(), C.c_int()
        assert T.lib().th_solver_colorize_device(s.h, 0, m.nz - 1, 40.0, p, C.byref(v), C.byref(cell)) == 0
        back = np.empty_like(img)
        assert hip.hipMemcpy(back.ctypes.data_as(C.c_void_p), p, C.c_size_t(nbytes), 2) == 0  # device to host
    finally:
        hip.hipFree(p)
    assert np.array_equal(back, img) and v.value == hot and cell.value == j * m.nx + i


def test_colorize_argument_errors():
    m = T.Model.example_case(16, 12)
    s = solved(m)
    for l0, l1, hi in [(0, m.nz, 10.0), (3, 2, 10.0), (-1, 2, 10.0), (0, 1, 0.0), (0, 1, -1.0)]:
        with pytest.raises(T.ThermalError):
            s.colorize(l0, l1, hi)
