"""th_solver_colorize_device_noalloc (the die map with its scratch in a static
device array) against th_solver_colorize_device (scratch allocated by the
library), through the C ABI: the same RGBA bytes and the same hottest cell,
on two calls in a row on one GPU, the second on a smaller grid; without an
image, the same hottest cell as th_solver_colorize; and a grid beyond the
static scratch is refused."""
import ctypes as C

import numpy as np
import pytest

import thermal_amd as T

pytestmark = pytest.mark.gpu

HIP = C.CDLL("libamdhip64.so.7")


def solved(nx, ny):
    m = T.Model.default(nx, ny)
    m.add_uniform_power(0, 22.0)
    m.add_hotspot(0, 3e-3, 4e-3, 1e-3, 3.0)
    m.add_uniform_power(2, 25.0)
    m.finalize()
    s = T.Solver(m)
    s.set_x_zero()
    s.set_rhs_steady()
    assert s.solve(1e-10, 200)["converged"]
    return m, s


def image_bytes(buf, nx, ny):
    out = np.empty(nx * ny * 4, np.uint8)
    assert HIP.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(buf.ptr), C.c_size_t(out.nbytes), 2) == 0
    return out


def die_map(entry, s, nx, ny, l0, l1, rise_hi, image):
    buf = T.DeviceBuffer(np.full(nx * ny, np.nan, np.float32)) if image else None
    v, cell = C.c_float(), C.c_int()
    rc = getattr(T.lib(), entry)(s.h, l0, l1, C.c_float(rise_hi), C.c_void_p(buf.ptr) if image else None,
                                 C.byref(v), C.byref(cell))
    assert rc == T.TH_OK, T.lib().th_last_error()
    return (image_bytes(buf, nx, ny) if image else None), v.value, cell.value


@pytest.mark.parametrize("image", [True, False])
def test_noalloc_matches_allocating_entry(image):
    results = []
    for nx, ny, l0, l1 in ((512, 384, 0, 1), (96, 72, 2, 3)):  # the second call on a smaller grid
        m, s = solved(nx, ny)
        rise_hi = float(s.get_x().max()) * 1.1
        # without an image, the allocating counterpart is th_solver_colorize
        # (th_solver_colorize_device refuses a null image)
        ref = die_map("th_solver_colorize_device" if image else "th_solver_colorize", s, nx, ny, l0, l1, rise_hi,
                      image)
        got = die_map("th_solver_colorize_device_noalloc", s, nx, ny, l0, l1, rise_hi, image)
        assert got[1] == ref[1] and got[2] == ref[2], (got[1:], ref[1:])
        if image:
            assert np.array_equal(got[0], ref[0])
            assert (got[0].reshape(-1, 4)[:, 3] == 255).all()
        # the hottest cell is the hottest column of planes l0..l1
        x = s.get_x().reshape(m.nz, ny, nx)[l0:l1 + 1].max(axis=0)
        assert got[1] == np.float32(max(0.0, x.max()))
        results.append(got)
        s.close()
    assert results[0][2] != results[1][2] or results[0][1] != results[1][1]


def test_noalloc_refuses_grids_beyond_the_scratch():
    m = T.Model.from_layers([("Si", 50e-6, 130.0, 1.66e6, 1, 1)], 2050, 2048)
    m.add_uniform_power(0, 10.0)
    m.finalize()
    s = T.Solver(m)
    v, cell = C.c_float(), C.c_int()
    rc = T.lib().th_solver_colorize_device_noalloc(s.h, 0, 0, C.c_float(1.0), None, C.byref(v), C.byref(cell))
    assert rc == 1
    assert b"exceeds the static scratch" in T.lib().th_last_error()
    s.close()
