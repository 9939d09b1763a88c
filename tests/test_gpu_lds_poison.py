"""Steady solves whose whole V-cycle runs in the LDS-resident tail, started
with NaN left in every CU's LDS.

The tail copies its top level's x back to global memory, and on a grid with
odd nx that level has padding slots (the half-row's last slot in the rows
whose colour has no column there) that no sweep writes in LDS. They must
come back as 0: the PCG's r.z runs over every slot, and r = 0 there times a
NaN would make the whole solve NaN. The tail once copied those slots from
LDS unwritten, and a solve on a 61 x 45 x 1 grid failed to converge once in
many runs, whenever an earlier workgroup had left an Inf or NaN bit pattern
there. Filling LDS with NaN before the solve makes any such read show every
time; a probe kernel first checks that what one kernel leaves in LDS is what
the next kernel finds (else the test is skipped).
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
import thermal_amd as T
from test_gpu_plane_counts import stack_model

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BLOCKS, BYTES = 2048, 160 * 1024


@pytest.fixture(scope="module")
def poison():
    path = os.path.join(HERE, "liblds_poison.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with make -C tests")
    lib = ctypes.CDLL(path)
    lib.lds_poison_probe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    assert lib.lds_poison_fill(BLOCKS, BYTES) == 0
    frac = ctypes.c_double()
    assert lib.lds_poison_probe(BLOCKS, BYTES, ctypes.byref(frac)) == 0
    print(f"LDS words still NaN in the next kernel: {frac.value:.3f}")
    if frac.value < 0.9:
        pytest.skip(f"LDS contents do not persist between kernels here ({frac.value:.2f})")
    return lib


@pytest.mark.parametrize("nx,ny,nz,precision", [(61, 45, 1, "mixed"), (41, 29, 1, "mixed"), (41, 29, 1, "fp64"),
                                                 (25, 19, 3, "mixed"), (25, 19, 3, "fp64")])
def test_lds_tail_solve_with_poisoned_lds(poison, nx, ny, nz, precision):
    m = stack_model(nz, nx, ny)
    a = m.arrays()
    ref = O.steady(a, tol=1e-10, mg_double=(precision == "fp64"), flexible=True, fast_restrict=True)
    s = T.Solver(m, precision=precision, true_residual=True)
    assert s.nlevels() >= 5
    for rep in range(3):
        s.set_x_zero()
        s.set_rhs_steady()
        assert poison.lds_poison_fill(BLOCKS, BYTES) == 0
        st = s.solve(1e-10, 200)
        x = s.get_x()
        assert st["converged"] and abs(st["iters"] - ref["iters"]) <= 1, (rep, st, ref["iters"])
        assert np.isfinite(x).all()
        assert np.abs(x - ref["T"]).max() < 1e-8 * ref["T"].max()
    s.close()
