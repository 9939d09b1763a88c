"""The benchmark's TSV stack (seeded floorplan of 12 blocks per die on a 10%
background, 205 TSV clusters per die in the Si and bond layers, 16 z-planes)
through the HIP solver, against the FP64 oracle.

Two oracles per case:
* the oracle with the solver's own options (flexible beta, lateral
  restriction, V-cycle in the solver's precision): the same arithmetic, so
  the iteration count within one and the field within 1e-8 relative;
* the reference's plain FP64 algorithm (FP64 V-cycle, Fletcher-Reeves beta,
  residual-form restriction, pcg3d_cpu.cpp:115-185 with -DMG_DOUBLE): at
  most 0.0005 C max-abs apart.
Every result also meets a true relative residual <= 1e-10 (flux form) and
energy balance (heat out = power in) to 1e-6.

Sizes: 128^2 (plane-parallel fine level), 256^2 (column walk with column
types: the clusters are whole cells), 300 x 212 (clusters cover cells in
part: the per-cell column walk), and the full 1024^2 x 16.
"""
import numpy as np
import pytest

import oracle as O
import thermal_amd as T
from thermal_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10
BAR_K = 5e-4


def build(nx, ny=None):
    m = W.build(W.config2_spec(1000), nx, ny)
    assert m.nz == 16
    return m


def gpu_solve(m, precision):
    s = T.Solver(m, precision=precision, beta="fr" if precision == "fp64" else "flexible", true_residual=True)
    s.set_x_zero()
    s.set_rhs_steady()
    st = s.solve(TOL, 200)
    x = s.get_x()
    info = dict(st, true_relres=s.true_relres(), heat_out=s.heat_out(), types=s.column_types())
    s.close()
    return x, info


def check_result(m, a, x, info, ref_plain):
    P = m.total_power()
    assert info["converged"], info
    assert info["true_relres"] <= TOL, info
    assert O.true_relres_flux(a, x) <= 1.05 * TOL
    assert abs(O.heat_out(a, x) - P) / P < 1e-6
    assert abs(info["heat_out"] - P) / P < 1e-6
    err = float(np.abs(x - ref_plain).max())
    assert err <= BAR_K, err
    return err


@pytest.mark.parametrize("nx,ny,typed", [(128, 128, None), (256, 256, True), (300, 212, False)])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_config2_vs_oracle(nx, ny, typed, precision):
    m = build(nx, ny)
    a = m.arrays()
    x, info = gpu_solve(m, precision)
    if typed is not None:
        assert (info["types"] > 0) == typed, info["types"]
    same = O.steady(a, tol=TOL, mg_double=(precision == "fp64"), flexible=(precision != "fp64"), fast_restrict=True)
    assert abs(info["iters"] - same["iters"]) <= 1, (info, same["iters"])
    assert np.abs(x - same["T"]).max() < 1e-8 * same["T"].max()
    plain = O.steady(a, tol=TOL, mg_double=True)
    err = check_result(m, a, x, info, plain["T"])
    print(f"config2 {nx}x{ny}x16 {precision}: {info['iters']} iterations (oracle {same['iters']}), "
          f"max |GPU - FP64 oracle| = {err:.2e} K")


@pytest.fixture(scope="module")
def full_size():
    """The 1024^2 x 16 model and the reference algorithm's FP64 answer (about
    15 s of one CPU core)."""
    m = build(1024)
    a = m.arrays()
    plain = O.steady(a, tol=TOL, mg_double=True)
    assert plain["converged"]
    return m, a, plain["T"]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_config2_full_size(full_size, precision):
    m, a, ref = full_size
    x, info = gpu_solve(m, precision)
    assert info["types"] > 0
    err = check_result(m, a, x, info, ref)
    print(f"config2 1024x1024x16 {precision}: {info['iters']} iterations, true rel-res {info['true_relres']:.2e}, "
          f"max |GPU - FP64 oracle| = {err:.2e} K")


def test_config2_transient_vs_oracle():
    """Backward-Euler steps of the TSV stack at 128^2 x 16 with the
    benchmark's dt rule (10 us growing 10% per step), 30 steps: t and dt
    bit-identical to the oracle's, the iteration count of every step within
    one, the final field within 1e-7 relative; per-step energy balance
    1e-6."""
    m = build(128)
    a = m.arrays()
    kw = dict(dt0=1e-5, growth=1.1, dt_max=1e30, t_end=1e30, max_steps=30, tol=1e-9)
    ref = O.run_transient(a, flexible=True, **kw)
    s = T.Solver(m)
    g = T.run_transient(s, **kw)
    s.close()
    assert g["converged"] and len(g["t"]) == 30
    assert np.array_equal(g["t"], ref["t"]) and np.array_equal(g["dt"], ref["dt"])
    assert np.abs(g["iters"] - ref["iters"]).max() <= 1
    assert np.abs(g["T"] - ref["T"]).max() < 1e-7 * ref["T"].max()
    P = m.total_power()
    assert (np.abs(g["q_out"] + g["q_stored"] - P) / P).max() < 1e-6
