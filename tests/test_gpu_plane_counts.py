"""Every plane-count instantiation of the column kernels against the oracle.

The other solver tests use 10 or 16 z-planes. Stacks of 1, 3, 6, 20, 40 and
64 planes reach the rest: the 4- and 8-lane plane-parallel sweeps and LDS
tail, the NZ = 4/8/24/32/48/64 column walks and, above 16 planes, the
L2-resident tail. Bars as in test_gpu_solver: one V-cycle bit-identical to
the oracle's (same operation order, no FMA contraction on either side), a
steady solve within one iteration of the oracle's count and 1e-8 of its
field, true relative residual <= 1e-10.
"""
import numpy as np
import pytest

import oracle as O
import thermal_amd as T

pytestmark = pytest.mark.gpu

SI = ("Si", 50e-6, 130.0, 1.66e6)
BOND = ("bond", 10e-6, 2.0, 1.70e6)
TIM = ("TIM", 25e-6, 5.0, 2.00e6)
CU = ("Cu", 1e-3, 400.0, 3.45e6)

# planes per layer of a Si / bond / Si / TIM / Cu stack (or a single Si die)
STACKS = {
    1: [SI + (1, 1)],
    3: [SI + (2, 1), CU + (1, 0)],
    6: [SI + (1, 1), BOND + (1, 0), SI + (1, 1), TIM + (1, 0), CU + (2, 0)],
    20: [SI + (6, 1), BOND + (2, 0), SI + (6, 1), TIM + (2, 0), CU + (4, 0)],
    40: [SI + (12, 1), BOND + (4, 0), SI + (12, 1), TIM + (4, 0), CU + (8, 0)],
    64: [SI + (20, 1), BOND + (6, 0), SI + (20, 1), TIM + (6, 0), CU + (12, 0)],
}


def stack_model(nz, nx, ny):
    layers = STACKS[nz]
    m = T.Model.from_layers(layers, nx, ny)
    active = [k for k, l in enumerate(layers) if l[5]]
    m.add_uniform_power(active[0], 20.0)
    m.add_hotspot(active[0], 3e-3, 4e-3, 1e-3, 3.0)
    if len(active) > 1:
        m.add_uniform_power(active[1], 15.0)
    m.add_tsv_region(active[0], 6e-3, 2e-3, 2e-3, 2e-3, 0.3)
    m.finalize()
    assert m.nz == nz
    return m


@pytest.mark.parametrize("nz", sorted(STACKS))
@pytest.mark.parametrize("precision,mg_double", [("mixed", False), ("fp64", True)])
@pytest.mark.parametrize("restriction", ["lateral", "residual"])
def test_vcycle_bitwise(nz, precision, mg_double, restriction):
    """61 x 45: the fine level runs as separate plane-parallel (nz <= 16) or
    column-walk (nz > 16) kernels, the coarse levels in a tail kernel."""
    m = stack_model(nz, 61, 45)
    s = T.Solver(m, precision=precision, restriction=restriction)
    o = O.Solver(m.arrays(), "mg", mg_double=mg_double, fast_restrict=(restriction == "lateral"))
    r = np.random.default_rng(nz).standard_normal(m.n)
    zg, zo = s.precondition(r), o.precond(r)
    assert np.array_equal(zg, zo), np.abs(zg - zo).max()


@pytest.mark.parametrize("nz", [1, 3, 6])
def test_vcycle_bitwise_column_walk(nz):
    """192 x 96: more than 8192 columns per colour, so the fine level runs
    the column-walk sweep instantiated for 4 and 8 planes."""
    m = stack_model(nz, 192, 96)
    s = T.Solver(m, precision="mixed")
    o = O.Solver(m.arrays(), "mg", mg_double=False, fast_restrict=True)
    r = np.random.default_rng(100 + nz).standard_normal(m.n)
    assert np.array_equal(s.precondition(r), o.precond(r))


@pytest.mark.parametrize("nz", sorted(STACKS))
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_steady_vs_oracle(nz, precision):
    m = stack_model(nz, 61, 45)
    a = m.arrays()
    ref = O.steady(a, tol=1e-10, mg_double=(precision == "fp64"), flexible=True, fast_restrict=True)
    s = T.Solver(m, precision=precision, true_residual=True)
    s.set_x_zero()
    s.set_rhs_steady()
    st = s.solve(1e-10, 200)
    x = s.get_x()
    assert st["converged"] and abs(st["iters"] - ref["iters"]) <= 1, (st, ref["iters"], O.true_relres(a, x))
    assert np.abs(x - ref["T"]).max() < 1e-8 * ref["T"].max()
    assert O.true_relres(a, x) <= 1.5e-10
    P = m.total_power()
    assert abs(O.heat_out(a, x) - P) / P < 1e-6


@pytest.mark.parametrize("nz", [1, 3, 6, 16])
@pytest.mark.parametrize("precision,mg_double", [("mixed", False), ("fp64", True)])
@pytest.mark.parametrize("n", [(301, 171), (522, 266)])
def test_vcycle_bitwise_large_levels(nz, precision, mg_double, n):
    """Levels with more than 8192 columns per colour run the column-walk
    sweeps: (301, 171) has one such level, (522, 266) two, with odd coarse
    sizes below them, steady and transient operators."""
    layers = STACKS[nz] if nz != 16 else [SI + (4, 1), BOND + (2, 0), SI + (4, 1), TIM + (2, 0), CU + (4, 0)]
    m = T.Model.from_layers(layers, *n)
    active = [k for k, l in enumerate(layers) if l[5]]
    m.add_uniform_power(active[0], 20.0)
    m.add_hotspot(active[0], 3e-3, 4e-3, 1e-3, 3.0)
    m.add_tsv_region(active[0], 6e-3, 2e-3, 2e-3, 2e-3, 0.3)
    m.finalize()
    s = T.Solver(m, precision=precision)
    assert s.column_types() > 0  # the fine level reads column types
    o = O.Solver(m.arrays(), "mg", mg_double=mg_double, fast_restrict=True)
    for dt in (0.0, 1e-3):
        s.set_dt(dt)
        o.set_dt(dt)
        r = np.random.default_rng(7 + nz).standard_normal(m.n)
        zg, zo = s.precondition(r), o.precond(r)
        assert np.array_equal(zg, zo), np.abs(zg - zo).max()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_steady_large_levels_vs_oracle(precision):
    m = T.Model.example_case(300, 212)
    a = m.arrays()
    ref = O.steady(a, tol=1e-10, mg_double=(precision == "fp64"), flexible=True, fast_restrict=True)
    s = T.Solver(m, precision=precision, true_residual=True)
    s.set_x_zero()
    s.set_rhs_steady()
    st = s.solve(1e-10, 200)
    x = s.get_x()
    assert st["converged"] and abs(st["iters"] - ref["iters"]) <= 1, (st, ref["iters"], O.true_relres(a, x))
    assert np.abs(x - ref["T"]).max() < 1e-8 * ref["T"].max()
    assert O.true_relres(a, x) <= 1.5e-10


@pytest.mark.parametrize("nz", [1, 6, 16])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("dt", [0.0, 1e-3])
def test_operator_bitwise_large_levels(nz, precision, dt):
    """The fine-level apply on a grid large enough for column types (the
    faces read from a per-type table): bit-identical to the oracle's."""
    layers = STACKS[nz] if nz != 16 else [SI + (4, 1), BOND + (2, 0), SI + (4, 1), TIM + (2, 0), CU + (4, 0)]
    m = T.Model.from_layers(layers, 301, 171)
    m.add_uniform_power(0, 20.0)
    m.add_tsv_region(0, 6e-3, 2e-3, 2e-3, 2e-3, 0.3)
    m.finalize()
    s = T.Solver(m, precision=precision)
    assert s.column_types() > 0
    s.set_dt(dt)
    o = O.Solver(m.arrays(), "mg")
    o.set_dt(dt)
    x = np.random.default_rng(nz).standard_normal(m.n)
    assert np.array_equal(s.apply(x), o.apply(x))


@pytest.mark.parametrize("nz", [1, 6, 16])
@pytest.mark.parametrize("x0", ["zero", "random"])
def test_true_residual_large_levels(nz, x0):
    """||P - A x|| / ||P|| on a grid with column types (the typed column-walk
    residual, flux form) against the oracle's flux-form residual, for x = 0
    and for a random x; then a steady solve on the same grid reaches a true
    relative residual <= 1e-10 through the typed start, replacement and
    check residuals."""
    layers = STACKS[nz] if nz != 16 else [SI + (4, 1), BOND + (2, 0), SI + (4, 1), TIM + (2, 0), CU + (4, 0)]
    m = T.Model.from_layers(layers, 301, 171)
    m.add_uniform_power(0, 20.0)
    m.add_hotspot(0, 3e-3, 4e-3, 1e-3, 3.0)
    m.add_tsv_region(0, 6e-3, 2e-3, 2e-3, 2e-3, 0.3)
    m.finalize()
    a = m.arrays()
    s = T.Solver(m, precision="mixed", true_residual=True)
    assert s.column_types() > 0
    x = np.zeros(m.n) if x0 == "zero" else 40.0 + np.random.default_rng(nz).standard_normal(m.n)
    s.set_x(x)
    tr, ref = s.true_relres(), O.true_relres_flux(a, x)
    # the norms are summed in another order than the oracle's (1e-12 of a
    # residual norm of ~3e3 for the random x); a wrong face would be O(1)
    assert abs(tr - ref) <= 1e-10 * ref, (tr, ref)
    s.set_x_zero()
    s.set_rhs_steady()
    st = s.solve(1e-10, 200)
    assert st["converged"]
    xs = s.get_x()
    assert O.true_relres_flux(a, xs) <= 1.5e-10
    assert abs(s.true_relres() - O.true_relres_flux(a, xs)) <= 1e-6 * s.true_relres()


@pytest.mark.parametrize("nz", [6, 16])
def test_transient_large_levels_vs_oracle(nz):
    """Backward-Euler steps on a grid with column types (the typed apply,
    sweeps and residual with their C/dt terms, the start residual from the
    previous step's x) against the oracle, step by step."""
    layers = STACKS[nz] if nz != 16 else [SI + (4, 1), BOND + (2, 0), SI + (4, 1), TIM + (2, 0), CU + (4, 0)]
    m = T.Model.from_layers(layers, 301, 171)
    m.add_uniform_power(0, 20.0)
    m.add_hotspot(0, 3e-3, 4e-3, 1e-3, 3.0)
    m.add_tsv_region(0, 6e-3, 2e-3, 2e-3, 2e-3, 0.3)
    m.finalize()
    a = m.arrays()
    kw = dict(dt0=1e-5, growth=1.5, dt_max=1e30, t_end=1e30, max_steps=8, tol=1e-9)
    ref = O.run_transient(a, flexible=True, **kw)
    s = T.Solver(m)
    assert s.column_types() > 0
    g = T.run_transient(s, **kw)
    assert g["converged"] and len(g["t"]) == 8
    assert np.array_equal(g["t"], ref["t"]) and np.array_equal(g["dt"], ref["dt"])
    assert np.abs(g["iters"] - ref["iters"]).max() <= 1
    assert np.abs(g["T"] - ref["T"]).max() < 1e-7 * ref["T"].max()
    P = m.total_power()
    eb = np.abs(g["q_out"] + g["q_stored"] - P) / P
    assert eb.max() < 1e-6
