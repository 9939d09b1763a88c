"""GPU parity tests: the HIP solver through the C ABI against the oracle (the
CPU restatement of the reference's pcg3d_cpu path) on the same inputs.

Bars (the reference's own checks, main3d.cu:94-282, tightened): solution
within 5e-4 C max-abs of the FP64 oracle;
true relative residual <= 1e-10; energy balance to 1e-6; the exact 1D
resistor chain under uniform power; FP64 operator apply and the FP32 / FP64
V-cycle bit-identical to the oracle's (same operation order, no FMA
contraction on either side).
"""
import numpy as np
import pytest

import oracle as O
import thermal_amd as T

pytestmark = pytest.mark.gpu

TOL_C = 5e-4  # max-abs difference vs the FP64 oracle (C)


def hotspot(n, ny=None, at=(3e-3, 3e-3)):
    m = T.Model.default(n, ny)
    m.add_uniform_power(0, 22.0)
    m.add_hotspot(0, at[0], at[1], 1e-3, 3.0)
    m.add_uniform_power(2, 25.0)
    return m.finalize()


@pytest.fixture(scope="module")
def example_odd():
    """The reference's example case (TSVs, off-grid blocks) on an odd grid."""
    return T.Model.example_case(37, 23)


def rng_vec(n, seed=0):
    return np.random.default_rng(seed).standard_normal(n)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("dt", [0.0, 1e-4])
def test_operator_bitwise(example_odd, precision, dt):
    a = example_odd.arrays()
    s = T.Solver(example_odd, precision=precision)
    s.set_dt(dt)
    o = O.Solver(a, "mg")
    o.set_dt(dt)
    x = rng_vec(example_odd.n, 1)
    assert np.array_equal(s.apply(x), o.apply(x))


@pytest.mark.parametrize("precision,mg_double", [("mixed", False), ("fp64", True)])
@pytest.mark.parametrize("dt", [0.0, 1e-3])
@pytest.mark.parametrize("restriction", ["lateral", "residual"])
@pytest.mark.parametrize("n", [(37, 23), (16, 12)])
def test_vcycle_bitwise(precision, mg_double, dt, restriction, n):
    """One V-cycle, bit for bit: (37, 23) runs level 0 as separate kernels,
    (16, 12) runs every level inside the single-workgroup tail."""
    m = T.Model.example_case(*n)
    a = m.arrays()
    s = T.Solver(m, precision=precision, restriction=restriction)
    s.set_dt(dt)
    o = O.Solver(a, "mg", mg_double=mg_double, fast_restrict=(restriction == "lateral"))
    o.set_dt(dt)
    r = rng_vec(m.n, 2)
    zg, zo = s.precondition(r), o.precond(r)
    assert np.array_equal(zg, zo), np.abs(zg - zo).max()


def test_jacobi_precond_matches(example_odd):
    a = example_odd.arrays()
    s = T.Solver(example_odd, precond="jacobi")
    o = O.Solver(a, "jacobi")
    r = rng_vec(example_odd.n, 3)
    assert np.array_equal(s.precondition(r), o.precond(r))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("beta", ["flexible", "fr"])
@pytest.mark.parametrize("restriction", ["lateral", "residual"])
def test_steady_vs_oracle(precision, beta, restriction):
    m = hotspot(64)
    a = m.arrays()
    ref = O.steady(a, tol=1e-10, mg_double=(precision == "fp64"), flexible=(beta == "flexible"),
                   fast_restrict=(restriction == "lateral"))
    s = T.Solver(m, precision=precision, beta=beta, restriction=restriction)
    s.set_x_zero()
    s.set_rhs_steady()
    st = s.solve(1e-10, 200)
    x = s.get_x()
    assert st["converged"]
    assert abs(st["iters"] - ref["iters"]) <= 1
    assert np.abs(x - ref["T"]).max() < TOL_C
    assert np.abs(x - ref["T"]).max() < 1e-8 * ref["T"].max()
    tr = s.true_relres()
    assert tr <= 1.5e-10
    assert abs(tr - O.true_relres(a, x)) <= 1e-3 * tr + 1e-15


@pytest.mark.parametrize("n", [(20, 12), (9, 7)])
def test_steady_small_grid_in_tail(n):
    """Grids whose every level runs in the single-workgroup tail, which then
    also carries the convergence test."""
    m = T.Model.example_case(*n)
    a = m.arrays()
    ref = O.steady(a, tol=1e-10, flexible=True, fast_restrict=True)
    r = T.steady_solve(m, tol=1e-10)
    assert r["converged"] and abs(r["iters"] - ref["iters"]) <= 1
    assert np.abs(r["T"] - ref["T"]).max() < 1e-8 * ref["T"].max()


def test_steady_odd_grid_with_tsvs(example_odd):
    a = example_odd.arrays()
    ref = O.steady(a, tol=1e-10, flexible=True)
    r = T.steady_solve(example_odd, tol=1e-10)
    assert r["converged"]
    assert np.abs(r["T"] - ref["T"]).max() < TOL_C
    assert O.true_relres(a, r["T"]) <= 1.5e-10
    P = example_odd.total_power()
    assert abs(O.heat_out(a, r["T"]) - P) / P < 1e-6


def test_uniform_power_is_1d_chain():
    m = T.Model.default(48)
    m.add_uniform_power(0, 25.0)
    m.add_uniform_power(2, 25.0)
    m.finalize()
    r = T.steady_solve(m, tol=1e-10)
    Tz = m.analytic_1d()
    err = np.abs(r["T"].reshape(m.nz, -1) - Tz[:, None]).max()
    assert err / Tz.max() < 1e-5
    P = m.total_power()
    assert abs(m.heat_out(r["T"]) - P) / P < 1e-6


def test_jacobi_steady_vs_oracle():
    m = hotspot(24)
    a = m.arrays()
    ref = O.steady(a, tol=1e-9, precond="jacobi")
    r = T.steady_solve(m, tol=1e-9, precond="jacobi")
    assert r["converged"]
    assert abs(r["iters"] - ref["iters"]) <= 2
    assert np.abs(r["T"] - ref["T"]).max() < 1e-6 * ref["T"].max()


def test_zero_power_and_maxit():
    m = T.Model.default(16).finalize()
    s = T.Solver(m)
    s.set_x(np.ones(m.n))
    s.set_rhs_steady()
    st = s.solve(1e-10, 100)
    assert st["converged"] and st["iters"] == 0
    assert not s.get_x().any()
    m2 = hotspot(32)
    s2 = T.Solver(m2)
    s2.set_x_zero()
    s2.set_rhs_steady()
    st2 = s2.solve(1e-12, 2)
    assert not st2["converged"] and st2["iters"] == 2


@pytest.mark.parametrize("k", [1, 2, 5])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_stopped_solve_returns_the_iterate(k, precision):
    """A solve stopped by maxit returns the oracle's k-th iterate: the last
    update's x += alpha p, which the multigrid path applies one kernel late,
    is in x when the solve returns."""
    m = hotspot(48)
    a = m.arrays()
    ref = O.steady(a, tol=1e-14, maxit=k, mg_double=(precision == "fp64"), flexible=True, fast_restrict=True)
    s = T.Solver(m, precision=precision, beta="flexible", restriction="lateral")
    s.set_x_zero()
    s.set_rhs_steady()
    st = s.solve(1e-14, k)
    assert st["iters"] == k and not st["converged"]
    x = s.get_x()
    err = np.abs(x - ref["T"]).max() / np.abs(ref["T"]).max()
    # FP64: only the dot products' summation order differs; mixed: FP32
    # search direction and V-cycle rounding
    assert err < (1e-12 if precision == "fp64" else 1e-5)
    tr = s.true_relres()
    assert abs(tr - O.true_relres(a, x)) <= 1e-6 * tr


@pytest.mark.parametrize("graphs", [True, False])
def test_transient_vs_oracle(graphs):
    m = hotspot(32)
    a = m.arrays()
    ref = O.run_transient(a, t_end=0.02, growth=1.2, tol=1e-9, flexible=True)
    s = T.Solver(m, graphs=graphs)
    g = T.run_transient(s, t_end=0.02, growth=1.2, tol=1e-9)
    assert g["converged"] and len(g["t"]) == len(ref["t"])
    assert np.array_equal(g["t"], ref["t"]) and np.array_equal(g["dt"], ref["dt"])
    assert np.abs(g["iters"] - ref["iters"]).max() <= 1
    assert np.abs(g["T"] - ref["T"]).max() < 1e-7 * ref["T"].max()
    assert np.allclose(g["T_max"], ref["T_max"], rtol=1e-7)
    P = m.total_power()
    eb = np.abs(g["q_out"] + g["q_stored"] - P) / P
    assert eb.max() < 1e-6


def test_transient_step_rule_vs_oracle():
    """The 100-step rule of the 800x800x10 transient (dt0 = 10 us growing 10%
    per step, no cap) on a 48x48 grid against the oracle, step by step."""
    m = hotspot(48, at=(5e-3, 5e-3))
    a = m.arrays()
    kw = dict(dt0=1e-5, growth=1.1, dt_max=1e30, t_end=1e30, max_steps=100, tol=1e-9)
    ref = O.run_transient(a, flexible=True, **kw)
    g = T.run_transient(T.Solver(m), **kw)
    assert g["converged"] and len(g["t"]) == 100
    assert np.array_equal(g["t"], ref["t"]) and np.array_equal(g["dt"], ref["dt"])
    assert np.isclose(g["t"][-1], 1e-5 * (1.1 ** 100 - 1) / 0.1, rtol=1e-12)
    assert np.abs(g["iters"] - ref["iters"]).max() <= 1
    assert np.abs(g["T"] - ref["T"]).max() < 1e-7 * ref["T"].max()
    assert np.allclose(g["T_max"], ref["T_max"], rtol=1e-7)
    assert np.allclose(g["q_out"], ref["q_out"], rtol=1e-7, atol=1e-9 * m.total_power())


@pytest.mark.parametrize("graphs", [True, False])
def test_transient_predictor(graphs):
    """Steps started from the extrapolation of the last two reach the same
    answers (to the solve tolerance) in fewer iterations."""
    m = hotspot(48, at=(5e-3, 5e-3))
    a = m.arrays()
    kw = dict(dt0=1e-5, growth=1.1, dt_max=1e30, t_end=1e30, max_steps=100, tol=1e-9)
    ref = O.run_transient(a, flexible=True, **kw)
    plain = T.run_transient(T.Solver(m, graphs=graphs), **kw)
    pred = T.run_transient(T.Solver(m, graphs=graphs), predictor=True, **kw)
    assert pred["converged"] and np.array_equal(pred["t"], ref["t"])
    assert np.abs(pred["T"] - ref["T"]).max() < 1e-7 * ref["T"].max()
    assert np.allclose(pred["T_max"], ref["T_max"], rtol=1e-7)
    P = m.total_power()
    assert (np.abs(pred["q_out"] + pred["q_stored"] - P) / P).max() < 1e-6
    assert pred["iters"].sum() < 0.95 * plain["iters"].sum(), (pred["iters"].sum(), plain["iters"].sum())


def test_transient_800x800x10_100_steps():
    """The 800x800x10 transient from ambient, 100 backward-Euler steps:
    per-step energy balance, mixed vs FP64 steps, and the final field
    against the steady state (the heat-sink mode, RC ~ 0.18 s, leaves a few
    1e-4 of the rise after 1.38 s)."""
    m = hotspot(800, at=(5e-3, 5e-3))
    P = m.total_power()
    kw = dict(dt0=1e-5, growth=1.1, dt_max=1e30, t_end=1e30, max_steps=100, tol=1e-9)
    res = {}
    for prec in ("fp64", "mixed"):
        g = T.run_transient(T.Solver(m, precision=prec), **kw)
        assert g["converged"] and len(g["t"]) == 100
        eb = np.abs(g["q_out"] + g["q_stored"] - P) / P
        assert eb.max() < 1e-6, (prec, eb.max())
        res[prec] = g
    assert np.abs(res["mixed"]["T"] - res["fp64"]["T"]).max() < 5e-4
    steady = T.steady_solve(m, tol=1e-10, precision="fp64")["T"]
    dev = np.abs(res["fp64"]["T"] - steady).max() / steady.max()
    assert 1e-5 < dev < 2e-3, dev
    # the deviation decays: the last steps approach the steady state
    assert np.all(np.diff(res["fp64"]["T_max"][-10:]) > 0)
    assert res["fp64"]["T_max"][-1] < steady.max()


def test_graphs_match_direct_launches():
    """Replayed graphs run the same kernels with deterministic reductions, so
    a transient with 1/dt changing between replays is bit-identical."""
    m = hotspot(40)
    out = []
    for graphs in (True, False):
        s = T.Solver(m, graphs=graphs)
        out.append(T.run_transient(s, t_end=0.01, growth=1.3, tol=1e-9))
    assert np.array_equal(out[0]["T"], out[1]["T"])
    assert np.array_equal(out[0]["iters"], out[1]["iters"])


def test_set_power_and_warm_start():
    m = hotspot(32)
    s = T.Solver(m)
    s.set_x_zero()
    s.set_rhs_steady()
    s.solve(1e-10)
    x1 = s.get_x()
    s.set_power(2.0 * m.get("P"))
    s.set_rhs_steady()
    st = s.solve(1e-10)  # warm start from x1
    x2 = s.get_x()
    assert st["converged"]
    assert np.abs(x2 - 2.0 * x1).max() < 1e-8 * x2.max()


@pytest.mark.parametrize("graphs", [True, False])
def test_rhs_steady_is_a_copy_of_p(graphs):
    """set_rhs_steady sets b = P (the solve reads P in place until b is next
    written): set_power after it leaves b at the old power, as the copy
    would, and the next set_rhs_steady takes the new one. The true residual
    check and its restart read the same right-hand side."""
    m = hotspot(48)
    P = m.get("P")
    ref = T.Solver(m, graphs=graphs, true_residual=True)
    ref.set_x_zero()
    ref.set_rhs_steady()
    st0 = ref.solve(1e-10)
    x0 = ref.get_x()
    s = T.Solver(m, graphs=graphs, true_residual=True)
    s.set_x_zero()
    s.set_rhs_steady()
    s.set_power(3.0 * P)
    st1 = s.solve(1e-10)
    assert st1["iters"] == st0["iters"] and np.array_equal(s.get_x(), x0)
    s.set_x_zero()
    s.set_rhs_steady()
    st2 = s.solve(1e-10)
    assert st2["converged"] and s.true_relres() < 1e-10
    assert np.abs(s.get_x() - 3.0 * x0).max() < 1e-8 * 3.0 * x0.max()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_true_residual_restart(precision):
    """With the true-residual check, the solve only reports convergence once
    ||P - A T|| / ||P|| itself is below tol (the recurrence drifts by
    rounding on large grids)."""
    m = hotspot(512, at=(5e-3, 5e-3))
    s = T.Solver(m, precision=precision, true_residual=True)
    s.set_x_zero()
    s.set_rhs_steady()
    st = s.solve(1e-10, 200)
    assert st["converged"]
    assert s.true_relres() < 1e-10
