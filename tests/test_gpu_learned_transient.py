"""The learned start of backward-Euler steps (th_transient_opts.predictor 2
and 3): each step starts from x + w (G - x), x = T_old (or the extrapolated
start), G the guide field (the surrogate's predicted field), and w >= 0 the
weight minimising the step's starting residual ||b - (C/dt + A) x||.

* The arithmetic: steps run with maxit = 0 leave their start as the result,
  which is compared with a numpy restatement of the rule built on the
  oracle's operator: w = (r0 . M d) / (M d . M d) with d = G - x, r0 = b - M x,
  M = C/dt + A, b = P + C/dt T_old.
* The converged trajectory equals the plain run's to the solve tolerance,
  with the per-step energy balance.
* The trained surrogate as the guide on its own grid.
"""
import os

import numpy as np
import pytest

import oracle as O
import thermal_amd as T

pytestmark = pytest.mark.gpu


def model(n):
    return T.hotspot_model(n)


def numpy_start(o, a, x, G, dt, b):
    o.set_dt(dt)
    Mx, MG = o.apply(x), o.apply(G)
    r0, Md = b - Mx, MG - Mx
    num, den = float(r0 @ Md), float(Md @ Md)
    w = num / den if num > 0 and den > 0 else 0.0
    return x + w * (G - x), w


@pytest.mark.parametrize("mode", ["learned", "extrapolate+learned"])
def test_learned_start_rule(mode):
    m = model(24)
    a = m.arrays()
    steady = O.steady(a, tol=1e-10)["T"]
    rng = np.random.default_rng(3)
    G = 0.9 * steady + 0.05 * rng.standard_normal(m.n)  # a guide with rough error, as a network's
    o = O.Solver(a, "mg")
    cap = a["cap"].astype(np.float64)
    P = a["P"]
    kw = dict(dt0=1e-3, growth=1.5, dt_max=1e30, t_end=1e30, maxit=0, tol=1e-9)
    for steps in (1, 2, 3):
        s = T.Solver(m)
        s.set_guide(G)
        g = T.run_transient(s, predictor=mode, max_steps=steps, **kw)
        used, w_mean = s.guide_stats()
        s.close()
        # restatement: the starts of the steps, each the result of the last
        x, xprev, dt_prev, ws = np.zeros(m.n), None, None, []
        for k in range(steps):
            dt = 1e-3 * 1.5 ** k
            b = P + cap / dt * x
            start = x.copy()
            if mode.startswith("extrapolate") and dt_prev is not None:
                start = x + (dt / dt_prev) * (x - xprev)
            new, w = numpy_start(o, a, start, G, dt, b)
            ws.append(w)
            xprev, x, dt_prev = x, new, dt
        assert np.abs(g["T"] - x).max() <= 1e-10 * np.abs(x).max(), (steps, np.abs(g["T"] - x).max())
        pos = [w for w in ws if w > 0]
        assert used == len(pos)
        assert abs(w_mean - (np.mean(pos) if pos else 0.0)) <= 1e-10 * max(1.0, abs(w_mean))


def test_learned_transient_matches_plain():
    """With the steady field as the guide (what a perfect surrogate would
    give), 30 steps from ambient: the same t and dt, the field of every run
    within the solve tolerance of the plain run's, energy balance 1e-6 per
    step, and no more iterations than the plain warm start."""
    m = model(96)
    a = m.arrays()
    steady = O.steady(a, tol=1e-10)["T"]
    kw = dict(dt0=1e-4, growth=1.3, dt_max=1e30, t_end=1e30, max_steps=30, tol=1e-9)
    runs = {}
    for mode in ("plain", "learned", "extrapolate", "extrapolate+learned"):
        s = T.Solver(m)
        s.set_guide(steady)
        runs[mode] = T.run_transient(s, predictor=mode, **kw)
        runs[mode]["guide"] = s.guide_stats()
        s.close()
    P = m.total_power()
    base = runs["plain"]
    for mode, g in runs.items():
        assert g["converged"] and np.array_equal(g["t"], base["t"])
        assert np.abs(g["T"] - base["T"]).max() < 1e-7 * base["T"].max(), mode
        assert (np.abs(g["q_out"] + g["q_stored"] - P) / P).max() < 1e-6, mode
        print(f"{mode}: {int(g['iters'].sum())} iterations, guide {g['guide']}")
    assert runs["learned"]["iters"].sum() <= base["iters"].sum()
    assert runs["learned"]["guide"][0] > 0


def test_surrogate_guided_transient():
    """The trained network's prediction (lifted) as the guide of a 256^2 x 10
    floorplan of the kind it was trained on: the run converges to the plain
    run's answer; the iterations of both are printed."""
    from thermal_amd import surrogate as TS

    if not os.path.exists(TS.WEIGHTS):
        pytest.skip("no trained weights")
    from thermal_amd import workloads as W

    spec = TS.sample_spec(424242, 7000)
    m = W.build(spec, TS.GRID)
    sur = TS.load()
    d = sur.predict_device(TS.features(spec, TS.GRID))
    kw = dict(dt0=1e-5, growth=1.1, dt_max=1e30, t_end=1e30, max_steps=100, tol=1e-9)
    runs = {}
    for mode in ("plain", "learned", "extrapolate", "extrapolate+learned"):
        s = T.Solver(m)
        s.set_guide_lifted(d, TS.GRID, TS.GRID, 16, TS.COARSE_Z, TS.plane_centres(spec["layers"]),
                           scale=T.Surrogate.T_SCALE, passes=1)
        runs[mode] = T.run_transient(s, predictor=mode, **kw)
        runs[mode]["guide"] = s.guide_stats()
        s.close()
    sur.close()
    P = m.total_power()
    base = runs["plain"]
    for mode, g in runs.items():
        assert g["converged"]
        assert np.abs(g["T"] - base["T"]).max() < 1e-7 * base["T"].max(), mode
        assert (np.abs(g["q_out"] + g["q_stored"] - P) / P).max() < 1e-6, mode
        print(f"surrogate-guided {mode}: {int(g['iters'].sum())} iterations, {g['solve_ms']:.1f} ms, "
              f"guide {g['guide']}")
