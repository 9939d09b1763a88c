"""GPU tests of the batched solver: per-instance results equal to the
single-instance solver's (same kernel bodies, same reductions), convergence
masks, the autotuner, and a small instance of the batched workload."""
import numpy as np
import pytest

import oracle as O
import thermal_amd as T
from thermal_amd import workloads as W

pytestmark = pytest.mark.gpu


def hotspot(n, ny=None, at=(3e-3, 3e-3)):
    m = T.Model.default(n, ny or n)
    m.add_uniform_power(0, 22.0)
    m.add_hotspot(0, at[0], at[1], 1e-3, 3.0)
    m.add_uniform_power(2, 25.0)
    return m.finalize()


def mixed_models():
    return [T.Model.example_case(37, 23), hotspot(64), hotspot(48, 40, at=(5e-3, 5e-3)),
            T.Model.example_case(16, 12), T.Model.example_case(100, 60)]


def single(m, tol):
    s = T.Solver(m)
    s.set_x_zero()
    s.set_rhs_steady()
    st = s.solve(tol, 200)
    return st, s.get_x()


def test_batch_equals_single_solves():
    """Instances of different sizes (one small enough that its whole
    V-cycle is the tail) converge in the same iterations to the same bits
    as separate solves."""
    ms = mixed_models()
    b = T.Batch(ms)
    r = b.solve(1e-10, 200)
    assert r["converged"].all()
    for i, m in enumerate(ms):
        st, x = single(m, 1e-10)
        assert r["iters"][i] == st["iters"], i
        assert np.array_equal(b.get_x(i), x), i


def test_typed_16_plane_instances_equal_single_solves():
    """The 16-plane TSV stack's floorplan at 256^2 (column types on its top
    levels), at 300 x 212 (as many columns, but the TSV clusters cover cells
    in part, so no types: per-cell arrays) and at 64^2, in one batch: each
    instance equals its single solve bit for bit, in the same iterations."""
    spec = W.config2_spec(1000)
    ms = [W.build(spec, 256), W.build(spec, 300, 212), W.build(spec, 64)]
    b = T.Batch(ms)
    r = b.solve(1e-10, 200)
    assert r["converged"].all()
    for i, m in enumerate(ms):
        s = T.Solver(m)
        assert (s.column_types() > 0) == (i == 0), i
        s.close()
        st, x = single(m, 1e-10)
        assert r["iters"][i] == st["iters"], i
        assert np.array_equal(b.get_x(i), x), i


def test_convergence_masks():
    """A zero-power instance finishes at the start (x = 0) while the others
    iterate; their results are unaffected."""
    ms = mixed_models()[:3]
    b = T.Batch(ms)
    b.set_power(1, np.zeros(64 * 64 * 10))
    r = b.solve(1e-10, 200)
    assert r["converged"].all() and r["iters"][1] == 0
    assert not b.get_x(1).any()
    st, x = single(ms[0], 1e-10)
    assert r["iters"][0] == st["iters"] and np.array_equal(b.get_x(0), x)


def test_autotuned_shapes_give_the_same_answers():
    ms = mixed_models()
    b = T.Batch(ms)
    tune = b.autotune(3)
    assert tune["rows"] in (1, 2, 4, 8) and tune["vec_threads"] in (256, 1024)
    assert b.shape() == (tune["rows"], tune["cols_per_wave"], tune["vec_threads"])
    r = b.solve(1e-10, 200)
    for i, m in enumerate(ms):
        st, x = single(m, 1e-10)
        assert abs(int(r["iters"][i]) - st["iters"]) <= 1
        assert np.abs(b.get_x(i) - x).max() <= 1e-8 * np.abs(x).max()


def test_small_batched_workload_vs_oracle():
    """Eight instances of the batched workload's family on small grids: each
    converges, and matches the oracle's FP64 answer."""
    inst = W.config3_instances(seed=42, count=8, sizes=(24, 32, 40))
    models = [W.build(spec, n) for n, spec in inst]
    b = T.Batch(models)
    r = b.solve(1e-10, 200)
    assert r["converged"].all()
    for i, m in enumerate(models):
        a = m.arrays()
        ref = O.steady(a, tol=1e-11, mg_double=True)
        x = b.get_x(i)
        assert np.abs(x - ref["T"]).max() < 5e-4
        assert O.true_relres(a, x) < 1e-9


def test_full_batched_workload():
    """The benchmark's batched workload itself: 64 instances of 128^2, 256^2
    or 512^2 x 10 cells (seeded, blocks on with p = 0.6; the 256^2 and
    512^2 instances have column types). With the default shapes (1024-thread
    vector blocks, as the single-instance solver) every instance equals its
    single-instance solve bit for bit, in the same iterations. Then one
    autotuned batched solve to 1e-10 (its vector blocks may sum the dot
    products in another order): every instance converges with energy balance
    1e-6 and a true relative residual <= 1e-9, matches the single solve to
    1e-10 of its peak rise and takes its iteration count within one; the
    smallest and the largest instance match the FP64 oracle to 5e-4 K."""
    inst = W.config3_instances(count=64)
    models = [W.build(spec, n) for n, spec in inst]
    b = T.Batch(models)
    r0 = b.solve(1e-10, 200)
    singles = [single(m, 1e-10) for m in models]
    for i, (st, xs) in enumerate(singles):
        assert r0["iters"][i] == st["iters"], i
        assert np.array_equal(b.get_x(i), xs), i
    tune = b.autotune(2)
    assert tune["cols_per_wave"] in (16, 32, 64)
    r = b.solve(1e-10, 200)
    assert r["converged"].all()
    sizes = [m.nx for m in models]
    for i, m in enumerate(models):
        x = b.get_x(i)
        P = m.total_power()
        assert abs(m.heat_out(x) - P) / P < 1e-6, i
        st, xs = singles[i]
        assert abs(int(r["iters"][i]) - st["iters"]) <= 1, (i, r["iters"][i], st["iters"])
        assert np.abs(x - xs).max() <= 1e-10 * np.abs(xs).max() + 1e-12, i
    for i in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        a = models[i].arrays()
        x = b.get_x(i)
        ref = O.steady(a, tol=1e-10, mg_double=True)
        assert np.abs(x - ref["T"]).max() < 5e-4
        assert O.true_relres_flux(a, x) <= 1e-9
    print(f"64 instances: autotune {tune}, batch {r['ms']:.1f} ms, iterations p50 "
          f"{np.percentile(r['iters'], 50):.0f} p99 {np.percentile(r['iters'], 99):.0f}")
