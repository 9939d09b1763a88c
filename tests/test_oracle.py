"""Pins the oracle (oracle/oracle.cpp, the CPU restatement of the reference's
MG-PCG path) before the GPU's results are compared with it:
  * against the golden fixtures written by the reference's own CPU build
    (tests/golden/make_golden.py), bit for bit where the reference is
    deterministic;
  * against what the mathematics fixes independently of the restatement:
    closed-form conductances, conservation, the row-sum identity, symmetry,
    dense assembly and dense solves, the exact 1D resistor chain.
CPU only.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN


def fx(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def hotspot(n, ny=None, at=(3e-3, 3e-3)):
    m = O.Model(O.DEFAULT_STACK, n, ny or n)
    m.add_uniform_power(0, 22.0)
    m.add_hotspot(0, at[0], at[1], 1e-3, 3.0)
    m.add_uniform_power(2, 25.0)
    return m.finalize()


MM = 1e-3
EXAMPLE_BLOCKS = [  # case3d.h:170-190 (same numbers as cases/example/floorplan.csv)
    (0, 0.5, 0.5, 2.0, 2.0, 6.0), (0, 3.0, 0.5, 2.0, 2.0, 6.0), (0, 0.5, 3.0, 2.0, 2.0, 6.0),
    (0, 3.0, 3.0, 2.0, 2.0, 6.0), (0, 5.5, 0.5, 4.0, 4.5, 3.0), (0, 0.5, 5.5, 9.0, 4.0, 3.0),
    (2, 0.0, 0.0, 10.0, 10.0, 12.0), (2, 4.0, 4.0, 2.0, 2.0, 3.0)]
EXAMPLE_TSVS = [(1, 2.55, 0.5, 0.40, 4.5, 0.15), (1, 0.5, 2.55, 4.5, 0.40, 0.15),
                (2, 2.55, 0.5, 0.40, 4.5, 0.15), (2, 0.5, 2.55, 4.5, 0.40, 0.15)]


def example(nx, ny):
    m = O.Model(O.DEFAULT_STACK, nx, ny)
    for L, x, y, w, h, W in EXAMPLE_BLOCKS:
        m.add_power_block(L, x * MM, y * MM, w * MM, h * MM, W)
    for L, x, y, w, h, f in EXAMPLE_TSVS:
        m.add_tsv_region(L, x * MM, y * MM, w * MM, h * MM, f)
    return m.finalize()


def assert_model_equal(a, f):
    for k in ("gE", "gN", "gU", "gB", "cap", "P", "tsv_frac", "dz"):
        assert np.array_equal(a[k], f[k]), k


# ---- model (stack3d.h) ------------------------------------------------------

def test_model_matches_reference_fixtures():
    assert_model_equal(hotspot(13, 9).arrays(), fx("model_hotspot_13x9"))
    assert_model_equal(example(20, 14).arrays(), fx("model_example_20x14"))


def test_conductances_closed_form():
    """One 3-plane slab (k = 130): lateral G = k dy dz / dx, vertical
    G = k A / dz, top G = A / (dz/2k + 1/h): the harmonic mean of equal k is k."""
    k, t, h = 130.0, 60e-6, 2e4
    m = O.Model([(t, k, 1.66e6, 3, 1)], 2, 2, Lx=1e-3, Ly=1e-3, h_top=h).finalize()
    a = m.arrays()
    dx = dy = 0.5e-3
    dz = t / 3
    A = dx * dy
    gE = a["gE"].reshape(3, 2, 2)
    gN = a["gN"].reshape(3, 2, 2)
    gU = a["gU"].reshape(3, 2, 2)
    assert np.allclose(gE[:, :, 0], k * dy * dz / dx, rtol=1e-7) and not gE[:, :, 1].any()
    assert np.allclose(gN[:, 0, :], k * dx * dz / dy, rtol=1e-7) and not gN[:, 1, :].any()
    assert np.allclose(gU[:2], k * A / dz, rtol=1e-7)
    assert np.allclose(gU[2], A / (dz / (2 * k) + 1 / h), rtol=1e-7)
    assert not a["gB"].any()  # adiabatic bottom
    assert np.allclose(a["cap"], 1.66e6 * A * dz, rtol=1e-7)
    # two materials in series across the vertical face: 1/G = d1/2k1 + d2/2k2
    m2 = O.Model([(10e-6, 2.0, 1.7e6, 1, 0), (50e-6, 130.0, 1.66e6, 1, 1)], 1, 1, Lx=1e-3, Ly=1e-3).finalize()
    g = m2.arrays()["gU"][0]
    assert np.isclose(g, 1e-6 / (10e-6 / 4.0 + 50e-6 / 260.0), rtol=2e-7)


def test_tsv_effective_medium_limits():
    for k in (2.0, 130.0):
        assert O.eff_kz(k, 400.0, 0.0) == k and O.eff_kxy(k, 400.0, 0.0) == k
        assert O.eff_kz(k, 400.0, 1.0) == 400.0 and np.isclose(O.eff_kxy(k, 400.0, 1.0), 400.0)
        # Maxwell-Garnett lies below the parallel (Voigt) mix for 0 < f < 1
        assert O.eff_kxy(k, 400.0, 0.3) < O.eff_kz(k, 400.0, 0.3)


def test_power_conserved_off_grid():
    """main3d.cu:232-241: blocks shifted off the cell edges keep their power."""
    m = O.Model(O.DEFAULT_STACK, 64, 64)
    want = 0.0
    for L, x, y, w, h, W in EXAMPLE_BLOCKS:
        sx = 0.037 if x + w + 0.037 <= 10 else 0.0
        sy = 0.011 if y + h + 0.011 <= 10 else 0.0
        m.add_power_block(L, (x + sx) * MM, (y + sy) * MM, w * MM, h * MM, W)
        want += W
    assert abs(m.total_power() - want) / want < 1e-12


@pytest.mark.parametrize("at,count", [((5e-3, 5e-3), 12), ((3e-3, 3e-3), 13)])
def test_hotspot_cell_centre_rule(at, count):
    m = O.Model(O.DEFAULT_STACK, 128, 128)
    m.add_hotspot(0, at[0], at[1], 1e-3, 3.0)
    P = m.arrays()["P"].reshape(10, 128, 128)
    hit = P[0] > 0
    assert hit.sum() == count * count and (P[1] > 0).sum() == count * count and not P[2:].any()
    assert np.isclose(P.sum(), 3.0, rtol=1e-14)


# ---- operator (operator3d.h) -----------------------------------------------

def assemble_dense(a, inv_dt=0.0):
    """Dense (A + C/dt) from the face list, independent of the stencil code."""
    nx, ny, nz = a["nx"], a["ny"], a["nz"]
    n = nx * ny * nz
    M = np.zeros((n, n))
    idx = lambda i, j, l: (l * ny + j) * nx + i  # noqa: E731
    def link(p, q, g):
        M[p, p] += g
        M[q, q] += g
        M[p, q] -= g
        M[q, p] -= g
    for l in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = idx(i, j, l)
                if i + 1 < nx:
                    link(c, idx(i + 1, j, l), float(a["gE"][c]))
                if j + 1 < ny:
                    link(c, idx(i, j + 1, l), float(a["gN"][c]))
                if l + 1 < nz:
                    link(c, idx(i, j, l + 1), float(a["gU"][c]))
                else:
                    M[c, c] += float(a["gU"][c])  # heat sink
                if l == 0:
                    M[c, c] += float(a["gB"][j * nx + i])
                M[c, c] += float(a["cap"][c]) * inv_dt
    return M


@pytest.mark.parametrize("inv_dt", [0.0, 1e3])
def test_operator_dense_assembly_and_identities(inv_dt):
    m = O.Model(O.DEFAULT_STACK[:3] + [(25e-6, 5.0, 2e6, 1, 0)], 4, 3, h_bot=1e3)
    m.add_uniform_power(0, 1.0)
    m.add_tsv_region(1, 2e-3, 2e-3, 4e-3, 3e-3, 0.4)
    a = m.finalize().arrays()
    s = O.Solver(a, "jacobi")
    s.set_dt(1.0 / inv_dt if inv_dt else 0.0)
    M = assemble_dense(a, inv_dt)
    x = np.random.default_rng(0).standard_normal(M.shape[0])
    assert np.allclose(s.apply(x), M @ x, rtol=1e-12, atol=1e-18)
    # symmetry x.Ay = y.Ax and row sums = ambient links + C/dt
    y = np.random.default_rng(1).standard_normal(M.shape[0])
    assert np.isclose(x @ s.apply(y), y @ s.apply(x), rtol=1e-12)
    ones = s.apply(np.ones(M.shape[0]))
    amb = np.zeros(M.shape[0])
    npl = a["nx"] * a["ny"]
    amb[-npl:] += a["gU"][-npl:]
    amb[:npl] += a["gB"]
    amb += a["cap"].astype(np.float64) * inv_dt
    assert np.allclose(ones, amb, rtol=1e-9, atol=1e-18)


def test_pcg_matches_dense_solve():
    a = hotspot(6, 5).arrays()
    T = np.linalg.solve(assemble_dense(a), a["P"])
    for pc in ("mg", "jacobi"):
        r = O.steady(a, tol=1e-13, precond=pc)
        assert np.allclose(r["T"], T, rtol=1e-10), pc


def test_backward_euler_step_matches_dense_solve():
    a = hotspot(6, 5).arrays()
    dt = 1e-4
    T1 = np.linalg.solve(assemble_dense(a, 1.0 / dt), a["P"])  # from T = 0
    tr = O.run_transient(a, dt0=dt, growth=1.0, dt_max=dt, t_end=dt, tol=1e-13)
    assert len(tr["t"]) == 1 and np.allclose(tr["T"], T1, rtol=1e-10)


# ---- multigrid (mg3d.h) -----------------------------------------------------

def test_levels_match_reference_fixture():
    f = fx("levels_example_21x14")
    s = O.Solver(example(21, 14).arrays(), "mg")
    assert s.nlevels() == int(f["nlevels"])
    for k in range(s.nlevels()):
        L = s.level(k)
        assert (L["nx"], L["ny"]) == (int(f[f"L{k}_nx"]), int(f[f"L{k}_ny"]))
        for name in ("gE", "gN", "gU", "gB", "cap"):
            assert np.array_equal(L[name], f[f"L{k}_{name}"]), (k, name)


def test_coarse_operator_of_uniform_material_is_the_coarse_grid():
    """mg3d.h:17-22: aggregating a uniform fine grid gives finalize() on the
    coarse grid; vertical, ambient and capacity sums are conserved."""
    layers = [(40e-6, 130.0, 1.66e6, 2, 1), (1e-3, 400.0, 3.45e6, 2, 0)]
    fine = O.Model(layers, 16, 12).finalize().arrays()
    coarse = O.Model(layers, 8, 6).finalize().arrays()
    L1 = O.Solver(fine, "mg").level(1)
    for name in ("gE", "gN", "gU", "cap"):
        assert np.allclose(L1[name], coarse[name], rtol=2e-6), name
    s = O.Solver(example(21, 14).arrays(), "mg")
    for k in range(1, s.nlevels()):
        for name in ("gU", "cap", "gB"):
            assert np.isclose(s.level(k)[name].astype(np.float64).sum(),
                              s.level(0)[name].astype(np.float64).sum(), rtol=1e-5), (k, name)


def test_column_factors_solve_the_column():
    """Thomas factors (mg3d.h:46-61) against a dense tridiagonal solve."""
    a = example(9, 7).arrays()
    s = O.Solver(a, "mg", mg_double=True)
    s.set_dt(1e-3)
    L = s.level(0)
    nx, ny, nz = 9, 7, 10
    i, j = 4, 3
    cells = [(l * ny + j) * nx + i for l in range(nz)]
    M = assemble_dense(a, 1e3)[np.ix_(cells, cells)]  # column block, laterals held fixed
    b = np.random.default_rng(5).standard_normal(nz)
    fcp, finv = L["fcp"][cells], L["finv"][cells]
    y = np.zeros(nz)
    prev = 0.0
    for l in range(nz):
        lower = M[l, l - 1] if l else 0.0
        prev = (b[l] - lower * prev) * finv[l]
        y[l] = prev
    for l in range(nz - 2, -1, -1):
        y[l] -= fcp[l] * y[l + 1]
    assert np.allclose(y, np.linalg.solve(M, b), rtol=1e-12)


def test_vcycle_matches_reference_fixture():
    f = fx("vcycle_example_21x14")
    s = O.Solver(example(21, 14).arrays(), "mg")
    assert np.array_equal(s.precond(f["r"]), f["z_steady"])
    s.set_dt(1e-3)
    assert np.array_equal(s.precond(f["r"]), f["z_dt1e3"])


def test_vcycle_is_symmetric():
    """Restriction is the transpose of prolongation and the smoothing order
    is mirrored, so M^-1 is symmetric (exact in FP64 up to rounding)."""
    s = O.Solver(example(21, 14).arrays(), "mg", mg_double=True)
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal((2, 21 * 14 * 10))
    assert np.isclose(u @ s.precond(v), v @ s.precond(u), rtol=1e-11)
    assert u @ s.precond(u) > 0


@pytest.mark.parametrize("dt", [0.0, 1e-3])
def test_lateral_restriction_equals_residual_restriction(dt):
    """After the red sweep from x = 0 and the black sweep, the residual of
    every black cell is zero and that of a red cell is its lateral sum G x;
    restricting the red children's lateral sums therefore gives the
    reference's restriction of b - A x, which the fixture above pins.  In
    FP64 the two V-cycles agree to rounding (a dropped child or a wrong sign
    would differ at O(1))."""
    a = example(21, 14).arrays()
    f, s = (O.Solver(a, "mg", mg_double=True, fast_restrict=k) for k in (True, False))
    f.set_dt(dt)
    s.set_dt(dt)
    r = np.random.default_rng(5).standard_normal(21 * 14 * 10)
    zf, zs = f.precond(r), s.precond(r)
    assert np.abs(zf - zs).max() <= 1e-12 * np.abs(zs).max()
    u = np.random.default_rng(6).standard_normal(r.size)
    assert np.isclose(u @ f.precond(r), r @ f.precond(u), rtol=1e-11)
    f32 = O.Solver(a, "mg", fast_restrict=True)
    f32.set_dt(dt)
    assert np.abs(f32.precond(r) - zs).max() <= 1e-5 * np.abs(zs).max()


# ---- PCG (pcg3d_cpu.cpp) ---------------------------------------------------

def test_pcg_matches_reference_fixture():
    f = fx("steady_hotspot_24")
    a = hotspot(24).arrays()
    for tol, key in ((1e-10, "mg_1e10"), (1e-9, "mg_1e9")):
        r = O.steady(a, tol=tol)
        assert r["iters"] == int(f[f"iters_{key}"])
        assert np.array_equal(r["T"], f[f"T_{key}"])
    r = O.steady(a, tol=1e-9, precond="jacobi")
    assert r["iters"] == int(f["iters_jacobi_1e9"]) and np.array_equal(r["T"], f["T_jacobi_1e9"])
    r = O.steady(a, tol=1e-10)
    assert O.true_relres(a, r["T"]) == float(f["true_relres_mg_1e10"])
    assert O.heat_out(a, r["T"]) == float(f["heat_out_mg_1e10"])


def test_config0_matches_reference_fixture():
    f = fx("steady_config0_128")
    a = hotspot(128, at=(5e-3, 5e-3)).arrays()
    r = O.steady(a, tol=1e-10)
    assert r["iters"] == int(f["iters"])
    assert np.array_equal(r["T"][f["sample_idx"]], f["sample_T"])
    assert np.array_equal(r["T"].reshape(10, -1).max(1), f["plane_max"])


def test_flexible_beta_agrees_with_fletcher_reeves():
    a = hotspot(32).arrays()
    fr, fl = O.steady(a, tol=1e-10), O.steady(a, tol=1e-10, flexible=True)
    assert abs(fr["iters"] - fl["iters"]) <= 1
    assert np.abs(fr["T"] - fl["T"]).max() < 1e-9 * fr["T"].max()


def test_uniform_power_is_the_1d_chain():
    f = fx("steady_uniform_128")
    m = O.Model(O.DEFAULT_STACK, 64, 64)
    m.add_uniform_power(0, 25.0)
    m.add_uniform_power(2, 25.0)
    m.finalize()
    Tz = m.analytic_1d()
    assert np.array_equal(Tz, f["analytic"])  # the chain is grid-independent
    r = O.steady(m.arrays(), tol=1e-10)
    assert np.abs(r["T"].reshape(10, -1) - Tz[:, None]).max() / Tz.max() < 1e-5
    assert np.allclose(f["plane_max"], Tz, rtol=1e-5)


def test_flux_form_residual():
    a = hotspot(32).arrays()
    r = O.steady(a, tol=1e-12)
    assert O.true_relres_flux(a, r["T"]) < 1e-11 and O.true_relres(a, r["T"]) < 1e-11
    x = np.random.default_rng(0).random(a["nx"] * a["ny"] * a["nz"])
    assert np.isclose(O.true_relres_flux(a, x), O.true_relres(a, x), rtol=1e-12)


# ---- transient (transient3d.h) ---------------------------------------------

def test_transient_matches_reference_fixture():
    f = fx("transient_hotspot_24")
    a = hotspot(24).arrays()
    tr = O.run_transient(a, t_end=0.02, growth=1.2, tol=1e-9)
    for k in ("t", "dt", "iters", "T_max", "q_out", "q_stored"):
        assert np.array_equal(tr[k], f[k]), k
    assert np.array_equal(tr["T"], f["T"])
    P = float(f["total_power"])
    assert (np.abs(tr["q_out"] + tr["q_stored"] - P) / P).max() < 1e-6


def test_transient_settles_to_steady_state():
    a = hotspot(16).arrays()
    tr = O.run_transient(a, t_end=3.0, dt_max=0.1, tol=1e-10)
    ss = O.steady(a, tol=1e-10)
    assert np.abs(tr["T"] - ss["T"]).max() / ss["T"].max() < 1e-5


def test_reference_harness_agrees_where_available():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    r = O.RefModel.example_case(17, 11)
    a = r.arrays()
    assert_model_equal(a, example(17, 11).arrays())
    x = O.steady(a, tol=1e-10)
    y = r.cpu_pcg(1e-10)
    assert x["iters"] == y["iters"] and np.array_equal(x["T"], y["T"])
