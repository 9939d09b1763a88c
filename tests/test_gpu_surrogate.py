"""GPU tests of the learned warm start: the surrogate's BF16 MFMA forward
and backward kernels against the CPU oracle (oracle/surrogate.py), training
on solver-generated data, and the lifted initial guess with its fallback."""
import os

import numpy as np
import pytest

import surrogate as S  # oracle/surrogate.py
import thermal_amd as T
from thermal_amd import surrogate as TS
from thermal_amd import workloads as W

pytestmark = pytest.mark.gpu


class DeviceArray:
    """A float32 array copied to device memory through the HIP runtime the
    library itself links (test plumbing)."""

    def __init__(self, a):
        import ctypes as C

        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        self.ptr = p.value

    def __del__(self):
        self.hip.hipFree(self.C.c_void_p(self.ptr))


def features_batch(B, n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((B, n, n, 3), np.float32)
    x[..., :2] = rng.gamma(1.0, 0.3, (B, n, n, 2))
    x[..., 2] = rng.random((B, n, n)) < 0.1
    return x


def padded(x):
    xp = np.zeros(x.shape[:3] + (16,), np.float32)
    xp[..., :3] = x
    return xp


@pytest.mark.parametrize("n,B", [(32, 2), (64, 1)])
def test_forward_matches_oracle(n, B):
    sur = T.Surrogate(n, n, seed=11)
    p = sur.get_params()
    x = features_batch(B, n, 1)
    out = sur.predict(x)
    ref, _ = S.forward(p, padded(x), round_bf16=True)
    ref = ref[..., :S.N_OUT]
    scale = np.abs(ref).max()
    assert scale > 0
    # FP32 accumulation order differs; BF16 storage then rounds the same
    # values, except for rare ties that propagate as ~1 bf16 ulp.
    assert np.abs(out - ref).max() < 2e-2 * scale, np.abs(out - ref).max() / scale
    assert np.sqrt(np.mean((out - ref) ** 2)) < 3e-3 * scale


@pytest.mark.parametrize("n,B", [(32, 2), (48, 3)])
def test_gradients_match_oracle(n, B):
    """Loss and every layer's gradient against the oracle. 48 x 48: rows of
    48, 24, 12, 6 and 3 pixels, so the weight gradient's K-blocks of 32
    pixels straddle rows and images and the last block of a chunk is
    partly past the batch."""
    sur = T.Surrogate(n, n, seed=12)
    p = sur.get_params()
    x = features_batch(B, n, 2)
    tgt = np.random.default_rng(3).random((B, n, n, S.N_OUT)).astype(np.float32)
    loss, g = sur.grad(x, tgt)
    ref_loss, ref_g = S.loss_and_grad(p, padded(x), tgt, round_bf16=True)
    assert abs(loss - ref_loss) < 1e-3 * ref_loss
    off = 0
    for name, _, ci, co, ks, _ in S.LAYERS:
        k = co * ks * ks * ci + co
        a, b = g[off:off + k].astype(np.float64), ref_g[off:off + k]
        off += k
        nb = np.linalg.norm(b)
        assert nb > 0, name
        # BF16 gradients between layers: a few parts in a hundred per layer
        assert np.linalg.norm(a - b) < 5e-2 * nb, (name, np.linalg.norm(a - b) / nb)
        assert a @ b > 0.998 * np.linalg.norm(a) * nb, name


def test_training_reduces_loss_on_solver_data():
    """A few hundred Adam steps on 64x64 solver-generated samples: the
    training loss falls and the held-out error beats the untrained net."""
    n = 64
    sur = T.Surrogate(n, n, seed=3)
    sur.reserve(72)
    info = TS.generate(sur, 72, n_tsv=3, seed0=500, log=None)
    assert info["samples"] == 72
    before = sur.eval(64, 8)
    losses, ms = sur.train(300, batch=8, lr=3e-3, lr_final=3e-4, seed=5, n_train=64)
    after = sur.eval(64, 8)
    assert np.mean(losses[-20:]) < 0.2 * np.mean(losses[:5])
    assert after < 0.2 * before, (before, after)


def test_lift_of_exact_field_and_fallback():
    """The exact solution given as a coarse field on the solver's own grid
    lifts to itself (a converged start); a zero field loses to the plain
    warm start and leaves x untouched."""
    n = 64
    spec = TS.sample_spec(9, 7000)
    m = W.build(spec, n)
    s = T.Solver(m, precision="fp64")
    s.set_x_zero()
    s.set_rhs_steady()
    s.solve(1e-12, 200)
    x = s.get_x()
    field = np.zeros((n, n, 16), np.float32)
    field[..., :10] = np.moveaxis(x.reshape(10, n, n), 0, -1)
    d = DeviceArray(field)
    zc = TS.plane_centres(W.config0_layers())
    s2 = T.Solver(m, precision="fp64")
    s2.set_x_zero()
    used, rp, rl = s2.set_x_lifted(d.ptr, n, n, 16, zc, zc, scale=1.0, passes=0)
    assert used and np.isclose(rp, 1.0) and rl < 1e-2
    s2.set_rhs_steady()
    st = s2.solve(1e-6, 50)
    assert st["converged"] and st["iters"] <= 4
    z = DeviceArray(np.zeros_like(field))
    s3 = T.Solver(m, precision="fp64")
    s3.set_x(x)
    used, rp, rl = s3.set_x_lifted(z.ptr, n, n, 16, zc, zc, scale=1.0, passes=1)
    assert not used and rp < 1e-10
    assert np.array_equal(s3.get_x(), x)


@pytest.mark.skipif(not os.path.exists(TS.WEIGHTS), reason="no trained weights")
def test_trained_surrogate_error_and_warm_start():
    """The trained network on floorplans it has not seen (256 x 256 x 10):
    its error against the converged PCG answer, and iterations to 1e-10
    from T = 0 and from the surrogate's lifted guess (with fallback). The
    MG-PCG reduces the error ~5x per iteration at every scale, so a guess
    this accurate saves no iteration: what the network buys is the answer
    itself, to ~2%, from one forward pass."""
    sur = TS.load()
    fine = W.config0_layers()
    errs, rms, cold, warm = [], [], [], []
    for k in range(4):
        spec = TS.sample_spec(990000 + k, 7000 + 2 * (60 + k))
        m = W.build(spec, 256)
        s = T.Solver(m)
        s.set_x_zero()
        s.set_rhs_steady()
        cold.append(s.solve(1e-10, 200)["iters"])
        xt = s.get_x().reshape(10, 256, 256)
        pred = sur.predict(TS.features(spec)[None])[0] * T.Surrogate.T_SCALE
        d = np.moveaxis(pred, -1, 0) - xt
        errs.append(np.abs(d).max() / xt.max())
        rms.append(np.sqrt(np.mean(d * d)) / xt.max())
        s.set_x_zero()
        used, rp, rl, _ = TS.warm_start(s, sur, spec, fine)
        s.set_rhs_steady()
        warm.append(s.solve(1e-10, 200)["iters"])
    print("surrogate max / rms rel error", errs, rms, "cold", cold, "warm", warm)
    # measured after 12000 steps on 20000 samples: rms ~2%, max 7-12% of the
    # peak rise (at the hottest small blocks)
    assert max(rms) < 0.05 and max(errs) < 0.2
    # the lifted guess never costs iterations (fallback on the start residual)
    assert all(w <= c for w, c in zip(warm, cold))


def _line_relax_np(a, nx, ny, nz, x, colour):
    """One colour of the FP64 column line relaxation of the steady operator
    (k_line_relax's arithmetic, in its order), vectorised over the columns;
    x is the natural (nz, ny, nx) field, updated in place."""
    sh = (nz, ny, nx)
    gE, gN, gU = (a[k].astype(np.float64).reshape(sh) for k in ("gE", "gN", "gU"))
    gB = a["gB"].astype(np.float64).reshape(ny, nx)
    P = a["P"].reshape(sh)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    sel = ((ii + jj) & 1) == colour
    hw, he, hs, hn = ii > 0, ii < nx - 1, jj > 0, jj < ny - 1
    zero = np.zeros((ny, nx))
    cp, dp = np.zeros(sh), np.zeros(sh)
    pcp = pdp = zero
    for l in range(nz):
        Gw = np.where(hw, np.roll(gE[l], 1, axis=1), 0.0)
        Ge = np.where(he, gE[l], 0.0)
        Gs = np.where(hs, np.roll(gN[l], 1, axis=0), 0.0)
        Gn = np.where(hn, gN[l], 0.0)
        Gu, Gd = gU[l], (gU[l - 1] if l > 0 else gB)
        xl = x[l]
        d, rhs = zero.copy(), P[l].copy()
        d = d + Gw; rhs = rhs + Gw * np.roll(xl, 1, axis=1)
        d = d + Ge; rhs = rhs + Ge * np.roll(xl, -1, axis=1)
        d = d + Gs; rhs = rhs + Gs * np.roll(xl, 1, axis=0)
        d = d + Gn; rhs = rhs + Gn * np.roll(xl, -1, axis=0)
        d = d + (Gu + Gd)
        lower = -Gd if l > 0 else zero
        upper = -Gu if l < nz - 1 else zero
        den = d - lower * pcp
        cp[l] = upper / den
        dp[l] = (rhs - lower * pdp) / den
        pcp, pdp = cp[l], dp[l]
    xu = dp[nz - 1]
    x[nz - 1][sel] = xu[sel]
    for l in range(nz - 2, -1, -1):
        xu = dp[l] - cp[l] * xu
        x[l][sel] = xu[sel]


def test_lift_relaxation_on_typed_grid():
    """160 x 160 x 10: the fine level reads column types, so the lifted
    guess's line relaxation runs the typed kernel. A lifted field is relaxed
    once (red then black columns) and compared bit for bit with the same
    relaxation restated in numpy."""
    n = 160
    spec = TS.sample_spec(9, 7001)
    m = W.build(spec, n)
    a = m.arrays()
    s = T.Solver(m, precision="mixed")
    assert s.column_types() > 0
    rng = np.random.default_rng(3)
    field = (20.0 + rng.random((n, n, 16))).astype(np.float32)
    d = DeviceArray(field)
    zc = TS.plane_centres(W.config0_layers())
    s0 = T.Solver(m, precision="mixed")
    s0.set_x(np.full(m.n, 1e9))  # a plain start so bad that the lifted guess is kept
    used, _, _ = s0.set_x_lifted(d.ptr, n, n, 16, zc, zc, scale=1.0, passes=0)
    assert used
    lifted = s0.get_x().reshape(m.nz, n, n)
    s.set_x(np.full(m.n, 1e9))
    used, _, _ = s.set_x_lifted(d.ptr, n, n, 16, zc, zc, scale=1.0, passes=1)
    assert used
    relaxed = s.get_x().reshape(m.nz, n, n)
    ref = lifted.copy()
    _line_relax_np(a, n, n, m.nz, ref, 0)
    _line_relax_np(a, n, n, m.nz, ref, 1)
    assert np.array_equal(relaxed, ref), np.abs(relaxed - ref).max()


def test_predict_device_input_matches_host_input():
    """th_surrogate_predict_device_input (input already in device memory)
    gives the output of th_surrogate_predict_device bit for bit."""
    n = 64
    sur = T.Surrogate(n, n, seed=11)
    x = np.random.default_rng(4).standard_normal((n, n, T.Surrogate.N_IN)).astype(np.float32)
    import ctypes as C

    def fetch(ptr):
        out = np.empty((n, n, 16), np.float32)
        hip = C.CDLL("libamdhip64.so.7")
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2) == 0
        return out

    a = fetch(sur.predict_device(x))
    d = T.DeviceBuffer(x)
    b = fetch(sur.predict_device_input(d.ptr))
    assert np.array_equal(a, b)
    sur.close()


def _lift_np(f, nx, ny, nz, zc, zf, scale):
    """The lift restated in numpy (float64, the kernels' operation order):
    each fine plane from the coarse channels through the plane map (linear
    between the two coarse planes around it, the nearest outside them, the
    weight rounded to float32), then bilinear over the die from cell
    centres, clamped at the edges. Returns [nz][ny][nx]."""
    Hc, Wc, _ = f.shape
    lo, w = np.zeros(nz, int), np.zeros(nz, np.float32)
    for l in range(nz):
        z, k = zf[l], 0
        while k + 1 < len(zc) and zc[k + 1] <= z:
            k += 1
        if k + 1 >= len(zc) or z <= zc[0]:
            lo[l] = 0 if z <= zc[0] else len(zc) - 1
        else:
            lo[l], w[l] = k, np.float32((z - zc[k]) / (zc[k + 1] - zc[k]))
    f64 = f.astype(np.float64)
    fz = np.empty((nz, Hc, Wc))
    for l in range(nz):
        wl = np.float64(w[l])
        fz[l] = (1.0 - wl) * f64[..., lo[l]] + wl * f64[..., lo[l] + 1] if wl > 0.0 else f64[..., lo[l]]
    xc = np.clip((np.arange(nx) + 0.5) * Wc / nx - 0.5, 0.0, Wc - 1.0)
    yc = np.clip((np.arange(ny) + 0.5) * Hc / ny - 0.5, 0.0, Hc - 1.0)
    x0, y0 = xc.astype(int), yc.astype(int)
    x1, y1 = np.minimum(x0 + 1, Wc - 1), np.minimum(y0 + 1, Hc - 1)
    fx, fy = (xc - x0)[None, :], (yc - y0)[:, None]
    out = np.empty((nz, ny, nx))
    for l in range(nz):
        F = fz[l]
        v = (1.0 - fy) * ((1.0 - fx) * F[y0][:, x0] + fx * F[y0][:, x1]) + fy * ((1.0 - fx) * F[y1][:, x0] + fx * F[y1][:, x1])
        out[l] = scale * v
    return out


def test_lift_matches_numpy_restatement():
    """The lifted field itself (no relaxation) against its numpy restatement,
    bit for bit: a 20 x 24 coarse field of 7 planes whose centres lie inside
    the stack (so the bottom and top planes take the nearest one and the
    others interpolate) onto 64 x 64 x 10, scaled."""
    n = 64
    m = W.build(TS.sample_spec(9, 7003), n)
    zf = np.asarray(TS.plane_centres(W.config0_layers()), np.float64)
    zc = np.linspace(zf[1] * 0.9 + zf[0] * 0.1, zf[-2] * 0.9 + zf[-1] * 0.1, 7)
    rng = np.random.default_rng(5)
    f = (10.0 + 5.0 * rng.random((20, 24, 16))).astype(np.float32)
    d = DeviceArray(f)
    s = T.Solver(m, precision="mixed")
    s.set_x(np.full(m.n, 1e9))  # so that the lifted field is kept
    used, _, _ = s.set_x_lifted(d.ptr, 20, 24, 16, zc, zf, scale=1.7, passes=0)
    assert used
    got = s.get_x().reshape(m.nz, n, n)
    ref = _lift_np(f, n, n, m.nz, zc, zf, 1.7)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    s.close()


@pytest.mark.parametrize("bad_start", [False, True])
def test_learned_start_async_matches_sync(bad_start):
    """th_solver_set_x_learned (forward pass, lift and fallback queued on the
    solver's stream, judged on the device) against the synchronous
    predict_device_input + set_x_lifted on the same input: the same judgement
    and starting residuals, then the same iterations and field bit for bit;
    once with the plain start x = 0 (the lifted guess of an untrained
    network loses) and once with a plain start so bad that it is kept."""
    n = 64
    m = W.build(TS.sample_spec(9, 7004), n)
    zf = TS.plane_centres(W.config0_layers())
    sur = T.Surrogate(n, n, seed=11)
    feats = DeviceArray(features_batch(1, n, 8)[0])
    out = []
    for mode in ("sync", "async"):
        s = T.Solver(m, precision="mixed", true_residual=True)
        if bad_start:
            s.set_x(np.full(m.n, 1e9))
        else:
            s.set_x_zero()
        if mode == "sync":
            d = sur.predict_device_input(feats.ptr)
            stats = s.set_x_lifted(d, n, n, 16, TS.COARSE_Z, zf, scale=T.Surrogate.T_SCALE, passes=1)
        else:
            s.set_x_learned(sur, feats.ptr, TS.COARSE_Z, zf, scale=T.Surrogate.T_SCALE, passes=1)
        s.set_rhs_steady()
        st = s.solve(1e-10, 200)
        if mode == "async":
            stats = s.lifted_stats()
        out.append((stats, st["iters"], st["converged"], s.get_x()))
        s.close()
    (sa, ia, ca, xa), (sb, ib, cb, xb) = out
    assert sa == sb and sa[0] == bad_start, (sa, sb)
    assert ia == ib and ca and cb
    assert np.array_equal(xa, xb)
    sur.close()
