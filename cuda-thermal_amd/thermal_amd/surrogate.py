"""The learned warm start: features, training data from the solver, training
and the lifted initial guess.

The network (csrc/surrogate.hip) sees a 256 x 256 picture of a two-tier
stack of the reference's default layers (stack3d.h:153-162): the power
density of each die (MW/m^2) and the TSV copper fraction, and predicts the
temperature rise of each of the stack's 10 planes (K / 50). Training data
are steady solutions of seeded random floorplans with TSV clusters, solved
by this library's MG-PCG; training runs the library's own forward, backward
and Adam kernels on the GPU.

At use, the prediction is lifted to the solver's grid (bilinear over the
die, linear in z between plane centres, then red-black column line
relaxations, each a tridiagonal solve per column) and kept as the initial
guess only if its starting residual ||P - A T|| is below the plain warm
start's (th_solver_set_x_lifted).
"""
import math
import os
import time

import numpy as np

from . import Model, Solver, Surrogate
from . import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(os.path.dirname(HERE), "weights", "surrogate_256.npy")
GRID = 256
DIE_LAYERS = (0, 2)       # Si tier 2 (bottom) and Si tier 1 (top)
TSV_LAYERS = (0, 1, 2)    # the layers TSV clusters cross


def plane_centres(layers):
    """z of each plane's centre (m), bottom to top."""
    z, out = 0.0, []
    for layer in layers:
        t, planes = layer[1], layer[4]
        out.extend(z + (q + 0.5) * t / planes for q in range(planes))
        z += t
    return np.array(out)


COARSE_Z = plane_centres(W.config0_layers())


def _overlap_1d(x0, w, n, L):
    """Fraction of [x0, x0 + w] in each of n cells over [0, L]."""
    e = np.linspace(0.0, L, n + 1)
    return np.clip(np.minimum(e[1:], x0 + w) - np.maximum(e[:-1], x0), 0.0, None) / w


def power_planes(spec, n, layers=None):
    """Power per cell (W) on an n x n grid, [nz][n][n], of a spec's blocks and
    uniform power, each spread over its area and over its layer's planes
    (the rule of stack3d.h:196-228)."""
    layers = layers or spec["layers"]
    first = np.cumsum([0] + [l[4] for l in layers])
    Lx, Ly = spec["Lx"], spec["Ly"]
    maps = {}
    for layer, watts in spec["uniform"]:
        maps[layer] = maps.get(layer, 0.0) + watts / (n * n)
    for b in spec["blocks"]:
        layer, x0, y0, w, h, watts = b[1:7]
        m = maps.get(layer, 0.0) + watts * np.outer(_overlap_1d(y0, h, n, Ly), _overlap_1d(x0, w, n, Lx))
        maps[layer] = m
    P = np.zeros((first[-1], n, n))
    for layer, m in maps.items():
        P[first[layer]:first[layer + 1]] = m / layers[layer][4]
    return P


def tsv_fraction(spec, n):
    """Copper fraction per column, averaged over the TSV layers' extent: the
    area fraction of the column covered by TSV regions (capped at 1 per
    layer, stack3d.h:210-221)."""
    f = np.zeros((len(spec["layers"]), n, n))
    for layer, x0, y0, w, h, frac in spec["tsvs"]:
        f[layer] += frac * np.outer(_overlap_1d(y0, h, n, spec["Ly"]) * h * n / spec["Ly"],
                                    _overlap_1d(x0, w, n, spec["Lx"]) * w * n / spec["Lx"])
    f = np.minimum(f, 1.0)
    return np.mean([f[k] for k in TSV_LAYERS], axis=0)


def features(spec, n=GRID, P=None, tsv=None):
    """Network input [n][n][3]: die power densities (MW/m^2), TSV fraction.
    P (power_planes) and tsv (tsv_fraction) may be passed precomputed."""
    P = power_planes(spec, n) if P is None else P
    first = np.cumsum([0] + [l[4] for l in spec["layers"]])
    area = (spec["Lx"] / n) * (spec["Ly"] / n)
    x = np.empty((n, n, 3), np.float32)
    for c, layer in enumerate(DIE_LAYERS):
        x[..., c] = P[first[layer]:first[layer + 1]].sum(axis=0) / area / 1e6
    x[..., 2] = tsv_fraction(spec, n) if tsv is None else tsv
    return x


def sample_spec(seed, tsv_seed):
    """A config-2-style floorplan (12 random blocks per die on a 10%
    background, 25 W per die) on the 10-plane default stack, with the TSV
    clusters of `tsv_seed`."""
    spec = W.config2_spec(seed, tsv=False)
    spec["layers"] = W.config0_layers()
    tsv = W.config2_spec(tsv_seed, tsv=True)["tsvs"]
    spec["tsvs"] = tsv
    return spec


def generate(sur, count, n_tsv=50, seed0=100000, chunk=128, log=print, precision="mixed"):
    """Adds `count` solved samples to the surrogate's device dataset:
    n_tsv TSV layouts (one solver each), count / n_tsv floorplans each;
    steady solves to 1e-8 warm-started from the previous floorplan."""
    n = sur.H
    per = int(math.ceil(count / n_tsv))
    xs, ts, done, t0 = [], [], 0, time.perf_counter()
    iters = []
    for k in range(n_tsv):
        tsv_seed = 7000 + 2 * k
        base = sample_spec(seed0 + k * per, tsv_seed)
        m = W.build(base, n)
        s = Solver(m, precision=precision)
        s.set_x_zero()
        tsv = tsv_fraction(base, n)
        for q in range(per):
            if done >= count:
                break
            spec = sample_spec(seed0 + k * per + q, tsv_seed)
            P = power_planes(spec, n)
            s.set_power(P.ravel())
            s.set_rhs_steady()
            st = s.solve(1e-8, 500)
            iters.append(st["iters"])
            T = s.get_x().reshape(-1, n, n)
            xs.append(features(spec, n, P, tsv))
            ts.append((np.moveaxis(T, 0, -1) / Surrogate.T_SCALE).astype(np.float32))
            done += 1
            if len(xs) == chunk:
                sur.add(np.stack(xs), np.stack(ts))
                xs, ts = [], []
        s.close()
        if log and (k + 1) % 10 == 0:
            log(f"  data: {done}/{count} samples, {time.perf_counter() - t0:.0f} s, "
                f"mean {np.mean(iters):.1f} iterations")
    if xs:
        sur.add(np.stack(xs), np.stack(ts))
    return dict(samples=done, seconds=time.perf_counter() - t0, mean_iters=float(np.mean(iters)))


def load(path=WEIGHTS, seed=1):
    """A surrogate with the trained weights (raises if the file is absent)."""
    p = np.load(path, allow_pickle=False)
    sur = Surrogate(GRID, GRID, seed)
    sur.set_params(p)
    return sur


def warm_start(solver, sur, spec, fine_layers, passes=2):
    """Predicts the stack's field from the spec's floorplan and sets it as
    the solver's initial guess unless the plain warm start has the smaller
    starting residual. Returns (used, relres_plain, relres_lift, ms)."""
    x = features(spec, sur.H)[None]
    t0 = time.perf_counter()
    d = sur.predict_device(x[0])
    used, rp, rl = solver.set_x_lifted(d, sur.H, sur.W, 16, COARSE_Z, plane_centres(fine_layers),
                                       scale=Surrogate.T_SCALE, passes=passes)
    return used, rp, rl, (time.perf_counter() - t0) * 1e3
