"""Trains the surrogate on solver-generated data, on one GPU.

  python -m thermal_amd.train_surrogate [--samples 20000] [--steps 6000]
      [--batch 16] [--out weights/surrogate_256.npy]

Generates the samples with the library's MG-PCG (features and targets stay
in device memory), trains with the library's forward/backward/Adam kernels,
reports the held-out error and the forward pass's MFMA rate, and writes the
flat FP32 parameter vector as .npy.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import Surrogate
from . import surrogate as S

BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def synthetic(sur, count, seed, chunk=64):
    """Smooth random features and targets (no solver): for timing the
    training kernels alone."""
    import time

    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    n = sur.H
    k = np.exp(-0.5 * (np.fft.fftfreq(n)[:, None] ** 2 + np.fft.fftfreq(n)[None, :] ** 2) / 0.02 ** 2)
    smooth = lambda a: np.real(np.fft.ifft2(np.fft.fft2(a, axes=(0, 1)) * k[..., None], axes=(0, 1)))  # noqa: E731
    done = 0
    while done < count:
        b = min(chunk, count - done)
        x = np.stack([np.abs(smooth(rng.standard_normal((n, n, 3)))) for _ in range(b)]).astype(np.float32)
        t = np.stack([smooth(rng.standard_normal((n, n, Surrogate.N_OUT))) for _ in range(b)]).astype(np.float32)
        sur.add(x, t)
        done += b
    return dict(samples=done, seconds=time.perf_counter() - t0, mean_iters=0.0)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--val", type=int, default=256, help="held-out samples (last of the set)")
    ap.add_argument("--n-tsv", type=int, default=50)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=5e-5)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=S.WEIGHTS)
    ap.add_argument("--report", default=None, help="JSON summary path")
    ap.add_argument("--synthetic", action="store_true",
                    help="smooth random features and targets instead of solver data (kernel timing only)")
    a = ap.parse_args(argv)

    sur = Surrogate(S.GRID, S.GRID, a.seed)
    total = a.samples + a.val
    sur.reserve(total)
    log(f"surrogate: {Surrogate.param_count()} parameters; generating {total} samples")
    gen = synthetic(sur, total, a.seed) if a.synthetic else S.generate(sur, total, n_tsv=a.n_tsv, log=log)
    log(f"  generated {gen['samples']} samples in {gen['seconds']:.0f} s")
    n_train = a.samples
    mse0 = sur.eval(n_train, a.val)
    chunks = 10
    losses, t_train = [], 0.0
    per = max(1, a.steps // chunks)
    for c in range(chunks):
        frac0, frac1 = c / chunks, (c + 1) / chunks
        # cosine schedule over the whole run, one segment per chunk
        lr0 = a.lr_final + 0.5 * (a.lr - a.lr_final) * (1 + np.cos(np.pi * frac0))
        lr1 = a.lr_final + 0.5 * (a.lr - a.lr_final) * (1 + np.cos(np.pi * frac1))
        l, ms = sur.train(per, a.batch, lr0, lr1, seed=a.seed + c, n_train=n_train)
        losses.extend(l.tolist())
        t_train += ms
        val = sur.eval(n_train, a.val)
        log(f"  steps {(c + 1) * per}: train loss {np.mean(l[-50:]):.3e}, held-out mse {val:.3e}, "
            f"{ms / per:.1f} ms/step")
    val = sur.eval(n_train, a.val)
    fwd_ms, fwd_flop = sur.time_forward(a.batch, 10)
    tflops = fwd_flop / (fwd_ms * 1e-3) / 1e12
    rms_K = float(np.sqrt(val)) * Surrogate.T_SCALE
    rep = dict(params=Surrogate.param_count(), samples=n_train, held_out=a.val, steps=len(losses), batch=a.batch,
               data_seconds=gen["seconds"], train_seconds=t_train / 1e3, ms_per_step=t_train / len(losses),
               held_out_mse_untrained=mse0, held_out_mse=val, held_out_rms_K=rms_K,
               loss_first=float(np.mean(losses[:20])), loss_last=float(np.mean(losses[-50:])),
               forward_ms_batch=fwd_ms, forward_tflops=tflops, mfma_frac=tflops / BF16_PEAK_TFLOPS)
    log(json.dumps(rep))
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.save(a.out, sur.get_params())
    if a.report:
        with open(a.report, "w") as f:
            json.dump(dict(rep, losses=losses[::10]), f)
    return rep


if __name__ == "__main__":
    main()
