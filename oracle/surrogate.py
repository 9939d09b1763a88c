"""CPU oracle of the thermal surrogate network (test infrastructure only).

Restates in numpy the U-Net whose forward, backward and training kernels
the library runs on the GPU (cuda-thermal_amd/csrc/surrogate.hip), so the
tests can check the GPU's outputs and gradients. The reference has no
learned part: this network is the one the benchmark description asks for
(4 down/up levels, 32-64 channels, BF16 weights and activations with FP32
accumulation), so parity here is against this restatement, itself pinned by
finite differences and by a direct-loop convolution in tests/test_surrogate_oracle.py.

Conventions (the same as the library's parameter layout):
  * activations NHWC; the input has 16 channels of which 3 are used
    (power density of die 0 and die 1 in MW/m^2, TSV copper fraction);
  * conv weights W[cout][tap][cin] with tap = ky * ks + kx, followed by the
    bias[cout]; layers in LAYERS order, concatenated into one flat vector;
  * conv: Y[y, x, co] = b[co] + sum W[co, ky*ks+kx, ci] X[y+ky-ks//2, x+kx-ks//2, ci]
    with zero padding; ReLU after every layer but the 1x1 head;
  * BF16 rounding (nearest even) of weights, of the network input and of
    every stored activation (conv outputs, pool outputs, up+skip sums); the
    head's output stays FP32.
"""
import numpy as np

# (name, level, cin, cout, kernel size, relu)
LAYERS = [
    ("e0a", 0, 16, 32, 3, True),
    ("e0b", 0, 32, 32, 3, True),
    ("e1a", 1, 32, 32, 3, True),
    ("e2a", 2, 32, 64, 3, True),
    ("e2b", 2, 64, 64, 3, True),
    ("e3a", 3, 64, 64, 3, True),
    ("b0", 4, 64, 64, 3, True),
    ("d3", 3, 64, 64, 3, True),
    ("d2", 2, 64, 32, 3, True),
    ("d1", 1, 32, 32, 3, True),
    ("d0", 0, 32, 32, 3, True),
    ("head", 0, 32, 16, 1, False),
]
N_IN, N_OUT = 3, 10  # used input / output channels


def param_count():
    return sum(co * ks * ks * ci + co for _, _, ci, co, ks, _ in LAYERS)


def unpack(flat):
    """Flat parameter vector -> {name: (W[cout][tap][cin], b[cout])}."""
    out, o = {}, 0
    for name, _, ci, co, ks, _ in LAYERS:
        nw = co * ks * ks * ci
        out[name] = (flat[o:o + nw].reshape(co, ks * ks, ci), flat[o + nw:o + nw + co])
        o += nw + co
    return out


def pack(grads):
    return np.concatenate([np.concatenate([grads[n][0].ravel(), grads[n][1].ravel()]) for n, *_ in LAYERS])


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def _shift(x, dy, dx):
    """out[b, y, x] = x[b, y + dy, x + dx] (zero outside)."""
    B, H, W, C = x.shape
    out = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    out[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


def conv(x, W, b, ks):
    """x [B,H,W,cin] -> [B,H,W,cout], accumulated in float64."""
    x = x.astype(np.float64)
    r = ks // 2
    y = np.zeros(x.shape[:3] + (W.shape[0],))
    for t in range(ks * ks):
        ky, kx = divmod(t, ks)
        y += _shift(x, ky - r, kx - r) @ W[:, t, :].T.astype(np.float64)
    return y + b


def conv_backward(x, W, gz, ks, need_dx=True):
    """Gradients of conv for dL/dY_pre = gz: (dW, db, dx)."""
    x = x.astype(np.float64)
    gz = gz.astype(np.float64)
    r = ks // 2
    dW = np.zeros(W.shape)
    dx = np.zeros(x.shape) if need_dx else None
    g2 = gz.reshape(-1, gz.shape[-1])
    for t in range(ks * ks):
        ky, kx = divmod(t, ks)
        xs = _shift(x, ky - r, kx - r)
        dW[:, t, :] = g2.T @ xs.reshape(-1, xs.shape[-1])
        if need_dx:
            dx += _shift(gz @ W[:, t, :].astype(np.float64), r - ky, r - kx)
    return dW, gz.sum(axis=(0, 1, 2)), dx


def pool(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_backward(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_backward(g):
    B, H, W, C = g.shape
    return g.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def forward(flat, x, round_bf16=True):
    """Network output [B,H,W,16] (FP32 head) and the stored activations."""
    P = unpack(np.asarray(flat, np.float32 if round_bf16 else np.float64))
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    L = {n: (q(W), b.astype(np.float64), ks, relu) for (n, _, _, _, ks, relu), (W, b) in
         zip(LAYERS, [P[n[0]] for n in LAYERS])}

    def lay(name, a):
        W, b, ks, relu = L[name]
        y = conv(a, W, b, ks)
        if relu:
            y = np.maximum(y, 0.0)
        return y if name == "head" else q(y)

    A = {"x": q(x)}
    A["a0"] = lay("e0a", A["x"])
    A["s0"] = lay("e0b", A["a0"])
    A["p0"] = q(pool(A["s0"]))
    A["s1"] = lay("e1a", A["p0"])
    A["p1"] = q(pool(A["s1"]))
    A["a2"] = lay("e2a", A["p1"])
    A["s2"] = lay("e2b", A["a2"])
    A["p2"] = q(pool(A["s2"]))
    A["s3"] = lay("e3a", A["p2"])
    A["p3"] = q(pool(A["s3"]))
    A["bb"] = lay("b0", A["p3"])
    A["u3"] = q(up(A["bb"]) + A["s3"])
    A["c3"] = lay("d3", A["u3"])
    A["u2"] = q(up(A["c3"]) + A["s2"])
    A["c2"] = lay("d2", A["u2"])
    A["u1"] = q(up(A["c2"]) + A["s1"])
    A["c1"] = lay("d1", A["u1"])
    A["u0"] = q(up(A["c1"]) + A["s0"])
    A["c0"] = lay("d0", A["u0"])
    A["out"] = lay("head", A["c0"])
    A["_L"] = L
    return A["out"], A


def loss_and_grad(flat, x, tgt, round_bf16=True):
    """MSE over the N_OUT used channels and its gradient (float64) with
    respect to the flat parameters, through the (rounded) forward pass."""
    out, A = forward(flat, x, round_bf16)
    L = A["_L"]
    B, H, W, _ = out.shape
    cnt = B * H * W * N_OUT
    d = out[..., :N_OUT] - tgt
    loss = float((d * d).sum() / cnt)
    g = np.zeros_like(out)
    g[..., :N_OUT] = 2.0 * d / cnt
    grads = {}

    def back(name, a_in, y, gy, need_dx=True):
        Wq, _, ks, relu = L[name]
        gz = gy * (y > 0) if relu else gy
        dW, db, dx = conv_backward(a_in, Wq, gz, ks, need_dx)
        grads[name] = (dW, db)
        return dx

    g_c0 = back("head", A["c0"], A["out"], g)
    g_u0 = back("d0", A["u0"], A["c0"], g_c0)
    g_u1 = back("d1", A["u1"], A["c1"], up_backward(g_u0))
    g_u2 = back("d2", A["u2"], A["c2"], up_backward(g_u1))
    g_u3 = back("d3", A["u3"], A["c3"], up_backward(g_u2))
    g_p3 = back("b0", A["p3"], A["bb"], up_backward(g_u3))
    g_p2 = back("e3a", A["p2"], A["s3"], pool_backward(g_p3) + g_u3)
    g_a2 = back("e2b", A["a2"], A["s2"], pool_backward(g_p2) + g_u2)
    g_p1 = back("e2a", A["p1"], A["a2"], g_a2)
    g_p0 = back("e1a", A["p0"], A["s1"], pool_backward(g_p1) + g_u1)
    g_a0 = back("e0b", A["a0"], A["s0"], pool_backward(g_p0) + g_u0)
    back("e0a", A["x"], A["a0"], g_a0, need_dx=False)
    return loss, pack(grads)
