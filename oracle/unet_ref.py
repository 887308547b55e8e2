"""CPU oracle of the learned denoiser (test infrastructure only).

A float64 numpy restatement of the network the HIP denoiser trains and runs:
a 5-level U-Net of 3x3 convolutions (zero padding, stride 1) with ReLU, 2x2
average pooling on the way down, nearest-neighbour 2x upsampling on the way
up, skip connections by channel concatenation (upsampled features first,
then the skip), and a final 3x3 convolution to 3 channels without ReLU.

Inputs (10 channels, NHWC): log(1 + radiance) RGB, albedo RGB, normal XYZ,
log(1 + depth).  Output: log(1 + radiance) RGB.  Loss: mean absolute error
in that log space over all pixels and the 3 channels.

There is no counterpart in the reference (its denoise stage calls a closed
library, denoiser.cpp); the network follows the stated design: widths
32/48/64/96/128.  Parity of the HIP kernels is pinned against this oracle,
and the oracle itself is pinned by finite differences (tests).

Parameter order (the flat vector the library exchanges): for every conv in
network order, W[tap][cout][cin] (tap = 3*(dy+1) + (dx+1)) then bias[cout],
using the real (unpadded) channel counts.
"""
import numpy as np

WIDTHS = (32, 48, 64, 96, 128)
IN_CH = 10
OUT_CH = 3


def layer_shapes(widths=WIDTHS, in_ch=IN_CH):
    """[(name, cin_a, cin_b, cout, relu)] in network order; cin_b is the skip part."""
    L = []
    c = in_ch
    for lvl, w in enumerate(widths):
        L.append((f"enc{lvl}a", c, 0, w, True))
        L.append((f"enc{lvl}b", w, 0, w, True))
        c = w
    for lvl in range(len(widths) - 2, -1, -1):
        w = widths[lvl]
        L.append((f"dec{lvl}a", c, w, w, True))
        L.append((f"dec{lvl}b", w, 0, w, True))
        c = w
    L.append(("out", c, 0, OUT_CH, False))
    return L


def num_params(widths=WIDTHS, in_ch=IN_CH):
    return sum(9 * (a + b) * o + o for _, a, b, o, _ in layer_shapes(widths, in_ch))


def unflatten(theta, widths=WIDTHS, in_ch=IN_CH):
    out, k = [], 0
    for name, a, b, o, relu in layer_shapes(widths, in_ch):
        n = 9 * (a + b) * o
        W = theta[k:k + n].reshape(9, o, a + b)
        k += n
        bias = theta[k:k + o]
        k += o
        out.append((name, W, bias, relu))
    assert k == theta.size
    return out


def init_params(seed=1, widths=WIDTHS, in_ch=IN_CH):
    """He-normal weights, zero bias (float32-representable values)."""
    rng = np.random.default_rng(seed)
    parts = []
    for _, a, b, o, relu in layer_shapes(widths, in_ch):
        fan = 9 * (a + b)
        std = np.sqrt(2.0 / fan) if relu else np.sqrt(1.0 / fan)
        parts.append((rng.standard_normal(fan * o) * std).astype(np.float32).astype(np.float64))
        parts.append(np.zeros(o))
    return np.concatenate(parts)


def _shift(x, dy, dx):
    """y[n, i, j] = x[n, i + dy, j + dx] with zeros outside."""
    N, H, W, C = x.shape
    y = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    y[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return y


def conv(x, W, b):
    """3x3 convolution, W[tap][cout][cin]."""
    y = np.zeros(x.shape[:3] + (W.shape[1],))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        y += _shift(x, dy, dx) @ W[t].T
    return y + b


def conv_backward(x, W, dy):
    """Gradients of conv w.r.t. input, weights, bias."""
    dx = np.zeros_like(x)
    dW = np.zeros_like(W)
    for t in range(9):
        ddy, ddx = t // 3 - 1, t % 3 - 1
        xs = _shift(x, ddy, ddx)
        dW[t] = np.einsum("nhwo,nhwi->oi", dy, xs)
        dx += _shift(dy @ W[t], -ddy, -ddx)
    return dx, dW, dy.sum((0, 1, 2))


def pool(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean((2, 4))


def pool_backward(dy):
    return np.repeat(np.repeat(dy, 2, 1), 2, 2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, 1), 2, 2)


def up_backward(dy):
    N, H, W, C = dy.shape
    return dy.reshape(N, H // 2, 2, W // 2, 2, C).sum((2, 4))


def features(radiance, albedo, normal, depth):
    """Network input from render outputs: radiance/albedo/normal (N, H, W, 3),
    depth (N, H, W)."""
    return np.concatenate([np.log1p(np.maximum(radiance, 0.0)), albedo, normal,
                           np.log1p(np.maximum(depth, 0.0))[..., None]], -1)


def forward(theta, x, widths=WIDTHS, keep=False):
    """x: (N, H, W, 10) -> (N, H, W, 3) prediction in log space."""
    layers = unflatten(theta, widths, x.shape[-1])
    acts = {}
    it = iter(layers)
    h = x
    skips = []
    nlev = len(widths)
    for lvl in range(nlev):
        if lvl:
            h = pool(h)
        for _ in range(2):
            name, W, b, relu = next(it)
            inp = h
            h = conv(inp, W, b)
            h = np.maximum(h, 0.0) if relu else h
            acts[name] = (inp, h)
        skips.append(h)
    for lvl in range(nlev - 2, -1, -1):
        name, W, b, relu = next(it)
        inp = np.concatenate([up(h), skips[lvl]], -1)
        h = np.maximum(conv(inp, W, b), 0.0)
        acts[name] = (inp, h)
        name, W, b, relu = next(it)
        inp = h
        h = np.maximum(conv(inp, W, b), 0.0)
        acts[name] = (inp, h)
    name, W, b, relu = next(it)
    acts[name] = (h, None)
    y = conv(h, W, b)
    return (y, acts, layers) if keep else y


def loss(pred, target_log):
    return np.abs(pred - target_log).mean()


def loss_and_grad(theta, x, target_log, widths=WIDTHS):
    """Mean L1 in log space and its gradient w.r.t. the flat parameters."""
    y, acts, layers = forward(theta, x, widths, keep=True)
    L = loss(y, target_log)
    g = np.sign(y - target_log) / y.size
    grads = {}
    by = {name: (W, b, relu) for name, W, b, relu in layers}
    nlev = len(widths)
    # output conv
    dx, dW, db = conv_backward(acts["out"][0], by["out"][0], g)
    grads["out"] = (dW, db)
    dh = dx
    dskip = [None] * nlev
    for lvl in range(nlev - 1):
        for name in (f"dec{lvl}b", f"dec{lvl}a"):
            inp, out = acts[name]
            dpre = dh * (out > 0)
            dx, dW, db = conv_backward(inp, by[name][0], dpre)
            grads[name] = (dW, db)
            dh = dx
        c_up = dh.shape[-1] - widths[lvl]
        dskip[lvl] = dh[..., c_up:]
        dh = up_backward(dh[..., :c_up])
    for lvl in range(nlev - 1, -1, -1):
        if dskip[lvl] is not None:
            dh = dh + dskip[lvl]
        for name in (f"enc{lvl}b", f"enc{lvl}a"):
            inp, out = acts[name]
            dpre = dh * (out > 0)
            dx, dW, db = conv_backward(inp, by[name][0], dpre)
            grads[name] = (dW, db)
            dh = dx
        if lvl:
            dh = pool_backward(dh)
    flat = []
    for name, W, b, relu in layers:
        dW, db = grads[name]
        flat += [dW.ravel(), db]
    return L, np.concatenate(flat)
