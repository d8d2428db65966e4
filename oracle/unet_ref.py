"""FP64 numpy restatement of the learned warm start's U-Net (test
infrastructure only: the checker of the GPU network, never the product).

Same architecture and parameter layout as the library (vgx_unet.hip):
14 3^3 convolutions without bias, ReLU after each, 2x2x2 average pooling,
nearest upsampling, decoder inputs [upsampled deeper output | skip], and a
1x1x1 head of 16 weights. Tensors are (Z, Y, X, C) arrays (voxel index
x + X (y + Y z), channels last). Parameters are layer by layer
[cout][27 taps][cin] with tap = (dz+1) 9 + (dy+1) 3 + (dx+1), then the head.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 48), (48, 48), (48, 64), (64, 64),
          (112, 48), (48, 48), (80, 32), (32, 32), (48, 16), (16, 16)]
IN_CHANNELS = 16
HEAD = 16
TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def param_count() -> int:
    return sum(co * 27 * ci for ci, co in LAYERS) + HEAD


def split(params: np.ndarray):
    """Per-layer weights (cout, 27, cin) and the head (16,)."""
    ws, o = [], 0
    for ci, co in LAYERS:
        k = co * 27 * ci
        ws.append(np.asarray(params[o:o + k], np.float64).reshape(co, 27, ci))
        o += k
    return ws, np.asarray(params[o:o + HEAD], np.float64)


def join(ws, head) -> np.ndarray:
    return np.concatenate([w.reshape(-1) for w in ws] + [np.asarray(head).reshape(-1)])


def _shift(x, dz, dy, dx):
    """x[z + dz, y + dy, x + dx] with zeros outside the grid."""
    p = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    Z, Y, X = x.shape[:3]
    return p[1 + dz:1 + dz + Z, 1 + dy:1 + dy + Y, 1 + dx:1 + dx + X]


def conv(x, w):
    out = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(x, dz, dy, dx) @ w[:, t, :].T
    return out


def conv_back(x, w, dz_):
    """Gradients of conv(x, w) given dL/d(out): (dL/dx, dL/dw)."""
    dx_ = np.zeros_like(x)
    dw = np.zeros_like(w)
    for t, (a, b, c) in enumerate(TAPS):
        sx = _shift(x, a, b, c)
        dw[:, t, :] = np.einsum("zyxo,zyxi->oi", dz_, sx)
        # out[v] gets x[v + off] w: the gradient of x[u] collects dz[u - off]
        dx_ += _shift(dz_ @ w[:, t, :], -a, -b, -c)
    return dx_, dw


def pool(x):
    Z, Y, X, C = x.shape
    return x.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).mean(axis=(1, 3, 5))


def pool_back(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_back(g):
    Z, Y, X, C = g.shape
    return g.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).sum(axis=(1, 3, 5))


def relu(x):
    return np.maximum(x, 0.0)


def make_input(field: np.ndarray, n: int, scale: float = 1.0) -> np.ndarray:
    """Channel 0 = field * scale, channel 1 = 1 at cells touching a wall."""
    x = np.zeros((n, n, n, IN_CHANNELS))
    x[..., 0] = np.asarray(field, np.float64).reshape(n, n, n) * scale
    wall = np.zeros((n, n, n))
    wall[0, :, :] = wall[-1, :, :] = wall[:, 0, :] = wall[:, -1, :] = wall[:, :, 0] = wall[:, :, -1] = 1
    x[..., 1] = wall
    return x


def forward(params, x0):
    """Returns the output (Z, Y, X) and the activations the backward pass needs."""
    w, head = split(params)
    a = {"x0": x0}
    a["a0"] = relu(conv(x0, w[0]))
    a["s0"] = relu(conv(a["a0"], w[1]))
    a["p0"] = pool(a["s0"])
    a["a1"] = relu(conv(a["p0"], w[2]))
    a["s1"] = relu(conv(a["a1"], w[3]))
    a["p1"] = pool(a["s1"])
    a["a2"] = relu(conv(a["p1"], w[4]))
    a["s2"] = relu(conv(a["a2"], w[5]))
    a["p2"] = pool(a["s2"])
    a["a3"] = relu(conv(a["p2"], w[6]))
    a["b3"] = relu(conv(a["a3"], w[7]))
    a["c2"] = np.concatenate([up(a["b3"]), a["s2"]], axis=-1)
    a["a4"] = relu(conv(a["c2"], w[8]))
    a["d2"] = relu(conv(a["a4"], w[9]))
    a["c1"] = np.concatenate([up(a["d2"]), a["s1"]], axis=-1)
    a["a5"] = relu(conv(a["c1"], w[10]))
    a["d1"] = relu(conv(a["a5"], w[11]))
    a["c0"] = np.concatenate([up(a["d1"]), a["s0"]], axis=-1)
    a["a6"] = relu(conv(a["c0"], w[12]))
    a["d0"] = relu(conv(a["a6"], w[13]))
    out = a["d0"] @ head
    return out, a


def loss_and_grad(params, x0, target):
    """Mean squared error over voxels and its gradient (flat, parameter order)."""
    w, head = split(params)
    out, a = forward(params, x0)
    target = np.asarray(target, np.float64).reshape(out.shape)
    diff = out - target
    loss = float(np.mean(diff * diff))
    dout = 2.0 * diff / diff.size
    g = [None] * len(LAYERS)
    ghead = np.einsum("zyx,zyxc->c", dout, a["d0"])
    gd0 = dout[..., None] * head

    def back(l, gy, y, x):
        dz_ = gy * (y > 0)
        dx_, g[l] = conv_back(x, w[l], dz_)
        return dx_

    ga6 = back(13, gd0, a["d0"], a["a6"])
    gc0 = back(12, ga6, a["a6"], a["c0"])
    gd1 = up_back(gc0[..., :32])
    ga5 = back(11, gd1, a["d1"], a["a5"])
    gc1 = back(10, ga5, a["a5"], a["c1"])
    gd2 = up_back(gc1[..., :48])
    ga4 = back(9, gd2, a["d2"], a["a4"])
    gc2 = back(8, ga4, a["a4"], a["c2"])
    gb3 = up_back(gc2[..., :64])
    ga3 = back(7, gb3, a["b3"], a["a3"])
    gp2 = back(6, ga3, a["a3"], a["p2"])
    gs2 = gc2[..., 64:] + pool_back(gp2)
    ga2 = back(5, gs2, a["s2"], a["a2"])
    gp1 = back(4, ga2, a["a2"], a["p1"])
    gs1 = gc1[..., 48:] + pool_back(gp1)
    ga1 = back(3, gs1, a["s1"], a["a1"])
    gp0 = back(2, ga1, a["a1"], a["p0"])
    gs0 = gc0[..., 32:] + pool_back(gp0)
    ga0 = back(1, gs0, a["s0"], a["a0"])
    back(0, ga0, a["a0"], a["x0"])
    return loss, join(g, ghead)


def adam(params, grads, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, Algorithm 1) at step count t >= 1."""
    m = beta1 * m + (1 - beta1) * grads
    v = beta2 * v + (1 - beta2) * grads * grads
    mhat = m / (1 - beta1 ** t)
    vhat = v / (1 - beta2 ** t)
    return params - lr * mhat / (np.sqrt(vhat) + eps), m, v
