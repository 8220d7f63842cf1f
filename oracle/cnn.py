"""FP64 oracle of the warm-start network's layers (test infrastructure only).

3x3 convolution with zero padding ("same"), cross-correlation convention:

    y[b, co, i, j] = bias[co] + sum_{ci, u, v} w[co, ci, u, v] x[b, ci, i + u - 1, j + v - 1]

and its two gradients.  Plain numpy in float64, written from the definition
(no third-party convolution routine); the checks in tests/test_cnn_oracle.py
pin it against brute-force loops and finite differences.
"""
import numpy as np


def _pad(x):
    return np.pad(np.asarray(x, np.float64), ((0, 0), (0, 0), (1, 1), (1, 1)))


def conv3x3(x, w, bias=None, relu=False, res=None):
    """x [B, Cin, H, W], w [Cout, Cin, 3, 3] -> [B, Cout, H, W] (float64)."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    B, Cin, H, W = x.shape
    xp = _pad(x)
    y = np.zeros((B, w.shape[0], H, W))
    for u in range(3):
        for v in range(3):
            # [B, Cin, H, W] x [Cout, Cin] -> [B, Cout, H, W]
            y += np.einsum("bchw,oc->bohw", xp[:, :, u:u + H, v:v + W], w[:, :, u, v], optimize=True)
    if bias is not None:
        y += np.asarray(bias, np.float64)[None, :, None, None]
    if relu:
        y = np.maximum(y, 0.0)
    if res is not None:
        y += np.asarray(res, np.float64)
    return y


def conv3x3_bwd_data(dy, w):
    """dL/dx of conv3x3 for dL/dy = dy: [B, Cout, H, W] -> [B, Cin, H, W]."""
    dy = np.asarray(dy, np.float64)
    w = np.asarray(w, np.float64)
    B, Cout, H, W = dy.shape
    dx = np.zeros((B, w.shape[1], H + 2, W + 2))
    for u in range(3):
        for v in range(3):
            dx[:, :, u:u + H, v:v + W] += np.einsum("bohw,oc->bchw", dy, w[:, :, u, v], optimize=True)
    return dx[:, :, 1:H + 1, 1:W + 1]


def conv3x3_bwd_weights(x, dy):
    """dL/dw [Cout, Cin, 3, 3] and dL/dbias [Cout] for input x and dL/dy = dy."""
    x = np.asarray(x, np.float64)
    dy = np.asarray(dy, np.float64)
    B, Cin, H, W = x.shape
    xp = _pad(x)
    dw = np.zeros((dy.shape[1], Cin, 3, 3))
    for u in range(3):
        for v in range(3):
            dw[:, :, u, v] = np.einsum("bohw,bchw->oc", dy, xp[:, :, u:u + H, v:v + W], optimize=True)
    return dw, dy.sum(axis=(0, 2, 3))


def conv3x3_loops(x, w, bias):
    """Brute-force loops of the definition (small sizes only)."""
    x = np.asarray(x, np.float64)
    B, Cin, H, W = x.shape
    Cout = w.shape[0]
    y = np.zeros((B, Cout, H, W))
    for b in range(B):
        for o in range(Cout):
            for i in range(H):
                for j in range(W):
                    s = bias[o]
                    for c in range(Cin):
                        for u in range(3):
                            for v in range(3):
                                ii, jj = i + u - 1, j + v - 1
                                if 0 <= ii < H and 0 <= jj < W:
                                    s += w[o, c, u, v] * x[b, c, ii, jj]
                    y[b, o, i, j] = s
    return y


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def net_channels(channels):
    return [2] + [channels] * 5 + [2]


def net_unpack(params, channels):
    """Packed [W0 b0 W1 b1 ... W5 b5] -> [(W, b), ...]."""
    ch = net_channels(channels)
    p = np.asarray(params, np.float64)
    out, o = [], 0
    for l in range(6):
        nw = ch[l + 1] * ch[l] * 9
        out.append((p[o:o + nw].reshape(ch[l + 1], ch[l], 3, 3), p[o + nw:o + nw + ch[l + 1]]))
        o += nw + ch[l + 1]
    return out


def _ident(a):
    return np.asarray(a, np.float64)


def net_forward(params, x, channels, rounding=True):
    """x + net(x) for the warm-start network (6 convolutions, ReLU after the
    first five).  rounding: operands of every convolution rounded to
    bfloat16, as the matrix-core kernels do.  Returns (out, layer inputs)."""
    rd = bf16 if rounding else _ident
    x = np.asarray(x, np.float64)
    acts = [x]
    a = x
    layers = net_unpack(params, channels)
    for l, (W, b) in enumerate(layers):
        z = conv3x3(rd(a), rd(W), b)
        if l == len(layers) - 1:
            return z + x, acts
        a = np.maximum(z, 0.0)
        if rounding:
            a = a.astype(np.float32).astype(np.float64)  # the kernels store layer outputs in float32
        acts.append(a)


def net_grads(params, x, t, channels, rounding=True):
    """Gradient of mean((x + net(x) - t)^2) with respect to the packed
    parameters, and the loss, by backpropagation."""
    rd = bf16 if rounding else _ident
    out, acts = net_forward(params, x, channels, rounding)
    n = out.size
    g = 2.0 * (out - np.asarray(t, np.float64)) / n
    if rounding:
        g = g.astype(np.float32).astype(np.float64)
    layers = net_unpack(params, channels)
    grads = [None] * 6
    for l in range(5, -1, -1):
        W, _ = layers[l]
        dW = conv3x3_bwd_weights(rd(acts[l]), rd(g))[0]
        grads[l] = (dW, g.sum(axis=(0, 2, 3)))
        if l > 0:
            g = conv3x3_bwd_data(rd(g), rd(W)) * (acts[l] > 0)
            if rounding:
                g = g.astype(np.float32).astype(np.float64)
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return flat, float(np.mean((out - t) ** 2))
