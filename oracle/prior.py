"""CPU oracle of the learned history-cost prior (test infrastructure only).

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this
module, and only to check the library.  The reference has no learned component,
so this oracle restates the network the library defines (apgx.h, "Learned
history-cost prior"): 15 3x3 convs in a 4-level encoder-decoder, average
pooling down, nearest upsampling plus skip up, masked mean-squared-error loss.

Everything is float64 numpy.  With ``bf16=True`` the operands of every
convolution (weights, activations, output gradients) are rounded to bfloat16
(round to nearest even from float32) exactly where the library's matrix-core
kernels round them, so that the library's fp32-accumulated results agree to
accumulation-order error; with ``bf16=False`` the oracle is the exact real
arithmetic model that finite differences check.

Pinning: forward convolution against a direct 7-loop definition; pooling and
upsampling against their adjoints (<A x, y> == <x, A^T y>); the full
backward pass against central finite differences (bf16 off).
"""
from __future__ import annotations

import numpy as np

C = 32
CIN = 16
COUT = 16
# (input tensor, output tensor, cin, cout, level, relu), forward order.
CONVS = [
    ("x", "e0a", CIN, C, 0, True), ("e0a", "e0", C, C, 0, True),
    ("p1", "e1a", C, C, 1, True), ("e1a", "e1", C, C, 1, True),
    ("p2", "e2a", C, C, 2, True), ("e2a", "e2", C, C, 2, True),
    ("p3", "e3a", C, C, 3, True), ("e3a", "e3", C, C, 3, True),
    ("u2", "d2a", C, C, 2, True), ("d2a", "d2", C, C, 2, True),
    ("u1", "d1a", C, C, 1, True), ("d1a", "d1", C, C, 1, True),
    ("u0", "d0a", C, C, 0, True), ("d0a", "d0", C, C, 0, True),
    ("d0", "out", C, COUT, 0, False),
]
OPS = [("conv", 0), ("conv", 1), ("pool", "e0", "p1"), ("conv", 2), ("conv", 3), ("pool", "e1", "p2"),
       ("conv", 4), ("conv", 5), ("pool", "e2", "p3"), ("conv", 6), ("conv", 7), ("up", "e3", "e2", "u2"),
       ("conv", 8), ("conv", 9), ("up", "d2", "e1", "u1"), ("conv", 10), ("conv", 11),
       ("up", "d1", "e0", "u0"), ("conv", 12), ("conv", 13), ("conv", 14)]


def param_layout():
    """[(weight offset, bias offset, cin, cout)] of the flat parameter vector."""
    out, off = [], 0
    for _, _, cin, cout, _, _ in CONVS:
        out.append((off, off + cout * 9 * cin, cin, cout))
        off += cout * 9 * cin + cout
    return out, off


def bf16(a):
    """float -> float32 -> bfloat16 (round to nearest even) -> float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    lay, _ = param_layout()
    ws, bs = [], []
    for w_off, b_off, cin, cout in lay:
        ws.append(np.asarray(params[w_off:b_off], dtype=np.float64).reshape(cout, 9, cin))
        bs.append(np.asarray(params[b_off:b_off + cout], dtype=np.float64))
    return ws, bs


def conv(x, w, b, round_ops):
    """x [n,h,w,cin], w [cout,9,cin] -> [n,h,w,cout] (3x3, zero pad 1)."""
    n, h, wd, _ = x.shape
    xr = bf16(x) if round_ops else x
    wr = bf16(w) if round_ops else w
    xp = np.pad(xr, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((n, h, wd, w.shape[0]))
    for t in range(9):
        ky, kx = divmod(t, 3)
        y += xp[:, ky:ky + h, kx:kx + wd, :] @ wr[:, t, :].T
    return y + b


def conv_direct(x, w, b):
    """The convolution's definition as explicit loops (tiny inputs only)."""
    n, h, wd, cin = x.shape
    cout = w.shape[0]
    y = np.zeros((n, h, wd, cout))
    for s in range(n):
        for yy in range(h):
            for xx in range(wd):
                for co in range(cout):
                    acc = b[co]
                    for ky in range(3):
                        for kx in range(3):
                            sy, sx = yy + ky - 1, xx + kx - 1
                            if 0 <= sy < h and 0 <= sx < wd:
                                for ci in range(cin):
                                    acc += x[s, sy, sx, ci] * w[co, ky * 3 + kx, ci]
                    y[s, yy, xx, co] = acc
    return y


def pool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_adjoint(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_adjoint(g):
    n, h, w, c = g.shape
    return g.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def to_nhwc(planar, cpad):
    """[n][c][h][w] -> [n][h][w][cpad] (zero padded channels)."""
    x = np.transpose(np.asarray(planar, dtype=np.float64), (0, 2, 3, 1))
    return np.concatenate([x, np.zeros(x.shape[:3] + (cpad - x.shape[3],))], axis=3)


def forward(params, inputs, round_ops=True):
    """inputs [n][2L][gh][gw] -> (outputs [n][L][gh][gw], activations)."""
    L = inputs.shape[1] // 2
    ws, bs = unpack(params)
    act = {"x": to_nhwc(inputs, CIN)}
    for op in OPS:
        if op[0] == "conv":
            i, o, _, _, _, relu = CONVS[op[1]]
            y = conv(act[i], ws[op[1]], bs[op[1]], round_ops)
            act[o] = np.maximum(y, 0.0) if relu else y
        elif op[0] == "pool":
            act[op[2]] = pool(act[op[1]])
        else:
            act[op[3]] = up(act[op[1]]) + act[op[2]]
    out = np.transpose(act["out"][..., :L], (0, 3, 1, 2))
    return out, act


def loss_and_gradients(params, inputs, targets, round_ops=True):
    """Masked MSE over the L used outputs and its gradient (flat layout)."""
    L = inputs.shape[1] // 2
    ws, bs = unpack(params)
    out, act = forward(params, inputs, round_ops)
    d = out - np.asarray(targets, dtype=np.float64)
    count = d.size
    loss = float((d * d).sum() / count)
    g = {"out": to_nhwc(2.0 * d / count, COUT)}
    gw = [None] * len(CONVS)
    gb = [None] * len(CONVS)
    for op in reversed(OPS):
        if op[0] == "conv":
            k = op[1]
            i, o, cin, cout, _, relu = CONVS[k]
            gy = g[o] * (act[o] > 0) if relu else g[o]
            gb[k] = gy.sum(axis=(0, 1, 2))
            gyr = bf16(gy) if round_ops else gy
            x = bf16(act[i]) if round_ops else act[i]
            wr = bf16(ws[k]) if round_ops else ws[k]
            n, h, wd, _ = x.shape
            xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
            gp = np.pad(gyr, ((0, 0), (1, 1), (1, 1), (0, 0)))
            dw = np.zeros((cout, 9, cin))
            gx = np.zeros(x.shape)
            for t in range(9):
                ky, kx = divmod(t, 3)
                xs = xp[:, ky:ky + h, kx:kx + wd, :].reshape(-1, cin)
                dw[:, t, :] = gyr.reshape(-1, cout).T @ xs
                gx += gp[:, 2 - ky:2 - ky + h, 2 - kx:2 - kx + wd, :] @ wr[:, t, :]
            gw[k] = dw
            if i != "x":
                g[i] = gx
        elif op[0] == "pool":
            g[op[1]] = g[op[1]] + pool_adjoint(g[op[2]])
        else:
            g[op[2]] = g[op[3]].copy()
            g[op[1]] = up_adjoint(g[op[3]])
    lay, total = param_layout()
    flat = np.zeros(total)
    for k, (w_off, b_off, cin, cout) in enumerate(lay):
        flat[w_off:b_off] = gw[k].reshape(-1)
        flat[b_off:b_off + cout] = gb[k]
    return loss, flat


def features(cost, pins, tile=64):
    """Input planes [2L][gh][gw]: pins per tile / 64 and blocked fraction
    (cells beyond the lattice edge count as blocked)."""
    L, H, W = cost.shape
    gh, gw = -(-H // tile), -(-W // tile)
    out = np.zeros((2 * L, gh, gw))
    pins = np.asarray(pins, dtype=np.int64)
    x, y, lay = pins % W, (pins // W) % H, pins // (W * H)
    np.add.at(out, (2 * lay, y // tile, x // tile), 1.0 / 64.0)
    padded = np.zeros((L, gh * tile, gw * tile), dtype=bool)
    padded[:] = True
    padded[:, :H, :W] = cost == 0
    out[1::2] = padded.reshape(L, gh, tile, gw, tile).mean(axis=(2, 4))
    return out


def targets(history, tile=64, unit=10.0):
    L, H, W = history.shape
    gh, gw = -(-H // tile), -(-W // tile)
    out = np.zeros((L, gh, gw))
    for ty in range(gh):
        for tx in range(gw):
            blk = history[:, ty * tile:(ty + 1) * tile, tx * tile:(tx + 1) * tile].astype(np.float64)
            out[:, ty, tx] = np.log1p(blk.mean(axis=(1, 2)) / unit)
    return out
