"""CPU oracle of the coarse-grid correction network (test infrastructure only).

numpy restatement of the 4-layer 3x3 convolutional network used by the
learned coarse-grid correction: 'same' zero padding, NHWC activations,
weights W[tap][co][ci] with tap = 3*(dy+1) + (dx+1), ReLU after the first
three layers, linear output, mean-squared-error loss.  Forward and backward
are written from the definitions (no framework); the backward pass is pinned
by finite differences in tests/test_nn_oracle.py.  The GPU kernels compute in
BF16 with FP32 accumulation; ``bf16`` below rounds to nearest-even BF16 so
the oracle can be fed the same rounded operands.

Only tests/ and bench.py's checks use this module.
"""
from __future__ import annotations

import numpy as np

TAPS = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).reshape(a.shape)


def shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """x[n, y+dy, x+dx, c] with zeros outside the image."""
    out = np.zeros_like(x)
    N, H, W, C = x.shape
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    out[:, ys:ye, xs:xe, :] = x[:, ys + dy:ye + dy, xs + dx:xe + dx, :]
    return out


def conv(x: np.ndarray, w: np.ndarray, b: np.ndarray = None) -> np.ndarray:
    """y[n,y,x,co] = sum_tap sum_ci x[n, y+dy, x+dx, ci] * w[tap, co, ci] (+ b[co])."""
    N, H, W, _ = x.shape
    co = w.shape[1]
    y = np.zeros((N, H, W, co), dtype=np.float64)
    for t, (dy, dx) in enumerate(TAPS):
        y += np.einsum("nhwi,oi->nhwo", shift(x, dy, dx).astype(np.float64), w[t].astype(np.float64))
    if b is not None:
        y += b.astype(np.float64)
    return y


def conv_input_grad(dy: np.ndarray, w: np.ndarray) -> np.ndarray:
    """dL/dx of conv: dx[n,y,x,ci] = sum_tap sum_co dy[n, y-dy, x-dx, co] w[tap, co, ci]."""
    N, H, W, _ = dy.shape
    ci = w.shape[2]
    dx_ = np.zeros((N, H, W, ci), dtype=np.float64)
    for t, (ddy, ddx) in enumerate(TAPS):
        dx_ += np.einsum("nhwo,oi->nhwi", shift(dy, -ddy, -ddx).astype(np.float64), w[t].astype(np.float64))
    return dx_


def conv_weight_grad(x: np.ndarray, dy: np.ndarray):
    """dL/dw[tap,co,ci] = sum_{n,y,x} dy[n,y,x,co] x[n,y+dy,x+dx,ci]; dL/db[co] = sum dy."""
    co, ci = dy.shape[3], x.shape[3]
    dw = np.zeros((9, co, ci), dtype=np.float64)
    for t, (ddy, ddx) in enumerate(TAPS):
        dw[t] = np.einsum("nhwo,nhwi->oi", dy.astype(np.float64), shift(x, ddy, ddx).astype(np.float64))
    db = dy.astype(np.float64).sum(axis=(0, 1, 2))
    return dw, db


def forward(x: np.ndarray, params, rnd=None):
    """params: list of (w, b) for the four layers; returns (output, activations).
    rnd, if given, rounds every stored hidden activation (the GPU stores them
    in BF16)."""
    acts = [x]
    h = x
    for k, (w, b) in enumerate(params):
        z = conv(h, w, b)
        if k < len(params) - 1:
            h = np.maximum(z, 0.0)
            if rnd is not None:
                h = rnd(h)
        else:
            h = z
        acts.append(h)
    return h, acts


def loss_and_grads(x: np.ndarray, target: np.ndarray, params, rnd=None):
    """Mean-squared error over all outputs and its parameter gradients.  rnd,
    if given, also rounds the back-propagated signals where the GPU stores
    them (the output gradient and each masked data gradient)."""
    out, acts = forward(x, params, rnd)
    diff = out - target
    loss = float(np.mean(diff * diff))
    d = 2.0 * diff / diff.size
    if rnd is not None:
        d = rnd(d)
    grads = [None] * len(params)
    for k in range(len(params) - 1, -1, -1):
        w, _ = params[k]
        dw, db = conv_weight_grad(acts[k], d)
        grads[k] = (dw, db)
        if k > 0:
            d = conv_input_grad(d, w) * (acts[k] > 0)
            if rnd is not None:
                d = rnd(d)
    return loss, grads


def bf16_64(a: np.ndarray) -> np.ndarray:
    """bf16() for float64 arrays (through float32, as the GPU rounds its FP32
    results), returned as float64."""
    return bf16(np.asarray(a, dtype=np.float32)).astype(np.float64)


def init_params(rng, cin: int = 8, width: int = 32, cout: int = 4):
    """He-normal init; layer k maps cin -> width -> width -> width -> cout."""
    dims = [(cin, width), (width, width), (width, width), (width, cout)]
    ps = []
    for ci, co in dims:
        w = rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), size=(9, co, ci)).astype(np.float32)
        b = np.zeros(co, dtype=np.float32)
        ps.append((w, b))
    return ps
