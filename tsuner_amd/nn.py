"""The correction network on the GPU: forward, backward and Adam through the
library's BF16 MFMA kernels (tsu_nn_* in include/tsuner.h).

torch provides device memory and the stream only; every arithmetic step runs
in libtsuner.  Layout: activations NHWC BF16, weights FP32 master copies in one
flat vector (per layer [9 taps][out][in] then the bias), packed to BF16 for
the kernels after every update.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check, lib

CIN, WIDTH, COUT = 8, 32, 4
# (rows held, input channels, rows with gradients) of the four weight tensors
LAYERS = [(WIDTH, CIN), (WIDTH, WIDTH), (WIDTH, WIDTH), (16, WIDTH)]
WGRAD_KIND = [5, 6, 6, 7]      # TSU_NN_WGRAD_IN, _MID, _MID, _OUT


def layout():
    """Offsets (w, b) of each layer in the flat parameter vector and its size."""
    offs, o = [], 0
    for m, k in LAYERS:
        w = o
        o += 9 * m * k
        b = o
        o += m
        offs.append((w, b))
    return offs, o


OFFS, NPARAM = layout()


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights over the real fan-in, zero biases; the output layer's
    unused rows 4..15 are zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for li, ((m, k), (wo, _)) in enumerate(zip(LAYERS, OFFS)):
        w = rng.normal(0.0, math.sqrt(2.0 / (9 * k)), size=(9, m, k)).astype(np.float32)
        if li == 3:
            w[:, COUT:, :] = 0.0
            w *= 0.1           # start near the identity correction (zero output)
        p[wo:wo + w.size] = w.ravel()
    return p


class Net:
    """Parameters, optimiser state and activation buffers for batches of up to
    `max_batch` images of h x w pixels."""

    def __init__(self, params: np.ndarray = None, max_batch: int = 8, h: int = 256, w: int = 256,
                 device: int = 0):
        import torch
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.h, self.w, self.max_batch = h, w, max_batch
        p = init_params() if params is None else np.asarray(params, np.float32)
        assert p.size == NPARAM
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.p = torch.from_numpy(p.copy()).to(self.dev)
        self.g = torch.zeros(NPARAM, **f32)
        self.m = torch.zeros(NPARAM, **f32)
        self.v = torch.zeros(NPARAM, **f32)
        self.t = 0
        bf = dict(dtype=torch.bfloat16, device=self.dev)
        self.wb = [torch.zeros(9 * m * k, **bf) for m, k in LAYERS]
        # per-tap transposes used by the data gradients of layers 2, 3, 4
        self.wt = [None, torch.zeros(9 * WIDTH * WIDTH, **bf), torch.zeros(9 * WIDTH * WIDTH, **bf),
                   torch.zeros(9 * WIDTH * 8, **bf)]
        npix = max_batch * h * w
        self.acts = [torch.zeros(npix * WIDTH, **bf) for _ in range(3)]
        self.out = torch.zeros(npix * COUT, **f32)
        self.dy4 = torch.zeros(npix * 8, **bf)
        self.dbuf = [torch.zeros(npix * WIDTH, **bf) for _ in range(2)]
        self.ws = torch.zeros(max(lib().tsu_nn_wgrad_workspace(k) for k in (5, 6, 7)), **f32)
        self.nblocks = 1024
        self.loss_part = torch.zeros(self.nblocks, dtype=torch.float64, device=self.dev)
        self.pack()

    # ---- parameters ------------------------------------------------------
    def pack(self):
        L, s = lib(), _stream()
        for li, (m, k) in enumerate(LAYERS):
            wo, _ = OFFS[li]
            wp = C.c_void_p(self.p.data_ptr() + 4 * wo)
            wt = _ptr(self.wt[li]) if self.wt[li] is not None else None
            mt = 8 if li == 3 else m
            check(L.tsu_nn_pack(wp, _ptr(self.wb[li]), wt, m, k, mt, s), "tsu_nn_pack")

    def params(self) -> np.ndarray:
        return self.p.detach().cpu().numpy().copy()

    def _bias(self, li):
        return C.c_void_p(self.p.data_ptr() + 4 * OFFS[li][1])

    # ---- forward / backward ----------------------------------------------
    def forward(self, x, n: int, out=None):
        """x: BF16 [n][h][w][8] (a torch tensor or raw device address); the
        FP32 [n][h][w][4] output goes to `out` (tensor or address), by default
        the network's own buffer, which is returned."""
        assert 1 <= n <= self.max_batch
        L, s = lib(), _stream()
        xin = C.c_void_p(x if isinstance(x, int) else x.data_ptr())
        dst = self.out if out is None else out
        dst_addr = dst if isinstance(dst, int) else dst.data_ptr()
        h, w = self.h, self.w
        check(L.tsu_nn_conv(0, xin, _ptr(self.wb[0]), self._bias(0), None, _ptr(self.acts[0]), n, h, w, s), "conv1")
        check(L.tsu_nn_conv(1, _ptr(self.acts[0]), _ptr(self.wb[1]), self._bias(1), None, _ptr(self.acts[1]),
                            n, h, w, s), "conv2")
        check(L.tsu_nn_conv(1, _ptr(self.acts[1]), _ptr(self.wb[2]), self._bias(2), None, _ptr(self.acts[2]),
                            n, h, w, s), "conv3")
        check(L.tsu_nn_conv(2, _ptr(self.acts[2]), _ptr(self.wb[3]), self._bias(3), None, C.c_void_p(dst_addr),
                            n, h, w, s), "conv4")
        return dst

    def backward(self, x, target, n: int):
        """Gradients of the mean-squared error into self.g; returns the loss
        partial-sum buffer (sum of squared errors, not yet divided)."""
        L, s = lib(), _stream()
        h, w = self.h, self.w
        npix = n * h * w
        xin = C.c_void_p(x if isinstance(x, int) else x.data_ptr())
        tin = C.c_void_p(target if isinstance(target, int) else target.data_ptr())
        check(L.tsu_nn_mse(_ptr(self.out), tin, _ptr(self.dy4), _ptr(self.loss_part), self.nblocks, npix,
                           C.c_float(2.0 / (4.0 * npix)), s), "mse")
        ws, wsn = _ptr(self.ws), self.ws.numel()
        inputs = [xin, _ptr(self.acts[0]), _ptr(self.acts[1]), _ptr(self.acts[2])]
        dy = _ptr(self.dy4)
        for li in (3, 2, 1, 0):
            wo, bo = OFFS[li]
            m = LAYERS[li][0]
            check(L.tsu_nn_wgrad(WGRAD_KIND[li], inputs[li], dy, C.c_void_p(self.g.data_ptr() + 4 * wo),
                                 C.c_void_p(self.g.data_ptr() + 4 * bo), m, n, h, w, ws, wsn, s), "wgrad%d" % li)
            if li > 0:
                dst = self.dbuf[li % 2]
                kind = 3 if li == 3 else 4
                check(L.tsu_nn_conv(kind, dy, _ptr(self.wt[li]), None, inputs[li], _ptr(dst), n, h, w, s),
                      "dgrad%d" % li)
                dy = _ptr(dst)
        return self.loss_part

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.t += 1
        lr_t = lr * math.sqrt(1.0 - b2 ** self.t) / (1.0 - b1 ** self.t)
        check(lib().tsu_nn_adam(_ptr(self.p), _ptr(self.g), _ptr(self.m), _ptr(self.v), NPARAM, C.c_float(lr_t),
                                C.c_float(b1), C.c_float(b2), C.c_float(eps), _stream()), "adam")
        self.pack()

    def train_step(self, x, target, n: int, lr: float) -> None:
        self.forward(x, n)
        self.backward(x, target, n)
        self.adam(lr)

    def loss(self, n: int) -> float:
        """Mean squared error of the last backward() call (synchronises)."""
        self.torch.cuda.current_stream().synchronize()
        return float(self.loss_part.cpu().numpy().sum()) / (4.0 * n * self.h * self.w)
