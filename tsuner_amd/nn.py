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

CIN, COUT = 8, 4
W64 = 16                       # TSU_NN_W64: the 64-channel kernel variants
# kernel kinds (TSU_NN_* in include/tsuner.h)
FWD_IN, FWD_MID, FWD_OUT, DGRAD_OUT, DGRAD_MID, WGRAD_IN, WGRAD_MID, WGRAD_OUT = range(8)


class Arch:
    """Layer shapes of a network 8 -> width (x mids + 1) -> 4: an input
    convolution, `mids` width -> width convolutions, an output convolution
    (weights held as 16 rows, 4 used), ReLU after all but the last."""

    def __init__(self, width: int = 32, mids: int = 2):
        if width not in (32, 64) or mids < 0:
            raise ValueError("width must be 32 or 64 and mids >= 0")
        self.width, self.mids = width, mids
        self.flag = W64 if width == 64 else 0
        # (rows held, input channels) of each weight tensor
        self.layers = [(width, CIN)] + [(width, width)] * mids + [(16, width)]
        self.offs, o = [], 0
        for m, k in self.layers:
            w = o
            o += 9 * m * k
            b = o
            o += m
            self.offs.append((w, b))
        self.nparam = o
        self.flop_per_pixel = 2 * 9 * (CIN * width + mids * width * width + width * COUT)

    @property
    def nlayers(self) -> int:
        return len(self.layers)

    def wgrad_kind(self, li: int) -> int:
        k = WGRAD_IN if li == 0 else (WGRAD_OUT if li == self.nlayers - 1 else WGRAD_MID)
        return k | self.flag

    def as_dict(self) -> dict:
        return {"width": self.width, "mids": self.mids}

    def __eq__(self, other):
        return isinstance(other, Arch) and self.as_dict() == other.as_dict()

    def __repr__(self):
        return "Arch(width=%d, mids=%d)" % (self.width, self.mids)


DEFAULT = Arch(32, 2)
# the default network's layer table (kept for callers of the fixed shape)
LAYERS, OFFS, NPARAM = DEFAULT.layers, DEFAULT.offs, DEFAULT.nparam
WIDTH = DEFAULT.width


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def init_params(seed: int = 0, arch: Arch = DEFAULT) -> np.ndarray:
    """He-normal weights over the real fan-in, zero biases; the output layer's
    unused rows 4..15 are zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(arch.nparam, np.float32)
    last = arch.nlayers - 1
    for li, ((m, k), (wo, _)) in enumerate(zip(arch.layers, arch.offs)):
        w = rng.normal(0.0, math.sqrt(2.0 / (9 * k)), size=(9, m, k)).astype(np.float32)
        if li == last:
            w[:, COUT:, :] = 0.0
            w *= 0.1           # start near the identity correction (zero output)
        p[wo:wo + w.size] = w.ravel()
    return p


class Net:
    """Parameters, optimiser state and activation buffers for batches of up to
    `max_batch` images of h x w pixels."""

    def __init__(self, params: np.ndarray = None, max_batch: int = 8, h: int = 256, w: int = 256,
                 device: int = 0, arch: Arch = DEFAULT):
        import torch
        self.torch = torch
        self.arch = arch
        self.dev = torch.device("cuda", device)
        self.h, self.w, self.max_batch = h, w, max_batch
        p = init_params(arch=arch) if params is None else np.asarray(params, np.float32)
        assert p.size == arch.nparam, "parameter vector does not match %r" % (arch,)
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.p = torch.from_numpy(p.copy()).to(self.dev)
        self.g = torch.zeros(arch.nparam, **f32)
        self.m = torch.zeros(arch.nparam, **f32)
        self.v = torch.zeros(arch.nparam, **f32)
        self.t = 0
        bf = dict(dtype=torch.bfloat16, device=self.dev)
        W, L = arch.width, arch.nlayers
        self.wb = [torch.zeros(9 * m * k, **bf) for m, k in arch.layers]
        # per-tap transposes used by the data gradients of layers 1 .. L-1
        self.wt = [None] + [torch.zeros(9 * W * W, **bf) for _ in range(arch.mids)] + [torch.zeros(9 * W * 8, **bf)]
        npix = max_batch * h * w
        self.acts = [torch.zeros(npix * W, **bf) for _ in range(L - 1)]
        self.out = torch.zeros(npix * COUT, **f32)
        self.dy4 = torch.zeros(npix * 8, **bf)
        self.dbuf = [torch.zeros(npix * W, **bf) for _ in range(2)]
        self.ws = torch.zeros(max(lib().tsu_nn_wgrad_workspace(arch.wgrad_kind(li)) for li in range(L)), **f32)
        self.nblocks = 1024
        self.loss_part = torch.zeros(self.nblocks, dtype=torch.float64, device=self.dev)
        self.pack()

    # ---- parameters ------------------------------------------------------
    def pack(self):
        L, s = lib(), _stream()
        last = self.arch.nlayers - 1
        for li, (m, k) in enumerate(self.arch.layers):
            wo, _ = self.arch.offs[li]
            wp = C.c_void_p(self.p.data_ptr() + 4 * wo)
            wt = _ptr(self.wt[li]) if self.wt[li] is not None else None
            mt = 8 if li == last else m
            check(L.tsu_nn_pack(wp, _ptr(self.wb[li]), wt, m, k, mt, s), "tsu_nn_pack")

    def params(self) -> np.ndarray:
        return self.p.detach().cpu().numpy().copy()

    def _bias(self, li):
        return C.c_void_p(self.p.data_ptr() + 4 * self.arch.offs[li][1])

    # ---- forward / backward ----------------------------------------------
    def forward(self, x, n: int, out=None):
        """x: BF16 [n][h][w][8] (a torch tensor or raw device address); the
        FP32 [n][h][w][4] output goes to `out` (tensor or address), by default
        the network's own buffer, which is returned."""
        assert 1 <= n <= self.max_batch
        L, s = lib(), _stream()
        f, last = self.arch.flag, self.arch.nlayers - 1
        xin = C.c_void_p(x if isinstance(x, int) else x.data_ptr())
        dst = self.out if out is None else out
        dst_addr = dst if isinstance(dst, int) else dst.data_ptr()
        h, w = self.h, self.w
        src = xin
        for li in range(last + 1):
            kind = (FWD_IN if li == 0 else FWD_OUT if li == last else FWD_MID) | f
            o = C.c_void_p(dst_addr) if li == last else _ptr(self.acts[li])
            check(L.tsu_nn_conv(kind, src, _ptr(self.wb[li]), self._bias(li), None, o, n, h, w, s), "conv%d" % li)
            src = o
        return dst

    def backward(self, x, target, n: int):
        """Gradients of the mean-squared error into self.g; returns the loss
        partial-sum buffer (sum of squared errors, not yet divided)."""
        L, s = lib(), _stream()
        A = self.arch
        h, w = self.h, self.w
        npix = n * h * w
        xin = C.c_void_p(x if isinstance(x, int) else x.data_ptr())
        tin = C.c_void_p(target if isinstance(target, int) else target.data_ptr())
        check(L.tsu_nn_mse(_ptr(self.out), tin, _ptr(self.dy4), _ptr(self.loss_part), self.nblocks, npix,
                           C.c_float(2.0 / (4.0 * npix)), s), "mse")
        ws, wsn = _ptr(self.ws), self.ws.numel()
        inputs = [xin] + [_ptr(a) for a in self.acts]
        dy = _ptr(self.dy4)
        last = A.nlayers - 1
        for li in range(last, -1, -1):
            wo, bo = A.offs[li]
            m = A.layers[li][0]
            check(L.tsu_nn_wgrad(A.wgrad_kind(li), inputs[li], dy, C.c_void_p(self.g.data_ptr() + 4 * wo),
                                 C.c_void_p(self.g.data_ptr() + 4 * bo), m, n, h, w, ws, wsn, s), "wgrad%d" % li)
            if li > 0:
                dst = self.dbuf[li % 2]
                kind = (DGRAD_OUT if li == last else DGRAD_MID) | A.flag
                check(L.tsu_nn_conv(kind, dy, _ptr(self.wt[li]), None, inputs[li], _ptr(dst), n, h, w, s),
                      "dgrad%d" % li)
                dy = _ptr(dst)
        return self.loss_part

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.t += 1
        lr_t = lr * math.sqrt(1.0 - b2 ** self.t) / (1.0 - b1 ** self.t)
        check(lib().tsu_nn_adam(_ptr(self.p), _ptr(self.g), _ptr(self.m), _ptr(self.v), self.arch.nparam, C.c_float(lr_t),
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
