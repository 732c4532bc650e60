"""tsu_nn_wgrad_noalloc against tsu_nn_wgrad through the C ABI: the same
weight and bias gradients, bit for bit, on two calls in a row, the second on
a smaller input (and another layer kind) than the first, so that anything
the first call left in the device scratch would show."""
import ctypes as C

import numpy as np
import pytest

from tsuner_amd._lib import lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

W64 = 16


def _bf16(rng, shape):
    return torch.from_numpy(rng.normal(0, 1, size=shape).astype(np.float32)).to(torch.bfloat16).cuda()


def _one(rng, kind, kx, my, mrows, n, h, w):
    print("no-alloc wgrad: kind %d, %d x %d x %d pixels" % (kind, n, h, w), flush=True)
    x = _bf16(rng, (n, h, w, kx))
    dy = _bf16(rng, (n, h, w, my))
    P = lambda t: C.c_void_p(t.data_ptr())
    s = C.c_void_p(0)
    nws = lib().tsu_nn_wgrad_workspace(kind)
    ws = torch.zeros(nws, dtype=torch.float32, device="cuda")
    dw_a = torch.full((9 * mrows * kx,), 3.0, dtype=torch.float32, device="cuda")
    db_a = torch.zeros(my, dtype=torch.float32, device="cuda")
    dw_b = torch.full((9 * mrows * kx,), 3.0, dtype=torch.float32, device="cuda")
    db_b = torch.zeros(my, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert lib().tsu_nn_wgrad(kind, P(x), P(dy), P(dw_a), P(db_a), mrows, n, h, w, P(ws), nws, s) == 0
    assert lib().tsu_nn_wgrad_noalloc(kind, P(x), P(dy), P(dw_b), P(db_b), mrows, n, h, w, s) == 0
    torch.cuda.synchronize()
    a, b = dw_a.cpu().numpy(), dw_b.cpu().numpy()
    assert np.abs(a).max() > 0
    np.testing.assert_array_equal(b, a)
    np.testing.assert_array_equal(db_b.cpu().numpy(), db_a.cpu().numpy())


def test_wgrad_noalloc_equals_wgrad_twice():
    rng = np.random.default_rng(5)
    _one(rng, 6 | W64, 64, 64, 64, 3, 20, 70)       # the largest workspace
    _one(rng, 5, 8, 32, 32, 1, 9, 33)               # a smaller input, another kind
