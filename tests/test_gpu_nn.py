"""Correction-network kernels (tsu_nn_* in include/tsuner.h) against the CPU
network oracle (oracle/nn_oracle.py), on BF16-rounded operands.

Tolerances: a BF16 output may differ from the FP64-accumulated oracle by one
BF16 rounding step (relative 2^-7) because the kernels accumulate in FP32;
FP32 outputs (the last layer, weight gradients) within 1e-5 of the largest
reference magnitude."""
import ctypes as C

import numpy as np
import pytest

import nn_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev_bf16(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).cuda()


def dev_f32(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.float().cpu().numpy()


_KEEP = []


def P(t):
    """Device address of t; t is kept alive until the test ends, so that
    temporaries passed inline are not recycled by the caching allocator while
    the enqueued kernels still use them."""
    _KEEP.append(t)
    return C.c_void_p(t.data_ptr())


@pytest.fixture(autouse=True)
def _release():
    yield
    torch.cuda.synchronize()
    _KEEP.clear()


def lib():
    from tsuner_amd import _lib
    return _lib.lib()


def ok(rc):
    from tsuner_amd import _lib
    _lib.check(rc, "call")


def assert_bf16_close(got, ref):
    err = np.abs(got - ref)
    tol = 2.0 ** -7 * np.abs(ref) + 1e-6 * max(1e-30, np.abs(ref).max())
    bad = err > tol
    assert not bad.any(), "max err %g at %s (ref %g got %g)" % (
        err.max(), np.argwhere(bad)[0], ref[tuple(np.argwhere(bad)[0])], got[tuple(np.argwhere(bad)[0])])


SHAPE = (2, 5, 70)     # n, h, w: two 64-pixel segments per row, one partial


def rand_bf16(rng, shape, scale=1.0):
    return O.bf16(rng.normal(0, scale, size=shape).astype(np.float32))


W64 = 16   # TSU_NN_W64


@pytest.mark.parametrize("width", [32, 64])
def test_mfma_conv_forward_kinds(width):
    rng = np.random.default_rng(1)
    n, h, w = SHAPE
    s = C.c_void_p(0)
    f = W64 if width == 64 else 0
    for kind, k, m, mreal in [(0, 8, width, width), (1, width, width, width), (2, width, 16, 4)]:
        x = rand_bf16(rng, (n, h, w, k))
        wt = rand_bf16(rng, (9, m, k), 0.2)
        if mreal < m:
            wt[:, mreal:, :] = 0
        b = rng.normal(0, 0.1, size=m).astype(np.float32)
        ref = O.conv(x, wt[:, :mreal, :], b[:mreal])
        if kind == 2:
            out = torch.zeros(n * h * w * 4, dtype=torch.float32, device="cuda")
        else:
            out = torch.zeros(n * h * w * m, dtype=torch.bfloat16, device="cuda")
        ok(lib().tsu_nn_conv(kind | f, P(dev_bf16(x)), P(dev_bf16(wt)), P(dev_f32(b)), None, P(out), n, h, w, s))
        got = host(out).reshape(n, h, w, -1)
        if kind == 2:
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
        else:
            assert_bf16_close(got, O.bf16(np.maximum(ref, 0).astype(np.float32)))


@pytest.mark.parametrize("width", [32, 64])
def test_mfma_conv_data_gradients_with_relu_mask(width):
    rng = np.random.default_rng(2)
    n, h, w = SHAPE
    s = C.c_void_p(0)
    f, W = (W64 if width == 64 else 0), width
    # through the output layer: dy has 8 channels (4 used), weights [9][16][W]
    w4 = rand_bf16(rng, (9, 16, W), 0.2)
    w4[:, 4:, :] = 0
    dy = np.zeros((n, h, w, 8), np.float32)
    dy[..., :4] = rand_bf16(rng, (n, h, w, 4))
    act = np.maximum(rand_bf16(rng, (n, h, w, W)), 0)
    wb = torch.zeros(9 * 16 * W, dtype=torch.bfloat16, device="cuda")
    wtr = torch.zeros(9 * W * 8, dtype=torch.bfloat16, device="cuda")
    ok(lib().tsu_nn_pack(P(dev_f32(w4)), P(wb), P(wtr), 16, W, 8, s))
    np.testing.assert_array_equal(host(wb).reshape(9, 16, W), w4)
    np.testing.assert_array_equal(host(wtr).reshape(9, W, 8), np.transpose(w4[:, :8, :], (0, 2, 1)))
    out = torch.zeros(n * h * w * W, dtype=torch.bfloat16, device="cuda")
    ok(lib().tsu_nn_conv(3 | f, P(dev_bf16(dy)), P(wtr), None, P(dev_bf16(act)), P(out), n, h, w, s))
    ref = O.conv_input_grad(dy[..., :4], w4[:, :4, :]) * (act > 0)
    assert_bf16_close(host(out).reshape(n, h, w, W), O.bf16(ref.astype(np.float32)))
    # middle layer
    w2 = rand_bf16(rng, (9, W, W), 0.2)
    d = rand_bf16(rng, (n, h, w, W))
    wt2 = torch.zeros(9 * W * W, dtype=torch.bfloat16, device="cuda")
    ok(lib().tsu_nn_pack(P(dev_f32(w2)), P(torch.zeros_like(wt2)), P(wt2), W, W, W, s))
    ok(lib().tsu_nn_conv(4 | f, P(dev_bf16(d)), P(wt2), None, P(dev_bf16(act)), P(out), n, h, w, s))
    ref = O.conv_input_grad(d, w2) * (act > 0)
    assert_bf16_close(host(out).reshape(n, h, w, W), O.bf16(ref.astype(np.float32)))


@pytest.mark.parametrize("width", [32, 64])
def test_mfma_weight_gradients(width):
    rng = np.random.default_rng(3)
    n, h, w = SHAPE
    s = C.c_void_p(0)
    f, W = (W64 if width == 64 else 0), width
    for kind, kx, my, mrows in [(5 | f, 8, W, W), (6 | f, W, W, W), (7 | f, W, 8, 16)]:
        x = rand_bf16(rng, (n, h, w, kx))
        dy = rand_bf16(rng, (n, h, w, my))
        dw = torch.full((9 * mrows * kx,), 7.0, dtype=torch.float32, device="cuda")
        db = torch.zeros(my, dtype=torch.float32, device="cuda")
        nws = lib().tsu_nn_wgrad_workspace(kind)
        ws = torch.zeros(nws, dtype=torch.float32, device="cuda")
        ok(lib().tsu_nn_wgrad(kind, P(dev_bf16(x)), P(dev_bf16(dy)), P(dw), P(db), mrows, n, h, w, P(ws), nws, s))
        rdw, rdb = O.conv_weight_grad(x, dy)
        got = host(dw).reshape(9, mrows, kx)
        np.testing.assert_allclose(got[:, :my, :], rdw, rtol=0, atol=1e-5 * np.abs(rdw).max())
        assert np.all(got[:, my:, :] == 7.0)      # rows beyond the gradient are left alone
        np.testing.assert_allclose(host(db), rdb, rtol=0, atol=1e-5 * np.abs(rdb).max())


def test_wide_conv_workgroups_loop_over_tiles():
    """The 64-channel kernels keep one workgroup per CU and loop over 8 x 64
    pixel tiles (prefetching the next one): an image set with more than
    twice as many tiles as CUs, ragged in both directions, against the
    oracle, forward and data gradient."""
    rng = np.random.default_rng(9)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    h, w = 45, 150                      # 6 x 3 tiles per image, partial ones at both edges
    n = (2 * cus) // 18 + 1
    s = C.c_void_p(0)
    x = rand_bf16(rng, (n, h, w, 64))
    wt = rand_bf16(rng, (9, 64, 64), 0.1)
    b = rng.normal(0, 0.1, size=64).astype(np.float32)
    out = torch.zeros(n * h * w * 64, dtype=torch.bfloat16, device="cuda")
    ok(lib().tsu_nn_conv(1 | W64, P(dev_bf16(x)), P(dev_bf16(wt)), P(dev_f32(b)), None, P(out), n, h, w, s))
    assert_bf16_close(host(out).reshape(n, h, w, 64), O.bf16(np.maximum(O.conv(x, wt, b), 0).astype(np.float32)))
    act = np.maximum(rand_bf16(rng, (n, h, w, 64)), 0)
    wtr = torch.zeros(9 * 64 * 64, dtype=torch.bfloat16, device="cuda")
    ok(lib().tsu_nn_pack(P(dev_f32(wt)), P(torch.zeros_like(wtr)), P(wtr), 64, 64, 64, s))
    ok(lib().tsu_nn_conv(4 | W64, P(dev_bf16(x)), P(wtr), None, P(dev_bf16(act)), P(out), n, h, w, s))
    ref = O.conv_input_grad(x, wt) * (act > 0)
    assert_bf16_close(host(out).reshape(n, h, w, 64), O.bf16(ref.astype(np.float32)))


def test_mse_and_adam_kernels():
    rng = np.random.default_rng(4)
    npix = 1000
    out = rng.normal(size=(npix, 4)).astype(np.float32)
    tgt = rng.normal(size=(npix, 4)).astype(np.float32)
    dy = torch.zeros(npix * 8, dtype=torch.bfloat16, device="cuda")
    lp = torch.zeros(64, dtype=torch.float64, device="cuda")
    ok(lib().tsu_nn_mse(P(dev_f32(out)), P(dev_f32(tgt)), P(dy), P(lp), 64, npix, C.c_float(0.5), C.c_void_p(0)))
    g = host(dy).reshape(npix, 8)
    np.testing.assert_array_equal(g[:, :4], O.bf16((0.5 * (out - tgt)).astype(np.float32)))
    assert np.all(g[:, 4:] == 0)
    d = (out - tgt).astype(np.float64)
    np.testing.assert_allclose(lp.cpu().numpy().sum(), (d * d).sum(), rtol=1e-12)
    # Adam: one step from m = v = 0 moves every parameter by lr_t * sign-ish
    p = rng.normal(size=300).astype(np.float32)
    gr = rng.normal(size=300).astype(np.float32)
    tp, tg = dev_f32(p), dev_f32(gr)
    tm, tv = torch.zeros_like(tp), torch.zeros_like(tp)
    ok(lib().tsu_nn_adam(P(tp), P(tg), P(tm), P(tv), 300, C.c_float(1e-2), C.c_float(0.9), C.c_float(0.999),
                         C.c_float(1e-8), C.c_void_p(0)))
    # the kernel's FP32 constants: (1 - 0.9f) and (1 - 0.999f) are not 0.1 and 0.001
    one, b1, b2 = np.float32(1), np.float32(0.9), np.float32(0.999)
    m = (one - b1) * gr
    v = (one - b2) * gr * gr
    np.testing.assert_allclose(host(tm), m, rtol=1e-6)
    np.testing.assert_allclose(host(tv), v, rtol=1e-6)
    np.testing.assert_allclose(host(tp), p - np.float32(1e-2) * m / (np.sqrt(v) + np.float32(1e-8)), rtol=1e-5,
                               atol=1e-7)


@pytest.mark.parametrize("width,mids", [(32, 2), (64, 1), (32, 4)])
def test_network_gradients_match_oracle(width, mids):
    """Full forward + backward on the GPU against the oracle's loss_and_grads
    (pinned by finite differences in test_nn_oracle.py), run with BF16
    rounding at the points where the GPU stores BF16 (hidden activations, the
    output gradient, the masked data gradients): norm-wise 1% per tensor, the
    rest being FP32 accumulation and the rare BF16 rounding flips it causes.
    Against the unrounded FP64 oracle the same gradients agree to 8%."""
    from tsuner_amd import nn
    arch = nn.Arch(width, mids)
    last = arch.nlayers - 1
    rng = np.random.default_rng(5)
    n, h, w = 2, 12, 20
    p = nn.init_params(7, arch)
    # give the output layer full-size weights so every layer carries gradient
    wo, bo = arch.offs[last]
    p[wo:wo + 9 * 16 * width] *= 10.0
    for li, (m, k) in enumerate(arch.layers):
        p[arch.offs[li][1]:arch.offs[li][1] + m] = rng.normal(0, 0.05, m)
    p[bo + 4:bo + 16] = 0
    p = O.bf16(p)
    net = nn.Net(p, max_batch=n, h=h, w=w, arch=arch)
    x = rand_bf16(rng, (n, h, w, 8))
    tgt = rng.normal(0, 1, size=(n, h, w, 4)).astype(np.float32)
    tx, tt = dev_bf16(x), dev_f32(tgt)
    net.forward(tx, n)
    net.backward(tx, tt, n)
    gpu_loss = net.loss(n)
    g = net.g.cpu().numpy()
    params = []
    for li, (m, k) in enumerate(arch.layers):
        wo, bo = arch.offs[li]
        W = p[wo:wo + 9 * m * k].reshape(9, m, k).astype(np.float64)
        b = p[bo:bo + m].astype(np.float64)
        if li == last:
            W, b = W[:, :4, :], b[:4]
        params.append((W, b))
    loss, grads = O.loss_and_grads(x.astype(np.float64), tgt.astype(np.float64), params, rnd=O.bf16_64)
    loss64, grads64 = O.loss_and_grads(x.astype(np.float64), tgt.astype(np.float64), params)
    assert abs(gpu_loss - loss) <= 1e-3 * loss
    for li, (m, k) in enumerate(arch.layers):
        wo, bo = arch.offs[li]
        gw = g[wo:wo + 9 * m * k].reshape(9, m, k)
        gb = g[bo:bo + m]
        if li == last:
            assert np.all(gw[:, 4:, :] == 0) and np.all(gb[4:] == 0)
            gw, gb = gw[:, :4, :], gb[:4]
        for (rw, rb), tol in ((grads[li], 0.01), (grads64[li], 0.08)):
            assert np.linalg.norm(gw - rw) <= tol * np.linalg.norm(rw), (li, tol)
            assert np.linalg.norm(gb - rb) <= tol * np.linalg.norm(rb), (li, tol)


def test_training_reduces_loss():
    """Adam on the GPU fits a target produced by a fixed random network of
    the same shape (realisable): the loss falls by 10x."""
    from tsuner_amd import nn
    rng = np.random.default_rng(6)
    n, h, w = 4, 16, 24
    teacher = nn.Net(nn.init_params(11), max_batch=n, h=h, w=w)
    x = dev_bf16(rand_bf16(rng, (n, h, w, 8)))
    tgt = teacher.forward(x, n).clone()
    student = nn.Net(nn.init_params(12), max_batch=n, h=h, w=w)
    student.forward(x, n)
    student.backward(x, tgt, n)
    first = student.loss(n)
    for _ in range(300):
        student.train_step(x, tgt, n, lr=3e-3)
    student.forward(x, n)
    student.backward(x, tgt, n)
    assert student.loss(n) < 0.1 * first, (first, student.loss(n))
