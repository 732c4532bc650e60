"""Pins of the network oracle (oracle/nn_oracle.py): the convolution against
a direct loop over pixels, the backward pass against finite differences, and
BF16 rounding against its definition."""
import numpy as np

import nn_oracle as N


def direct_conv(x, w, b):
    n, H, W, ci = x.shape
    co = w.shape[1]
    y = np.zeros((n, H, W, co))
    for i in range(n):
        for r in range(H):
            for c in range(W):
                for t, (dy, dx) in enumerate(N.TAPS):
                    rr, cc = r + dy, c + dx
                    if 0 <= rr < H and 0 <= cc < W:
                        y[i, r, c] += w[t] @ x[i, rr, cc]
                y[i, r, c] += b
    return y


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 7, 3))
    w = rng.normal(size=(9, 4, 3))
    b = rng.normal(size=4)
    np.testing.assert_allclose(N.conv(x, w, b), direct_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    params = N.init_params(rng, cin=3, width=4, cout=2)
    params = [(w.astype(np.float64), rng.normal(0, 0.1, b.shape)) for w, b in params]
    x = rng.normal(size=(2, 6, 5, 3))
    tgt = rng.normal(size=(2, 6, 5, 2))
    loss, grads = N.loss_and_grads(x, tgt, params)
    h = 1e-6
    for k in range(4):
        w, b = params[k]
        for idx in [(0, 0, 0), (4, 1, 2), (8, w.shape[1] - 1, w.shape[2] - 1)]:
            wp = w.copy()
            wp[idx] += h
            wm = w.copy()
            wm[idx] -= h
            pp = list(params)
            pp[k] = (wp, b)
            lp, _ = N.loss_and_grads(x, tgt, pp)
            pp[k] = (wm, b)
            lm, _ = N.loss_and_grads(x, tgt, pp)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[k][0][idx]) <= 1e-6 * max(1.0, abs(fd)), (k, idx, fd, grads[k][0][idx])
        bp = b.copy()
        bp[0] += h
        bm = b.copy()
        bm[0] -= h
        pp = list(params)
        pp[k] = (w, bp)
        lp, _ = N.loss_and_grads(x, tgt, pp)
        pp[k] = (w, bm)
        lm, _ = N.loss_and_grads(x, tgt, pp)
        assert abs((lp - lm) / (2 * h) - grads[k][1][0]) <= 1e-6


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265], dtype=np.float32)
    got = N.bf16(vals)
    want = np.array([1.0, 1.0 + 2 ** -7 * 0 + 2 ** -8 * 0 + 1.0 * 0, 1.0, 1.0 + 2 ** -7, -2.5, 3.140625],
                    dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa -> 1
    want[1] = 1.0
    np.testing.assert_array_equal(got, want)
    assert np.all((N.bf16(vals).view(np.uint32) & 0xFFFF) == 0)
