"""Pins the NumPy oracle of the learned correction network (oracle/unet_ref.py) against
what the mathematics fixes: a brute-force loop convolution, adjoint identities, exact
interpolation of linear fields, and central finite differences of the full loss."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_ref as U  # noqa: E402


def brute_conv(x, W, b, stride):
    B, Z, Y, X, ci = x.shape
    oz, oy, ox = U.out_shape((Z, Y, X), stride)
    y = np.zeros((B, oz, oy, ox, W.shape[2]))
    for n in range(B):
        for k in range(oz):
            for j in range(oy):
                for i in range(ox):
                    acc = b.copy()
                    for t in range(27):
                        dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
                        a, bb, c = stride * i + dx, stride * j + dy, stride * k + dz
                        if 0 <= a < X and 0 <= bb < Y and 0 <= c < Z:
                            acc = acc + x[n, c, bb, a] @ W[t]
                    y[n, k, j, i] = acc
    return y


@pytest.mark.parametrize("stride,shape", [(1, (3, 4, 5)), (2, (5, 4, 7)), (2, (4, 6, 5))])
def test_conv_matches_brute_force(stride, shape):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2,) + shape + (3,))
    W = rng.normal(size=(27, 3, 5))
    b = rng.normal(size=5)
    np.testing.assert_allclose(U.conv3d(x, W, b, stride), brute_conv(x, W, b, stride), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("stride,shape", [(1, (4, 5, 6)), (2, (5, 6, 7)), (2, (6, 4, 5))])
def test_conv_adjoints(stride, shape):
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2,) + shape + (4,))
    W = rng.normal(size=(27, 4, 6))
    y = U.conv3d(x, W, None, stride)
    dy = rng.normal(size=y.shape)
    # <conv(x), dy> = <x, dgrad(dy)> = <W, wgrad(x, dy)>
    lhs = np.sum(y * dy)
    assert abs(lhs - np.sum(x * U.conv3d_dgrad(dy, W, shape, stride))) < 1e-9 * abs(lhs) + 1e-9
    assert abs(lhs - np.sum(W * U.conv3d_wgrad(x, dy, stride))) < 1e-9 * abs(lhs) + 1e-9


@pytest.mark.parametrize("coarse,fine", [((3, 4, 5), (5, 7, 9)), ((3, 3, 4), (6, 5, 8))])
def test_upsample_linear_and_adjoint(coarse, fine):
    rng = np.random.default_rng(3)
    zc, yc, xc = np.meshgrid(*[np.arange(n) for n in coarse], indexing="ij")
    lin = (0.3 + 0.5 * xc - 0.25 * yc + 2.0 * zc)[None, ..., None]
    up = U.upsample(lin, fine)
    zf, yf, xf = np.meshgrid(*[np.arange(n) / 2.0 for n in fine], indexing="ij")
    exact = 0.3 + 0.5 * xf - 0.25 * yf + 2.0 * zf
    inside = (2 * (np.array(coarse) - 1) >= np.array(fine) - 1).all()
    if inside:  # odd fine box: every fine node lies between existing coarse nodes
        np.testing.assert_allclose(up[0, ..., 0], exact, atol=1e-12)
    else:  # interior part is still exact
        sl = tuple(slice(0, 2 * (c - 1) + 1) for c in coarse)
        np.testing.assert_allclose(up[0][sl][..., 0], exact[sl], atol=1e-12)
    x = rng.normal(size=(2,) + coarse + (3,))
    dy = rng.normal(size=(2,) + fine + (3,))
    assert abs(np.sum(U.upsample(x, fine) * dy) - np.sum(x * U.upsample_adj(dy, coarse))) < 1e-10


def test_unet_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    params = U.init_params(5)
    # enlarge the last layer so the loss depends on every layer visibly
    params[-1] = (params[-1][0] * 10, rng.normal(0, 0.1, 3))
    shape = (5, 6, 7)
    feat = rng.normal(size=(2,) + shape + (4,))
    tgt = rng.normal(size=(2,) + shape + (3,))
    mask = (rng.random((2,) + shape) > 0.2).astype(np.uint8)
    loss, grads = U.loss_and_grad(params, feat, tgt, mask, bf=False)
    theta = U.pack(params)
    g = U.pack(grads)
    idx = rng.choice(theta.size, 40, replace=False)
    # include biases of the first and last layer and a stride-2 layer's weight
    offs = np.cumsum([0] + [27 * ci * co + co for ci, co, _, _ in U.layer_table()])
    idx = np.concatenate([idx, [offs[1] - 1, offs[3] + 5, offs[-1] - 2]])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        lp = U.loss_and_grad(U.unpack(tp), feat, tgt, mask)[0]
        lm = U.loss_and_grad(U.unpack(tm), feat, tgt, mask)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)) + 1e-7, (int(i), fd, g[i])


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0, 0.0], np.float32)
    r = U.bf16(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1.0); 1.005859375 = 1 + 1.5 * 2^-8 -> 1 + 2^-7
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert abs(r[3] / -3.0e-3 - 1) < 2 ** -8 and r[5] == 0.0
