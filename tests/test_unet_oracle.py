"""Pins of the denoiser's CPU oracle (oracle/unet_ref.py), which the GPU
kernels are checked against: analytic gradients equal central finite
differences, the layer table matches the stated design, and the building
blocks satisfy their closed forms."""
import numpy as np

from oracle import unet_ref as U


def test_layer_table_is_the_stated_unet():
    L = U.layer_shapes()
    assert [n for n, *_ in L][:4] == ["enc0a", "enc0b", "enc1a", "enc1b"]
    assert L[0][1:4] == (10, 0, 32) and L[9][1:4] == (128, 0, 128)
    assert L[10] == ("dec3a", 128, 96, 96, True)  # up(128) ++ skip(96)
    assert L[-3][1:4] == (48, 32, 32) and L[-1] == ("out", 32, 0, 3, False)
    assert U.num_params() == sum(9 * (a + b) * o + o for _, a, b, o, _ in L)


def test_gradient_equals_finite_differences():
    w = (3, 4, 5)
    th = U.init_params(3, w)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(1, 8, 8, 10))
    t = rng.normal(size=(1, 8, 8, 3))
    _, g = U.loss_and_grad(th, x, t, w)
    eps = 1e-6
    for i in rng.choice(th.size, 40, replace=False):
        e = np.zeros_like(th)
        e[i] = eps
        fd = (U.loss(U.forward(th + e, x, w), t) - U.loss(U.forward(th - e, x, w), t)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd)


def test_conv_is_a_correlation_with_zero_padding():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 5, 6, 2))
    W = rng.normal(size=(9, 3, 2))
    y = U.conv(x, W, np.zeros(3))
    # brute force at every pixel
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    for i in range(5):
        for j in range(6):
            ref = sum(xp[0, i + 1 + t // 3 - 1, j + 1 + t % 3 - 1] @ W[t].T for t in range(9))
            np.testing.assert_allclose(y[0, i, j], ref, rtol=1e-12, atol=1e-12)


def test_pool_and_upsample_are_adjoint_pairs():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(2, 4, 6, 3))
    b = rng.normal(size=(2, 2, 3, 3))
    # <pool(a), b> = <a, pool_backward(b)>; <up(b), a> = <b, up_backward(a)>
    assert np.isclose((U.pool(a) * b).sum(), (a * U.pool_backward(b)).sum())
    assert np.isclose((U.up(b) * a).sum(), (b * U.up_backward(a)).sum())
