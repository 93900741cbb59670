"""Pins of the FP64 U-Net oracle (oracle/unet_ref.py) on the CPU.

The oracle is the checker of the HIP U-Net kernels.  It is pinned here by
(1) a brute-force loop convolution with explicit zero padding, (2) the
adjoint identity <conv(x), y> = <x, conv^T(y)> that ties the data gradient
to the forward convolution, (3) central finite differences of the whole
network's parameter gradient, and (4) the parameter layout and count.
"""
import numpy as np

from oracle import unet_ref as U


def brute_conv(x, W, b):
    Z, Y, X, _ = x.shape
    co_n, ntap, _ = W.shape
    taps = U._taps(ntap)
    y = np.zeros((Z, Y, X, co_n))
    for z in range(Z):
        for yy in range(Y):
            for xx in range(X):
                acc = b.copy()
                for t, (dz, dy, dx) in enumerate(taps):
                    zz, y2, x2 = z + dz, yy + dy, xx + dx
                    if 0 <= zz < Z and 0 <= y2 < Y and 0 <= x2 < X:
                        acc = acc + W[:, t, :] @ x[zz, y2, x2]
                y[z, yy, xx] = acc
    return y


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 4, 5, 8))
    for ntap in (27, 1):
        W = rng.standard_normal((6, ntap, 8))
        b = rng.standard_normal(6)
        assert np.allclose(U.conv(x, W, b), brute_conv(x, W, b), atol=1e-12)


def test_conv_adjoint_identity():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 4, 6, 8))
    W = rng.standard_normal((16, 27, 8))
    dy = rng.standard_normal((4, 4, 6, 16))
    dx, dW, db = U.conv_back(x, W, dy)
    lhs = np.sum(U.conv(x, W, np.zeros(16)) * dy)
    assert np.isclose(lhs, np.sum(x * dx), rtol=1e-12)
    assert np.isclose(lhs, np.sum(W * dW), rtol=1e-12)
    assert np.allclose(db, dy.sum(axis=(0, 1, 2)))


def test_pool_up_adjoints():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((4, 4, 4, 3))
    c = rng.standard_normal((2, 2, 2, 3))
    assert np.isclose(np.sum(U.pool(a) * c), np.sum(a * U.pool_back(c)))
    assert np.isclose(np.sum(U.up(c) * a), np.sum(c * U.up_back(a)))
    assert np.allclose(U.pool(U.up(c)), c)


def test_network_gradient_finite_differences():
    rng = np.random.default_rng(3)
    p = U.init_params(5).astype(np.float64)
    p += 0.01 * rng.standard_normal(p.size)  # nonzero biases
    x = rng.standard_normal((4, 4, 4, 8))
    out, cache = U.forward(p, x, round_bf16=False)
    d = rng.standard_normal(out.shape)
    g = U.backward(p, cache, d)
    gn = np.linalg.norm(g)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (np.sum(U.forward(p + e, x, False)[0] * d) - np.sum(U.forward(p - e, x, False)[0] * d)) / (2 * h)
        assert abs(fd - g[i]) < 1e-5 * gn + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 1e-3])
    r = U.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7
    assert abs(r[3] + 3.140625) < 1e-12
    assert abs(r[4] - 1e-3) < 1e-3 * 2 ** -8


def test_layout_matches_package():
    from giem2g_amd import learned
    assert U.n_params() == learned.n_params() == 356390
    assert [(ci, co, t) for _, ci, co, t, _ in U.LAYERS] == learned.ARCH
