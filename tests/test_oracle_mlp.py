"""Pins the FP64 MLP reference (oracle/mlp_ref.py) against central finite
differences: parameter gradients of the MSE loss and input Jacobians."""
import numpy as np

from mlp_ref import MLPRef, bf16


def _net(seed=3, n_in=5, pad=64, H=64, n_out=64):
    rng = np.random.default_rng(seed)
    dims = [pad, H, H, H, n_out]
    n = sum(dims[l] * dims[l + 1] + dims[l + 1] for l in range(4))
    p = rng.normal(0, 0.2, n)
    net = MLPRef(p, n_in, pad, H, n_out)
    net.W[0][n_in:] = 0.0
    return net, net.pack(net.W, net.b)


def test_param_gradient_matches_fd():
    net, p = _net()
    rng = np.random.default_rng(4)
    X = rng.normal(size=(16, 5))
    T = rng.normal(size=(16, 64))
    _, g = net.loss_grad(X, T)
    idx = rng.choice(np.flatnonzero(np.abs(g) > 1e-4 * np.abs(g).max()), 40, replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = MLPRef(pp, 5, 64, 64, 64).loss_grad(X, T)
        lm, _ = MLPRef(pm, 5, 64, 64, 64).loss_grad(X, T)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * np.abs(g).max() + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_input_jacobian_matches_fd():
    net, _ = _net(5)
    rng = np.random.default_rng(6)
    X = rng.normal(size=(4, 5))
    J = net.input_jacobian(X)
    h = 1e-6
    for i in range(5):
        Xp, Xm = X.copy(), X.copy()
        Xp[:, i] += h
        Xm[:, i] -= h
        fd = (net.forward(Xp) - net.forward(Xm)) / (2 * h)
        assert np.max(np.abs(fd - J[:, :, i])) < 1e-7 * max(1.0, np.abs(J).max())


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 1e-40, 65536.0 * 1.5])
    r = bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0           # halfway -> even
    assert r[2] == 1.0078125                      # above halfway -> up
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 + 1e-38)
