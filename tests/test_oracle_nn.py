"""CPU pins of the surrogate oracle (oracle/nn_oracle.py): im2col
convolution == direct loops, col2im is im2col's adjoint, and the network's
analytic gradients == central finite differences."""

import numpy as np

from oracle import nn_oracle as N


def test_conv_matches_direct_loops():
    rng = np.random.default_rng(0)
    for s in (1, 2):
        x = rng.normal(size=(2, 5, 6, 7, 3))
        w = rng.normal(size=(4, 3, 3, 3, 3))
        b = rng.normal(size=4)
        np.testing.assert_allclose(N.conv_fwd(x, w, b, s), N.conv_direct(x, w, b, s), rtol=1e-12, atol=1e-12)


def test_col2im_is_adjoint_of_im2col():
    rng = np.random.default_rng(1)
    for s in (1, 2):
        x = rng.normal(size=(2, 6, 5, 7, 2))
        g = rng.normal(size=N.im2col(x, s).shape)
        assert np.isclose(np.sum(N.im2col(x, s) * g), np.sum(x * N.col2im(g, x.shape, s)), rtol=1e-12)


def tiny_params(rng):
    return {"conv": [(rng.normal(size=(3, 3, 3, 3, 1)) * 0.5, rng.normal(size=3) * 0.1),
                     (rng.normal(size=(4, 3, 3, 3, 3)) * 0.3, rng.normal(size=4) * 0.1)],
            "fc1": (rng.normal(size=(5, 4)) * 0.5, rng.normal(size=5) * 0.1),
            "fc2": (rng.normal(size=(4, 5)) * 0.5, rng.normal(size=4) * 0.1)}


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    params = tiny_params(rng)
    strides = (2, 2)
    x = rng.normal(size=(3, 6, 7, 5, 1))
    t = rng.uniform(size=(3, 4))

    def loss_of(p):
        pred, _ = N.network_forward(p, x, strides)
        return N.quadratic(pred, t)[0]

    pred, cache = N.network_forward(params, x, strides)
    _, dpred = N.quadratic(pred, t)
    grads = N.network_backward(params, cache, dpred, strides)
    eps = 1e-6
    checks = [("conv", 0, 0), ("conv", 1, 0), ("conv", 0, 1), ("conv", 1, 1), ("fc1", None, 0), ("fc1", None, 1),
              ("fc2", None, 0), ("fc2", None, 1)]
    for name, li, which in checks:
        arr = params[name][li][which] if li is not None else params[name][which]
        g = grads[name][li][which] if li is not None else grads[name][which]
        for _ in range(4):
            idx = tuple(int(rng.integers(0, n)) for n in arr.shape)
            old = arr[idx]
            arr[idx] = old + eps
            lp = loss_of(params)
            arr[idx] = old - eps
            lm = loss_of(params)
            arr[idx] = old
            assert np.isclose(g[idx], (lp - lm) / (2 * eps), rtol=1e-5, atol=1e-8), (name, li, which, idx)


def test_training_loss_gradient():
    """The curve part of the training gradient is the squared error's; the
    variance part is d/ds of the likelihood term with se fixed."""
    rng = np.random.default_rng(3)
    pred = rng.normal(size=(4, 6))
    t = rng.uniform(size=(4, 5))
    _, _, d = N.gnll(pred, t)
    eps = 1e-6
    for b in range(4):
        for j in range(5):
            p2 = pred.copy()
            p2[b, j] += eps
            lp = 0.5 * ((p2[:, :5] - t) ** 2).sum() / 4
            p2[b, j] -= 2 * eps
            lm = 0.5 * ((p2[:, :5] - t) ** 2).sum() / 4
            assert np.isclose(d[b, j], (lp - lm) / (2 * eps), rtol=1e-6)
        se = ((pred[b, :5] - t[b]) ** 2).sum()
        f = lambda s: 0.5 * (se * np.exp(-s) + 5 * s) / 4  # noqa: E731
        assert np.isclose(d[b, 5], (f(pred[b, 5] + eps) - f(pred[b, 5] - eps)) / (2 * eps), rtol=1e-6)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    r = N.bf16(vals)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8
