"""Pin the MLP oracle: analytic gradients against central finite differences."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = mlp.init(1)
    X = rng.standard_normal((37, mlp.IN))
    t = rng.standard_normal(37)
    loss, g, _ = mlp.loss_and_grad(theta, X, t)
    idx = rng.choice(mlp.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [0, 1023, 1024 + 63, mlp.NPARAM - 1, mlp.NPARAM - 2]])
    h = 1e-6
    for k in idx:
        e = np.zeros(mlp.NPARAM)
        e[k] = h
        lp, _, _ = mlp.loss_and_grad(theta + e, X, t)
        lm, _, _ = mlp.loss_and_grad(theta - e, X, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_pack_roundtrip_and_count():
    theta = mlp.init(2)
    assert mlp.NPARAM == 9473
    np.testing.assert_array_equal(mlp.pack(mlp.unpack(theta)), theta)


def test_bf16_rounding_is_nearest_even():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265])
    r = mlp.to_bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0          # tie -> even
    assert r[2] == 1.0 + 2 ** -7               # above the tie -> up
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12
