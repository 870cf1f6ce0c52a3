"""The MLP checker's backward is the gradient of its forward (float64, no
quantisation), by central finite differences."""
import numpy as np

from oracle import mlp_ref


def test_mlp_checker_gradient_fd():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.2, 37893)
    X = rng.normal(size=(7, 32))
    T = rng.normal(size=(7, 5))
    M = (rng.uniform(size=(7, 5)) > 0.2).astype(float)
    mu, sd = rng.normal(size=32) * 0.1, rng.uniform(0.5, 2.0, 32)
    loss, g = mlp_ref.loss_grad(p, X, T, M, mu, sd, quant=False, dtype=np.float64)
    idx = np.concatenate([rng.integers(0, 37893, 60), [4096, 20608, 37120, 37248, 37888, 37892]])
    for i in idx:
        e = np.zeros_like(p); e[i] = 1e-6
        lp, _ = mlp_ref.loss_grad(p + e, X, T, M, mu, sd, quant=False, dtype=np.float64)
        lm, _ = mlp_ref.loss_grad(p - e, X, T, M, mu, sd, quant=False, dtype=np.float64)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.1415926, 65504.0, 1e-20], np.float32)
    y = mlp_ref.bf16(x)
    # bf16 keeps 8 mantissa bits: 1 + 2^-8 is exact, ties go to even
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == np.float32(1.0078125)
    assert abs(y[3] - np.float32(-3.140625)) == 0
    assert np.all(np.abs(y - x) <= np.abs(x) * 2.0 ** -8)
