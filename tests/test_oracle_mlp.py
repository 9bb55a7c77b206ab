"""The CPU reference network's backward pass against central finite differences
(float64, BF16 rounding switched off by monkeypatching it to the identity)."""
import numpy as np

from oracle import mlp_ref


def test_backward_matches_finite_differences(monkeypatch):
    monkeypatch.setattr(mlp_ref, "bf16", lambda x: np.asarray(x, dtype=np.float64))
    rng = np.random.default_rng(0)
    params = rng.normal(scale=0.3, size=mlp_ref.NPARAMS)
    X = rng.uniform(-1, 1, (32, 4))
    T = rng.uniform(0.5, 1.5, 32)
    _, _, g = mlp_ref.loss_and_grad(params, X, T)
    h = 1e-6
    for k in rng.choice(mlp_ref.NPARAMS, 60, replace=False):
        p1, p2 = params.copy(), params.copy()
        p1[k] += h
        p2[k] -= h
        fd = (mlp_ref.loss_and_grad(p1, X, T)[0] - mlp_ref.loss_and_grad(p2, X, T)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-3], dtype=np.float32)
    r = mlp_ref.bf16(vals)
    # representable in 8 bits of mantissa, nearest even
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12
    assert np.all(np.abs(r - vals) <= np.abs(vals) * 2 ** -8)
