"""Shared helpers of the GPU parity tests."""
import numpy as np

R, RCELL, LDIV = 0.8, 0.5, 2.4
THR = 0.5 * (2 * R + LDIV)


def rel_err(got, ref):
    """Normwise per-column relative error: max_i |got - ref| / max_i |ref| (column-wise)."""
    scale = np.maximum(np.max(np.abs(ref), axis=0), 1e-300)
    return float(np.max(np.abs(got - ref) / scale))


def random_box(n, side, seed, nz_min=0.0, zmin=0.0):
    rng = np.random.default_rng(seed)
    cells = np.zeros((n, 22))
    cells[:, 0] = np.arange(n)
    cells[:, 1:3] = rng.uniform(0, side, (n, 2))
    cells[:, 3] = zmin + rng.uniform(0, side / 2, n)
    v = rng.normal(size=(n, 3))
    v[:, 2] = np.abs(v[:, 2]) + nz_min
    cells[:, 4:7] = v / np.linalg.norm(v, axis=1)[:, None]
    hl = rng.uniform(0.4, 1.6, n)
    cells[:, 13], cells[:, 14] = 2 * hl, hl
    cells[:, 16], cells[:, 17], cells[:, 18], cells[:, 19], cells[:, 20] = 1.5, 0.7, 720, 100, 1
    cells[:, 21] = 100.0
    cells[:, 7:13] = rng.normal(scale=5.0, size=(n, 6))
    return np.asfortranarray(cells)
