"""Seeded radiance-cache query batches (16-float rows, see include/nuray.h) and smooth targets."""
import numpy as np


def queries(n, seed=0, extent=1000.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    q = np.zeros((n, 16), np.float32)
    q[:, 0:3] = rng.uniform(0, extent, (n, 3))
    q[:, 1] *= 0.2
    q[:, 3] = rng.uniform(0.05, 0.8, n)
    q[:, 4:8] = rng.uniform(0, 1, (n, 4))
    q[:, 8:11] = rng.uniform(0.1, 0.9, (n, 3))
    return q


def smooth_target(q):
    """A smooth radiance-like field of position, direction and albedo (non-negative)."""
    p = q[:, 0:3].astype(np.float64) / 1000.0
    base = 0.6 + 0.4 * np.sin(6.0 * p[:, 0]) * np.cos(5.0 * p[:, 2]) + 0.3 * q[:, 4]
    return (base[:, None] * q[:, 8:11] * (1.0 + 0.5 * q[:, 5:6])).astype(np.float32)


def identity_weights(nparam):
    """Weights whose output is the albedo: layer 0 rows 0..2 pick inputs 41..43, hidden layers
    pass units 0..2 through, layer 5 rows 0..2 read units 0..2."""
    w = np.zeros(nparam, np.float32)
    for c in range(3):
        w[0 * 4096 + c * 64 + 41 + c] = 1.0
        for l in range(1, 5):
            w[l * 4096 + c * 64 + c] = 1.0
        w[5 * 4096 + c * 64 + c] = 1.0
    return w
