"""Seeded synthetic volumes for benchmarks and examples (numpy, host).

ellipsoid_phantom: a piecewise-smooth volume, the sum of random ellipsoids
(centres uniform in the volume, semi-axes uniform in [8, 40] voxels, random
rotations, intensities uniform in [0, 1]) plus Gaussian noise.
"""
import numpy as np


def _rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def ellipsoid_phantom(dims, seed, n_ellipsoids=8, axes=(8.0, 40.0), noise=0.05, dtype=np.float32):
    rng = np.random.default_rng(seed)
    dims = tuple(int(d) for d in dims)
    grid = np.stack(np.meshgrid(*(np.arange(d, dtype=np.float64) for d in dims), indexing="ij"), axis=-1)
    vol = np.zeros(dims, dtype=np.float64)
    for _ in range(n_ellipsoids):
        c = rng.uniform(0, 1, 3) * np.array(dims)
        a = rng.uniform(axes[0], axes[1], 3)
        R = _rotation(rng)
        inten = rng.uniform(0, 1)
        d = (grid - c) @ R
        inside = np.sum((d / a) ** 2, axis=-1) <= 1.0
        vol[inside] += inten
    vol += rng.normal(0, noise, dims)
    return vol.astype(dtype)
