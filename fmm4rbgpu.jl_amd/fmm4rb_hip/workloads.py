"""Synthetic beams of the benchmark configurations (seeded numpy generators).

uniform_cube : positions i.i.d. U[0,1)^3, momenta zero (N = 8192 case)
bunch        : relativistic Gaussian bunch, sigma = (1e-3, 1e-3, 1e-4),
               p_z ~ N(100, 1), p_x, p_y ~ N(0, 0.01) in mc units
ellipsoid    : uniform in the ellipsoid with semi-axes (1e-3, 1e-3, 1e-4),
               p_z ~ N(200, 2), p_x, p_y ~ N(0, 0.02)
All return (pos, mom) as (N, 3) float64 arrays.
"""
from __future__ import annotations

import numpy as np


def uniform_cube(N: int, seed: int = 1):
    rng = np.random.default_rng(seed)
    return rng.random((N, 3)), np.zeros((N, 3))


def bunch(N: int, seed: int = 2, sigma=(1e-3, 1e-3, 1e-4), pz=(100.0, 1.0), pt=0.01):
    rng = np.random.default_rng(seed)
    pos = rng.standard_normal((N, 3)) * np.asarray(sigma)
    mom = np.empty((N, 3))
    mom[:, :2] = rng.standard_normal((N, 2)) * pt
    mom[:, 2] = pz[0] + pz[1] * rng.standard_normal(N)
    return pos, mom


def ellipsoid(N: int, seed: int = 5, axes=(1e-3, 1e-3, 1e-4), pz=(200.0, 2.0), pt=0.02):
    rng = np.random.default_rng(seed)
    out = np.empty((0, 3))
    while len(out) < N:
        u = rng.uniform(-1.0, 1.0, size=(2 * (N - len(out)) + 16, 3))
        u = u[(u * u).sum(axis=1) <= 1.0]
        out = np.concatenate([out, u])
    pos = out[:N] * np.asarray(axes)
    mom = np.empty((N, 3))
    mom[:, :2] = rng.standard_normal((N, 2)) * pt
    mom[:, 2] = pz[0] + pz[1] * rng.standard_normal(N)
    return pos, mom


def rel_l2(a, b):
    """Relative L2 error ||a - b|| / ||b|| over all particles and components."""
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b))) / (nb if nb > 0 else 1.0)
