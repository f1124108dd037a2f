"""Synthetic workloads of the benchmark and the batch tests.

* ``sod_3d``: the 3D Cartesian Sod tube along x (``3d_sod_x``,
  torchSolver3D.py:271-274): (rho, p) = (1, 1) for cell centres x <= 0.5,
  (0.125, 0.1) beyond, at rest; open x ends, reflective y/z walls.
* ``uneven_problems``: seeded mixes of 1D (N in 256..8192), 2D (64^2..512^2)
  and 3D (32^3..160^3) boxes, dimension cycling 1, 2, 3, each with random
  two-state planar or spherical Riemann data (rho in [0.1, 2], |u| <= 1 in a
  uniform direction, p in [0.05, 5]), gamma in {1.4, 1.67}, a random limiter
  of the five and open boundaries.

Problems are generated one at a time so a large mix never needs all its
host arrays at once.
"""

import numpy as np

LIMITERS = ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos")
SIZES = {1: (256, 8192), 2: (64, 512), 3: (32, 160)}


def sod_3d(n):
    """Primitive (5, n, n, n) of the 3D Sod tube on the unit cube."""
    x = (np.arange(n) + 0.5) / n
    left = (x <= 0.5)[None, None, :]
    prim = np.zeros((5, n, n, n))
    prim[0] = np.where(left, 1.0, 0.125)
    prim[4] = np.where(left, 1.0, 0.1)
    return prim


def _centres(counts):
    """Cell-centre coordinates in x, y, z order as arrays of the state shape."""
    grids = np.meshgrid(*[(np.arange(n) + 0.5) / n for n in counts[::-1]], indexing="ij", sparse=True)
    return grids[::-1]


def _direction(rng, dim):
    v = rng.normal(size=dim)
    norm = np.linalg.norm(v)
    return v / norm if norm > 0 else np.eye(dim)[0]


def uneven_problem(rng, dim):
    """One random two-state problem; returns (primitive, options)."""
    lo, hi = SIZES[dim]
    n = int(rng.integers(lo, hi + 1))
    counts = [n] * dim
    xyz = _centres(counts)
    if rng.random() < 0.5:
        axis = int(rng.integers(dim))
        inside = xyz[axis] < 0.5
        shape_desc = f"planar axis {axis}"
    else:
        radius = rng.uniform(0.1, 0.3)
        inside = sum((x - 0.5) ** 2 for x in xyz) < radius * radius
        shape_desc = f"sphere r={radius:.3f}"
    shape = tuple(counts[::-1])
    inside = np.broadcast_to(inside, shape)
    prim = np.empty((dim + 2,) + shape)
    for mask in (inside, ~inside):
        rho = rng.uniform(0.1, 2.0)
        speed = rng.uniform(0.0, 1.0)
        vel = speed * _direction(rng, dim)
        p = rng.uniform(0.05, 5.0)
        prim[0][mask] = rho
        for k in range(dim):
            prim[1 + k][mask] = vel[k]
        prim[-1][mask] = p
    gamma = 1.4 if rng.random() < 0.5 else 1.67
    limiter = LIMITERS[int(rng.integers(len(LIMITERS)))]
    opts = dict(lengths=[1.0] * dim, gamma=gamma, limiter=limiter, boundaries=["Open"] * dim, cfl=0.4,
                desc=f"{dim}D n={n} {shape_desc} {limiter} gamma={gamma}")
    return prim, opts


def uneven_problems(seed=508, count=256):
    """Yields (index, primitive, options) for the batched uneven mix."""
    rng = np.random.default_rng(seed)
    for i in range(count):
        dim = (1, 2, 3)[i % 3]
        prim, opts = uneven_problem(rng, dim)
        yield i, prim, opts

