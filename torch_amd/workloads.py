"""Synthetic workloads of the benchmark and the batch tests.

* ``sod_3d``: the 3D Cartesian Sod tube along x (``3d_sod_x``,
  torchSolver3D.py:271-274): (rho, p) = (1, 1) for cell centres x <= 0.5,
  (0.125, 0.1) beyond, at rest; open x ends, reflective y/z walls.
* ``quadrant_2d`` / ``sphere_3d``: the seeded 2D quadrant Riemann problem and
  the 3D pressure sphere (rho, p) = (1, 10) within radius 0.1 of the centre,
  (0.125, 0.1) outside, at rest.
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


def quadrant_2d(n, seed=20261203):
    """Primitive (4, n, n) of the seeded 2D quadrant Riemann problem: four
    constant states in the quadrants of the unit square split at (0.5, 0.5),
    drawn in the order upper right, upper left, lower left, lower right, each
    as (rho, u, v, p) with rho in [0.1, 1.5], u and v in [-1, 1] and p in
    [0.03, 1.5]."""
    rng = np.random.default_rng(seed)
    states = [(rng.uniform(0.1, 1.5), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.03, 1.5))
              for _ in range(4)]
    c = (np.arange(n) + 0.5) / n
    right, upper = (c > 0.5)[None, :], (c > 0.5)[:, None]
    prim = np.zeros((4, n, n))
    for q, (hi_y, hi_x) in enumerate(((True, True), (True, False), (False, False), (False, True))):
        mask = (upper == hi_y) & (right == hi_x)
        for k in range(4):
            prim[k] = np.where(mask, states[q][k], prim[k])
    return prim


def sphere_3d(n):
    """Primitive (5, n, n, n) of the 3D pressure sphere: rho = 1, p = 10 for
    cell centres within 0.1 of (0.5, 0.5, 0.5), rho = 0.125, p = 0.1 outside,
    at rest."""
    x, y, z = _centres([n, n, n])
    inside = (x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2 <= 0.1 ** 2
    prim = np.zeros((5, n, n, n))
    prim[0] = np.where(inside, 1.0, 0.125)
    prim[4] = np.where(inside, 10.0, 0.1)
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



# ---------------------------------------------------------------------------
# Served reactive ignition requests
# ---------------------------------------------------------------------------
R_UNIVERSAL = 8314.46  # J/(kmol K), the kinetics' gas constant (torchSolver.py:90-114)


def molar_mass(lam):
    """Mixture molar mass of the kinetics at reactant fraction lam."""
    return lam * (2.0 / 3.0) * 2.016 + lam * (1.0 / 3.0) * 32.0 + (1.0 - lam) * 18.016


def gamma_of(lam):
    return (4.5 - lam) / (3.5 - lam)


def induction_integral(theta, q, t_hat, rho0=1.0, p0=1.0, points=4001):
    """Time for a homogeneous reactor at the hot-spot peak to burn from
    lambda = 1 to 0.5 with unit pre-exponential factor.

    At rest and with constant density the internal energy is
    E(lam) = E0 + Q rho (1 - lam) and p = (gamma(lam) - 1) E, so the
    temperature and the rate are functions of lam alone and the induction
    time is the integral of dlam / (lam exp(-Ea / (8.314 T(lam)))) over
    [0.5, 1]; a pre-exponential factor A divides it by A.
    """
    t0 = p0 * molar_mass(1.0) / (rho0 * R_UNIVERSAL)
    ea = theta * 8.314 * t0
    q_heat = q * (R_UNIVERSAL / molar_mass(1.0)) * t0
    lam = np.linspace(0.5, 1.0, points)
    e0 = t_hat * p0 / (gamma_of(1.0) - 1.0)
    e = e0 + q_heat * rho0 * (1.0 - lam)
    p = (gamma_of(lam) - 1.0) * e
    temp = p * molar_mass(lam) / (rho0 * R_UNIVERSAL)
    f = 1.0 / (lam * np.exp(-ea / (8.314 * temp)))
    return float(np.sum(0.5 * (f[1:] + f[:-1]) * np.diff(lam)))


def served_requests(seed=355, count=64, t_run=0.1, znd_fraction=0.1, sizes=(32, 160)):
    """Seeded 3D ignition requests (parameters only; ``ignition_problem``
    builds the arrays). Nondimensional background (rho0, p0) = (1, 1) at
    lambda = 1; Ea/(R T0) in [20, 30]; Q/(R T0) in [10, 50]; a Gaussian
    hot spot with peak T/T0 in [1.5, 3] and width 0.05-0.15 of the box at a
    random centre in [0.25, 0.75]^3; ZND for a tenth of the requests; the
    pre-exponential factor puts the peak's induction time at 0.1 t_run."""
    rng = np.random.default_rng(seed)
    t0 = molar_mass(1.0) / R_UNIVERSAL
    out = []
    for i in range(count):
        n = int(rng.integers(sizes[0], sizes[1] + 1))
        theta = rng.uniform(20.0, 30.0)
        q = rng.uniform(10.0, 50.0)
        t_hat = rng.uniform(1.5, 3.0)
        sigma = rng.uniform(0.05, 0.15)
        centre = rng.uniform(0.25, 0.75, size=3)
        model = "znd" if rng.random() < znd_fraction else "arrhenius"
        tau = induction_integral(theta, q, t_hat)
        out.append(dict(index=i, n=n, theta=theta, q=q, t_hat=t_hat, sigma=sigma, centre=centre, model=model,
                        t_run=t_run, constants=dict(activation_energy=theta * 8.314 * t0,
                                                    q_heat=q * (R_UNIVERSAL / molar_mass(1.0)) * t0,
                                                    pre_exponential=tau / (0.1 * t_run))))
    return out


def ignition_problem(req):
    """Primitive state and lambda of one request on the unit cube."""
    n = req["n"]
    x, y, z = _centres([n, n, n])
    cx, cy, cz = req["centre"]
    r2 = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2
    prim = np.zeros((5, n, n, n))
    prim[0] = 1.0
    prim[4] = 1.0 + (req["t_hat"] - 1.0) * np.exp(-r2 / (2.0 * req["sigma"] ** 2))
    lam = np.ones((n, n, n))
    return prim, lam
