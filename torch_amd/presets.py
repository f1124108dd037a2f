"""Initial states of the reference's benchmark presets (host side).

The preset tables and builders of python/cfdSolver/torchSolver3D.py:259-286
(3D), torchSolver2D.py:538-639 (2D), torchSolver.py:817-910 and
torchSession.py:653-694 (1D and the session's divider rule), used by
``Simulation.preset``.
"""

import numpy as np

PRESETS_1D = {
    # name: (t_end, cfl, rho_l, u_l, p_l, rho_r, u_r, p_r, limiter, reactive)
    "sod_standard": (0.2, 0.4, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, "VanLeer", False),
    "detonation_cj": (0.01, 0.4, 1.0, 0.0, 50.0 * 101325.0, 0.125, 0.0, 0.1 * 101325.0, "VanLeer", True),
    "rarefaction_wave": (0.15, 0.4, 1.0, -2.0, 0.4, 1.0, 2.0, 0.4, "Albada", False),
    "strong_shock": (0.015, 0.4, 1.0, 0.0, 1000.0, 1.0, 0.0, 0.01, "VanLeer", False),
    "toro_test4": (0.035, 0.4, 1.0, 0.0, 0.01, 1.0, 0.0, 100.0, "VanLeer", False),
}


def preset_1d(name, n):
    t_end, cfl, rl, ul, pl, rr, ur, pr, limiter, reactive = PRESETS_1D[name]
    divider = max(1, min(n - 1, int(np.round(0.5 * n))))
    p = np.empty((3, n))
    p[:, :divider] = np.array([rl, ul, pl])[:, None]
    p[:, divider:] = np.array([rr, ur, pr])[:, None]
    return p, dict(t_end=t_end, cfl=cfl, limiter=limiter, reactive=reactive)


def preset_2d(name, nx, ny):
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    rho, u, v, p = np.ones((ny, nx)), np.zeros((ny, nx)), np.zeros((ny, nx)), np.ones((ny, nx))
    reactive = False
    key = name.lower()
    if key in ("2d_explosion", "cylindrical_explosion"):
        r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
        rho, p = np.where(r <= 0.2, 1.0, 0.125), np.where(r <= 0.2, 1.0, 0.1)
        t_end, bcs = 0.25, ["Open", "Open"]
    elif key in ("2d_implosion", "implosion"):
        r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
        rho, p = np.where(r <= 0.2, 0.125, 1.0), np.where(r <= 0.2, 0.14, 1.0)
        t_end, bcs = 0.35, ["Reflective", "Reflective"]
    elif key in ("2d_riemann_quadrant", "riemann_4quadrant"):
        q1, q2 = (X >= 0.5) & (Y >= 0.5), (X < 0.5) & (Y >= 0.5)
        q3 = (X < 0.5) & (Y < 0.5)
        rho = np.where(q1, 0.5313, np.where(q2, 1.0, np.where(q3, 0.8, 1.0)))
        u = np.where(q2, 0.7276, 0.0)
        v = np.where(q1 | q2 | q3, 0.0, 0.7276)
        p = np.where(q1, 0.4, 1.0)
        t_end, bcs = 0.25, ["Open", "Open"]
    elif key in ("2d_shock_tube", "planar_shock"):
        rho, p = np.where(X <= 0.5, 1.0, 0.125), np.where(X <= 0.5, 1.0, 0.1)
        t_end, bcs = 0.2, ["Open", "Reflective"]
    elif key in ("2d_rde_chamber", "rde_detonation"):
        d = np.sqrt((X - 0.2) ** 2 + (Y - 0.1) ** 2)
        rho = np.where(d <= 0.1, 2.5, 1.0)
        u = np.where(d <= 0.1, 500.0, 0.0)
        p = np.where(d <= 0.1, 30.0, 1.0) * 101325.0
        t_end, bcs, reactive = 0.005, ["Periodic", "Inflow"], True
    else:
        raise ValueError(f"Unknown 2D preset: {name}")
    return np.stack([rho, u, v, p]), dict(t_end=t_end, boundaries=bcs, reactive=reactive)


PRESETS_3D = ("3d_sod_x", "3d_sod_y", "3d_sod_z", "3d_explosion", "3d_ignition")


def get_3d_preset(name, nx=20, ny=20, nz=20):
    """torchSolver3D.py:259-286: primitive, reactive flag and end time."""
    if name not in PRESETS_3D:
        raise ValueError(f"Unknown 3D preset: {name}")
    prim, d = preset_3d(name, nx, ny, nz)
    return {"primitive": prim, "is_reactive": d["reactive"], "t_end": d["t_end"]}


def preset_3d(name, nx, ny, nz):
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny,
                          (np.arange(nx) + 0.5) / nx, indexing="ij")
    prim = np.zeros((5, nz, ny, nx))
    if name.startswith("3d_sod_"):
        c = {"x": x, "y": y, "z": z}[name[-1]]
        prim[0], prim[4] = np.where(c <= 0.5, 1, 0.125), np.where(c <= 0.5, 1, 0.1)
    elif name == "3d_explosion":
        inside = (x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2 <= 0.2**2
        prim[0], prim[4] = np.where(inside, 1, 0.125), np.where(inside, 1, 0.1)
    elif name == "3d_ignition":
        prim[0] = 1
        prim[4] = np.where(x <= 0.5, 1e6, 101325)
    else:
        raise ValueError(f"Unknown 3D preset: {name}")
    ign = name == "3d_ignition"
    return prim, dict(t_end=1e-5 if ign else 0.05, reactive=ign)


def preset(name, counts, **options):
    """Build (primitive, defaults) as ``Simulation.preset`` (torchSession.py:653-694)."""
    counts = list(np.atleast_1d(counts))
    if len(counts) == 1:
        prim, defaults = preset_1d(name, counts[0])
    elif len(counts) == 2:
        prim, defaults = preset_2d(name, *counts)
    else:
        prim, defaults = preset_3d(name, *counts)
    defaults.update(options)
    return prim, defaults
