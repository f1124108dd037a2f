"""Limiter edge operands through the device step, against the oracle.

The limited slope phi(minus / plus) * plus of the reference
(torchSolver3D.py:86-111, torchSolver2D.py:124-168) divides by +1e-14
instead of plus when |plus| < 1e-14, so r can be huge, and Albada squares
it: for plus == 0 and minus above ~1.34e140 the reference's phi is
inf / inf = NaN, the predicted face is NaN and the whole step takes the
first-order fallback (torchSolver3D.py:160-170; 2D: the LLF retry,
torchSolver2D.py:412-462). The device's division-free limiter must take the
same decisions. Three states per dimension, every limiter, FP64 at the
1e-12 bar with the fallback flag exact:

- tiny: densities 1 + k 2^-52 (|k| <= 20), so every density slope has
  0 < |plus| < 1e-14 or plus == 0 and runs the literal quotient branch;
- huge: the right part of the box at p = 1e150 with equal neighbours, so
  plus == 0 with minus ~ 2.5e150 at the jump (Albada: NaN -> fallback;
  the other limiters stay finite);
- nan: one NaN operand; both sides raise the nonphysical-state error.
"""

import numpy as np
import pytest

from oracle import solver2d as o2
from oracle import solver3d as o3
from parity import assert_fields

pytestmark = pytest.mark.gpu

LIMITERS = ("VanLeer", "Albada", "MINBEE", "Superbee", "Unos")


def state3(case, seed=3):
    rng = np.random.default_rng(seed)
    nz, ny, nx = 5, 6, 14
    p = np.empty((5, nz, ny, nx))
    p[0] = 1.0 + rng.integers(-20, 21, (nz, ny, nx)) * 2.0 ** -52
    p[1:4] = rng.uniform(-0.1, 0.1, (3, nz, ny, nx))
    p[4] = rng.uniform(0.8, 1.2, (nz, ny, nx))
    if case == "huge":
        p[1:4, :, :, 7:] = 0.0
        p[4, :, :, 7:] = 1e150
        p[0, :, :, 7:] = 1.0
    u = o3.to_cons(p, 1.4)
    if case == "nan":
        u[3, 2, 3, 4] = np.nan
    return u


def state2(case, seed=4):
    rng = np.random.default_rng(seed)
    ny, nx = 7, 13
    p = np.empty((4, ny, nx))
    p[0] = 1.0 + rng.integers(-20, 21, (ny, nx)) * 2.0 ** -52
    p[1:3] = rng.uniform(-0.1, 0.1, (2, ny, nx))
    p[3] = rng.uniform(0.8, 1.2, (ny, nx))
    if case == "huge":
        p[1:3, :, 6:] = 0.0
        p[3, :, 6:] = 1e150
        p[0, :, 6:] = 1.0
    u = o2.to_cons(p, 1.4)
    if case == "nan":
        u[2, 3, 4] = np.nan
    return u


def device_hydro(u, spacing, dt, limiter, bcs_xyz):
    from torch_amd import _abi
    from torch_amd.torchExecution import DeviceExecution
    from torch_amd.torchSolver3D import config

    dim = u.ndim - 1
    counts = u.shape[1:][::-1]
    ex = DeviceExecution()
    cfg = config(counts, spacing, dim=dim, limiter=limiter, bcs=bcs_xyz, precision="fp64")
    dev = _abi.DeviceState(ex.context, cfg)
    try:
        dev.upload(u)
        info = dev.hydro(dt, want_fluxes=True)
        return dev.download()[0], dev.fluxes(), bool(info.fallback)
    finally:
        dev.close()
        ex.close()


@pytest.mark.parametrize("limiter", LIMITERS)
@pytest.mark.parametrize("case", ["tiny", "huge"])
def test_limiter_edges_3d(limiter, case):
    u = state3(case)
    spacing = (1 / 14, 1 / 6, 1 / 5)
    dt = o3.cfl_dt(u, 1.4, spacing, 0.4)
    with np.errstate(all="ignore"):
        ref, rflux, rfb = o3.hydro(u, 1.4, spacing, dt, limiter, ("Reflective", "Periodic", "Open"))
    out, flux, fb = device_hydro(u, spacing, dt, limiter, ("Open", "Periodic", "Reflective"))
    assert fb == rfb, f"{case}/{limiter}: device fallback {fb}, reference {rfb}"
    if case == "huge":
        assert rfb == (limiter == "Albada")
    assert_fields(out, ref, what=f"3D {case} {limiter}")
    for a, b, name in zip(flux, rflux, "xyz"):
        assert_fields(a, b, what=f"3D {case} {limiter} F{name}")


@pytest.mark.parametrize("limiter", LIMITERS)
@pytest.mark.parametrize("case", ["tiny", "huge"])
def test_limiter_edges_2d(limiter, case):
    u = state2(case)
    dx, dy = 1 / 13, 1 / 7
    dt = o2.cfl_dt(u, 1.4, dx, dy, 0.4, "Open", "Periodic")
    with np.errstate(all="ignore"):
        ref, rfx, rfy, rfb = o2.hydro(u, dx, dy, dt, 1.4, limiter, "Open", "Periodic")
    out, (fx, fy), fb = device_hydro(u, (dx, dy), dt, limiter, ("Open", "Periodic"))
    assert fb == rfb, f"{case}/{limiter}: device fallback {fb}, reference {rfb}"
    if case == "huge":
        assert rfb == (limiter == "Albada")
    assert_fields(out, ref, what=f"2D {case} {limiter}")
    assert_fields(fx, rfx, what=f"2D {case} {limiter} Fx")
    assert_fields(fy, rfy, what=f"2D {case} {limiter} Fy")


@pytest.mark.parametrize("limiter", LIMITERS)
def test_limiter_nan_operand_raises(limiter):
    u3 = state3("nan")
    with np.errstate(all="ignore"), pytest.raises(RuntimeError, match="3D hydrodynamic update produced"):
        o3.hydro(u3, 1.4, (1 / 14, 1 / 6, 1 / 5), 1e-3, limiter, ("Reflective", "Periodic", "Open"))
    with pytest.raises(RuntimeError, match="3D hydrodynamic update produced a nonphysical state"):
        device_hydro(u3, (1 / 14, 1 / 6, 1 / 5), 1e-3, limiter, ("Open", "Periodic", "Reflective"))
    u2 = state2("nan")
    with np.errstate(all="ignore"), pytest.raises(RuntimeError, match="Hydrodynamic update produced"):
        o2.hydro(u2, 1 / 13, 1 / 7, 1e-3, 1.4, limiter, "Open", "Periodic")
    with pytest.raises(RuntimeError, match="ydrodynamic update produced a nonphysical state"):
        device_hydro(u2, (1 / 13, 1 / 7), 1e-3, limiter, ("Open", "Periodic"))
