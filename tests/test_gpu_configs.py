"""The named workloads that are not bench lines, as parity cases.

- 1D Sod, 400 cells, Van Leer, CFL 0.9 (outside the reference's validated
  CFL range, run unchecked), open ends, t_end 0.2, FP64: every step against
  the oracle at 1e-12, and the density's L1 error against the exact Riemann
  solution.
- 2D quadrant Riemann problem (seeded four states), Superbee, CFL 0.4, open:
  FP64 against the oracle at 1e-12 and FP32 against the oracle at a size
  the oracle runs in seconds (128^2, 20 steps); at the full 1024^2, 100
  steps, FP32 against FP64 on the device, and the balance of mass and energy.
- 3D pressure sphere, Superbee, CFL 0.4, open: FP64 and FP32 against the
  oracle at 48^3 (10 steps); at the full 192^3, 100 steps, FP32 against
  FP64 (1e-5 relative L1 on every field, momenta included), the balances,
  and the sphere's symmetry (x, y and z exchanged;
  1e-10, since the update sums the directions in a fixed order).

The per-field errors are printed (the numbers stated for these workloads).
"""

import numpy as np
import pytest

from oracle import exact
from oracle import session as osession
from parity import assert_fields, l1_rel

pytestmark = pytest.mark.gpu


def pair(prim, lengths, steps, precision="fp64", oracle=True, **opts):
    from torch_amd import Simulation

    dev = Simulation(prim, lengths, precision=precision, **opts)
    ref = osession.Session(prim, lengths, **opts) if oracle else None
    for _ in range(steps):
        dev.step()
        if ref is not None:
            ref.step()
            assert dev.fallbacks == ref.fallbacks
            assert dev.time == pytest.approx(ref.time, rel=1e-12 if precision == "fp64" else 1e-5)
    return dev, ref


def balance(sim):
    """Conservation residual of every field (final - initial + net outflow),
    relative to the largest of the three terms."""
    scale = np.maximum.reduce([np.abs(sim.initial_integral), np.abs(sim.integral), np.abs(sim.boundary_integral)])
    return np.abs(sim.balance_residual) / np.maximum(scale, 1e-300)


def test_sod1d_400_cells_cfl09():
    from torch_amd import Simulation

    n = 400
    x = (np.arange(n) + 0.5) / n
    prim = np.array([np.where(x < 0.5, 1.0, 0.125), np.zeros(n), np.where(x < 0.5, 1.0, 0.1)])
    opts = dict(t_end=0.2, cfl=0.9, limiter="VanLeer", boundaries=["Open"], checked=False)
    dev = Simulation(prim, [1.0], **opts)
    ref = osession.Session(prim, [1.0], **opts)
    steps = 0
    while not ref.completed:
        ref.step()
        dev.step()
        steps += 1
        assert dev.time == pytest.approx(ref.time, rel=1e-13)
        assert_fields(dev.state, ref.state, what=f"1D Sod step {steps}")
    assert dev.completed and dev.fallbacks == ref.fallbacks
    rho = exact.sample(x, dev.time)[0]
    err = np.mean(np.abs(dev.state[0] - rho))
    print(f"1D Sod 400 CFL 0.9: {steps} steps, density L1 vs exact {err:.4e}")
    assert err < 0.02
    dev.close()


def test_quadrant2d_against_the_oracle():
    from torch_amd.workloads import quadrant_2d

    prim = quadrant_2d(128)
    opts = dict(t_end=10.0, cfl=0.4, limiter="Superbee", boundaries=["Open", "Open"])
    dev, ref = pair(prim, [1.0, 1.0], 20, **opts)
    assert_fields(dev.state, ref.state, what="quadrant FP64")
    dev32, _ = pair(prim, [1.0, 1.0], 20, precision="fp32", oracle=False, **opts)
    err = l1_rel(dev32.state, ref.state)
    print("quadrant 128^2 FP32 vs oracle L1 rel per field:", err)
    assert np.all(err < 1e-5)
    dev.close()
    dev32.close()


def test_quadrant2d_1024_fp32_against_fp64():
    from torch_amd.workloads import quadrant_2d

    prim = quadrant_2d(1024)
    opts = dict(t_end=10.0, cfl=0.4, limiter="Superbee", boundaries=["Open", "Open"])
    d64, _ = pair(prim, [1.0, 1.0], 100, oracle=False, **opts)
    d32, _ = pair(prim, [1.0, 1.0], 100, precision="fp32", oracle=False, **opts)
    err = l1_rel(d32.state, d64.state)
    print("quadrant 1024^2, 100 steps, FP32 vs FP64 L1 rel per field:", err,
          "balance FP64:", balance(d64), "FP32:", balance(d32))
    assert np.all(err < 1e-5)
    assert np.all(balance(d64)[[0, 3]] < 1e-12) and np.all(balance(d32)[[0, 3]] < 1e-5)
    d64.close()
    d32.close()


def test_sphere3d_against_the_oracle():
    from torch_amd.workloads import sphere_3d

    prim = sphere_3d(48)
    opts = dict(t_end=10.0, cfl=0.4, limiter="Superbee", boundaries=["Open"] * 3)
    dev, ref = pair(prim, [1.0] * 3, 10, **opts)
    assert_fields(dev.state, ref.state, what="sphere FP64")
    dev32, _ = pair(prim, [1.0] * 3, 10, precision="fp32", oracle=False, **opts)
    err = l1_rel(dev32.state, ref.state)
    print("sphere 48^3 FP32 vs oracle L1 rel per field:", err)
    assert np.all(err < 1e-5)
    dev.close()
    dev32.close()


def test_sphere3d_192_fp32_against_fp64_and_symmetry():
    from torch_amd.workloads import sphere_3d

    prim = sphere_3d(192)
    opts = dict(t_end=10.0, cfl=0.4, limiter="Superbee", boundaries=["Open"] * 3)
    d64, _ = pair(prim, [1.0] * 3, 100, oracle=False, **opts)
    d32, _ = pair(prim, [1.0] * 3, 100, precision="fp32", oracle=False, **opts)
    s64, s32 = d64.state, d32.state
    err = l1_rel(s32, s64)
    print("sphere 192^3, 100 steps, FP32 vs FP64 L1 rel per field:", err,
          "balance FP64:", balance(d64), "FP32:", balance(d32))
    assert np.all(err < 1e-5)
    assert np.all(balance(d64)[[0, 4]] < 1e-12)
    # State axes are (z, y, x): exchanging two axes exchanges the matching momenta.
    rho, e = s64[0], s64[4]
    for axes in ((2, 1, 0), (1, 0, 2), (0, 2, 1)):
        sym = max(np.max(np.abs(rho - rho.transpose(axes))) / np.max(rho),
                  np.max(np.abs(e - e.transpose(axes))) / np.max(e))
        print("sphere symmetry", axes, sym)
        assert sym < 1e-10, (axes, sym)
    mx, my, mz = s64[1], s64[2], s64[3]
    assert np.max(np.abs(mx - my.transpose(0, 2, 1))) <= 1e-10 * np.max(np.abs(mx))
    assert np.max(np.abs(mx - mz.transpose(2, 1, 0))) <= 1e-10 * np.max(np.abs(mx))
    d64.close()
    d32.close()
