"""In-place edits of a session's arrays reach the device.

The reference's Simulation.state, lam and beta are plain arrays that callers
edit in place (torchSession.py:145-158; testAdaptive.py:127-128 sets
``s.state[:]`` and ``s.lam[:]`` before a regrid, :288 edits ``s.state``
before a save). The device session hands out host copies, so an edit must
be uploaded before the next device operation: a step after an edit equals
the oracle's step after the same edit, to the FP64 bar.
"""

import numpy as np
import pytest

from oracle import presets as opresets
from oracle import session as osession
from parity import assert_fields

pytestmark = pytest.mark.gpu


def test_edited_state_steps_like_the_oracle():
    from torch_amd import Simulation

    prim, opt = opresets.preset("3d_explosion", [12, 10, 8], limiter="Superbee", t_end=0.2)
    opt.pop("reactive")
    ref = osession.Session(prim, [1.0] * 3, **opt)
    dev = Simulation(prim, [1.0] * 3, **opt)
    try:
        # Edit before any step (the uploaded copy), then after steps (a download).
        for edit in range(3):
            for sim in (ref, dev):
                s = sim.state
                s[0, 2, 3, 4] *= 1.25
                s[4, :, 1, :] += 0.05 * (edit + 1)
            np.testing.assert_allclose(dev.integral, ref.integral, rtol=1e-12,
                                       atol=1e-14 * np.max(np.abs(ref.integral)))
            assert dev.dt_limit() == pytest.approx(ref.dt_limit(ref.state, ref.gamma), rel=1e-13)
            for _ in range(2):
                ref.step()
                dev.step()
            assert dev.time == pytest.approx(ref.time, rel=1e-13)
            assert_fields(dev.state, ref.state, what=f"after edit {edit}")
        # Replacing the array through the setter keeps working.
        ref.state = ref.state * 1.0
        dev.state = dev.state * 1.0
        ref.step()
        dev.step()
        assert_fields(dev.state, ref.state, what="after setter")
    finally:
        dev.close()


def test_edited_fraction_steps_like_the_oracle():
    from torch_amd import Simulation

    d = np.load(__import__("conftest").GOLDEN + "/session.npz")
    prim = d["ignition3d_prim"]
    opts = dict(t_end=1e-5, reactive=True, kinetic_model="znd", boundaries=["Open", "Periodic", "Reflective"])
    lam0 = d["ignition3d_znd_lam0"]
    ref = osession.Session(prim, [1.0] * 3, lambda_init=lam0, **opts)
    dev = Simulation(prim, [1.0] * 3, lambda_init=lam0, **opts)
    try:
        ref.step()
        dev.step()
        for sim in (ref, dev):
            sim.lam[:, :, : sim.lam.shape[2] // 2] = 0.5
            sim.beta[0] = 0.25
        for _ in range(2):
            ref.step()
            dev.step()
        assert_fields(dev.state, ref.state, what="state")
        assert_fields(dev.lam.ravel(), ref.lam.ravel(), what="lambda")
        assert_fields(dev.beta.ravel(), ref.beta.ravel(), what="beta")
    finally:
        dev.close()


def test_adaptive_edit_then_regrid_coarsens():
    """testAdaptive.py:111-137 on the device: a constant state and fraction
    written in place make every complete sibling group eligible."""
    from torch_amd import Simulation

    for dim in (1, 2, 3):
        counts = [8, 4, 4][:dim]
        p = np.ones((dim + 2, *counts[::-1]))
        p[1:-1] = 0
        p[0] += 0.1 * np.arange(counts[0]) / counts[0]
        s = Simulation(p, [1.0] * dim, t_end=0.02, reactive=True).to_adaptive(max_level=2, refine_threshold=1)
        try:
            s.regrid(refine=np.ones(len(s.mesh.keys), dtype=bool), coarsen=False)
            s.state[:] = s.state[:, :1]
            s.lam[:] = 0.37
            before = s.integral.copy()
            species = np.sum(s.state[0] * s.lam * s.mesh.volumes)
            count = len(s.mesh.keys)
            s.regrid(refine=np.zeros(len(s.mesh.keys), dtype=bool))
            assert len(s.mesh.keys) < count, dim
            np.testing.assert_allclose(s.integral, before, rtol=5e-15)
            assert np.sum(s.state[0] * s.lam * s.mesh.volumes) == pytest.approx(species, rel=1e-14)
            s.step()
            np.testing.assert_allclose(s.primitive[1:-1], 0, atol=1e-13)
        finally:
            s.close()
