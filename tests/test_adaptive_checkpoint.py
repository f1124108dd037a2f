"""Adaptive (schema 3) checkpoints on the host: the datasets written, the
validating reader and its rejections (torchAdaptiveIO.py:86-300,
testAdaptive.py:272-322). Sessions here are host-side shells (mesh, state,
counters) built from the reference's adaptive runs, so no device is needed;
exact restart on the device is in test_gpu_adaptive.py."""

import copy
import json

import numpy as np
import pytest

from conftest import golden
from test_adaptive_host import HostLeaves, run_config


def shell(tag, steps=2):
    from torch_amd.torchAdaptive import AdaptiveSimulation

    d = golden("adaptive")
    c, p = run_config(d, tag)
    run = HostLeaves(c, p)
    for _ in range(steps):
        run.step()
    s = AdaptiveSimulation.__new__(AdaptiveSimulation)
    s._device = s._leaves = s._leaves_key = s._mesh_on_device = s._execution = None
    c = copy.deepcopy(c)
    c.update(limiter="VanLeer", controller="Sound Speed Controlled", variant="Variante 1 - Original",
             wavespeed="Pressure Based", inlet_primitive=[1.0, 0.0, 100.0, 101325.0])
    s.config, s.mesh = c, run.mesh
    s.precision, s.checked, s.constants, s.tile = "fp64", True, {}, (0, 0, 0)
    lam = run.fractions[0] if c["reactive"] else None
    beta = run.fractions[1] if len(run.fractions) > 1 else None
    s._host = (run.state, lam, beta)
    s.time, s.steps, s.run_id, s.fallbacks = run.time, run.steps, "run-1", 0
    s.regrids, s.mesh_revision, s.budget_limited = run.regrids, 1, False
    s.initial_integral = np.sum(run.state * run.mesh.volumes, axis=1) + 0.5
    s.boundary_integral, s.source_integral = run.boundary.copy(), run.source.copy()
    return s


@pytest.mark.parametrize("tag", ["sod2d", "znd2d", "stag3d"])
def test_adaptive_checkpoint_round_trip(tag):
    from torch_amd.torchIO import content, record

    s = shell(tag)
    data = content(s)
    meta = json.loads(bytes(data["metadata"]).decode())
    assert meta["schema_version"] == 3 and meta["layout"] == "leaf-level-z-y-x"
    assert data["mesh/levels"].dtype == np.dtype("<i8") and data["mesh/indices"].dtype == np.dtype("<i8")
    back = record(data)
    np.testing.assert_array_equal(back.mesh.keys, s.mesh.keys)
    for a, b in zip(back._host, s._host):
        if b is None:
            assert a is None
        else:
            np.testing.assert_array_equal(a, b)
    for key in ("initial_integral", "boundary_integral", "source_integral"):
        np.testing.assert_array_equal(getattr(back, key), getattr(s, key))
    assert (back.time, back.steps, back.regrids, back.mesh_revision, back.run_id) == (
        s.time, s.steps, s.regrids, s.mesh_revision, s.run_id)


@pytest.mark.parametrize("mode", ["layout", "composition", "counter", "integral", "indices", "noncanonical"])
def test_corrupt_adaptive_checkpoints_are_rejected(mode):
    """testAdaptive.py:293-322 on the NumPy container."""
    from torch_amd.torchIO import content, record

    s = shell("znd2d" if mode == "composition" else "sod2d")
    data = dict(content(s))
    meta = json.loads(bytes(data["metadata"]).decode())
    if mode == "layout":
        meta["layout"] = "x-fastest"
    if mode == "composition":
        del data["state/lambda"]
    if mode == "counter":
        meta["mesh_revision"] = meta["regrids"] + 1
    if mode == "integral":
        meta["current_integral"] = [1.0]
    if mode == "indices":
        data["mesh/indices"] = data["mesh/indices"].copy()
        data["mesh/indices"][0] = 1
    if mode == "noncanonical":
        order = np.arange(len(data["mesh/levels"]))[::-1]
        d = s.config["dimension"]
        data["mesh/levels"] = data["mesh/levels"][order]
        data["mesh/indices"] = data["mesh/indices"].reshape(-1, d)[order].ravel()
    data["metadata"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    with pytest.raises(ValueError):
        record(data)


def test_saving_a_nonphysical_state_is_refused_and_leaves_the_file(tmp_path):
    from torch_amd.torchIO import write_checkpoint

    s = shell("sod2d")
    path = tmp_path / "state.npz"
    write_checkpoint(path, s)
    old = path.read_bytes()
    s._host[0][0, 0] = -1.0
    with pytest.raises(ValueError):
        write_checkpoint(path, s)
    assert path.read_bytes() == old
