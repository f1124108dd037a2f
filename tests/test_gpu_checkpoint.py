"""Checkpoint and exact restart of device sessions.

The contract of the reference's testSession.py:80-121: a run paused after a
few steps, checkpointed, read back and run on ends bitwise where the
uninterrupted run ends (state, composition, balances, step count, time), in
every dimension and chemistry model. Here also in the FP32 form, and for a
checkpoint resumed inside a device batch.
"""

import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODELS = ("inert", "arrhenius", "znd")


def example(dim, model="inert", precision="fp64"):
    """The reference's restart example (testSession.py:38-54)."""
    from torch_amd import Simulation

    reactive = model != "inert"
    counts = [24] if dim == 1 else ([8, 6] if dim == 2 else [6, 4, 4])
    p = np.ones((dim + 2, *counts[::-1]))
    p[1:-1] = 0.2
    x = (np.arange(counts[0]) + 0.5) / counts[0]
    p[0] = 1 + 0.1 * np.sin(2 * np.pi * x)
    p[-1] = np.where(x < 0.5, 1e6, 101325) if reactive else 1
    return Simulation(p, [1.0] * dim, t_end=1e-4 if reactive else 0.08, boundaries=["Periodic"] * dim,
                      reactive=reactive, kinetic_model="arrhenius" if model == "inert" else model,
                      precision=precision)


def assert_same_run(a, b):
    np.testing.assert_array_equal(a.state, b.state)
    for name in ("lam", "beta"):
        x, y = getattr(a, name), getattr(b, name)
        assert (x is None) == (y is None)
        if x is not None:
            np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a.boundary_integral, b.boundary_integral)
    np.testing.assert_array_equal(a.source_integral, b.source_integral)
    np.testing.assert_array_equal(a.initial_integral, b.initial_integral)
    assert a.steps == b.steps and a.time == b.time


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_restart_is_exact_in_all_dimensions_and_models(tmp_path, dim, precision):
    from torch_amd import read_checkpoint

    for model in MODELS:
        a = example(dim, model, precision)
        continuous = copy.deepcopy(a)
        continuous.run()
        assert continuous.completed and continuous.steps > 3
        path = tmp_path / f"checkpoint_{model}.npz"
        a.run(step_limit=2, checkpoint=path)
        assert not a.completed and a.steps == 2
        restarted = read_checkpoint(path)
        assert restarted.precision == precision and restarted.run_id == a.run_id
        restarted.run()
        assert_same_run(restarted, continuous)
        for sim in (a, continuous, restarted):
            sim.close()


def test_periodic_checkpoints_and_later_end_time(tmp_path):
    from torch_amd import Simulation

    a = example(3, "znd")
    path = tmp_path / "every.npz"
    a.run(step_limit=3, checkpoint=path, checkpoint_every=2)
    assert a.steps == 3 and Simulation.load(path).steps == 3  # the final write follows the periodic one
    continuous = copy.deepcopy(a)
    continuous.config["t_end"] = 2e-4
    continuous.run()
    extended = Simulation.load(path, t_end=2e-4)
    assert extended.config["t_end"] == 2e-4
    extended.run()
    assert continuous.completed
    assert_same_run(extended, continuous)
    with pytest.raises(ValueError):
        Simulation.load(path, t_end=0.0)


def test_solver_family_12_loads_and_is_upgraded_on_save(tmp_path):
    """testSession.py:55-78 with the .npz container."""
    import json

    from torch_amd import Simulation
    from torch_amd.torchIO import load_content

    original = example(1)
    original.run(step_limit=3)
    path = tmp_path / "family12.npz"
    original.save(path)
    data = load_content(path)
    meta = json.loads(bytes(data["metadata"]).decode())
    assert meta["solver_version"] == "1.4"
    meta["solver_version"] = "1.2"
    data["metadata"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez(path, **data)
    loaded = Simulation.load(path)
    np.testing.assert_array_equal(loaded.state, original.state)
    np.testing.assert_array_equal(loaded.boundary_integral, original.boundary_integral)
    loaded.run()
    loaded.save(path)
    assert json.loads(bytes(load_content(path)["metadata"]).decode())["solver_version"] == "1.4"


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_checkpoint_resumes_inside_a_batch(tmp_path, precision):
    """A checkpoint admitted to a device batch ends where the session ends alone."""
    from torch_amd import read_checkpoint
    from torch_amd.torchBatch import Batch

    for dim, model in ((3, "znd"), (3, "arrhenius"), (2, "inert"), (1, "inert")):
        a = example(dim, model, precision)
        continuous = copy.deepcopy(a)
        continuous.run()
        path = tmp_path / f"batch_{dim}_{model}.npz"
        a.run(step_limit=2, checkpoint=path)
        resumed = read_checkpoint(path)
        batch = Batch(precision=precision)
        ident = batch.add(resumed)
        batch.run(continuous.steps + 5)
        info = batch.info(ident)
        assert info["completed"] and info["status"] == 0
        batch.commit(ident, resumed)
        batch.close()
        assert_same_run(resumed, continuous)


# Metadata keys of the reference's writer (torchIO.py:66-119); the device
# writer adds one record, "device", which the reference's reader ignores.
REFERENCE_KEYS = {"format", "schema_version", "solver_version", "geometry", "units", "layout", "config", "run_id",
                  "time", "steps", "initial_integral", "boundary_integral", "source_integral", "current_integral"}


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_device_checkpoint_resumed_by_the_oracle(tmp_path, dim):
    """A checkpoint the device wrote, read with plain NumPy (no device code),
    resumed by the oracle's session: it ends within 1e-12 of the device's
    uninterrupted run (testSession.py:80-121 across the two implementations),
    and its layout is the reference writer's (torchIO.py:66-119, 192-214)."""
    import json

    from oracle import session as osession
    from parity import assert_fields

    for model in MODELS:
        a = example(dim, model)
        continuous = copy.deepcopy(a)
        continuous.run()
        path = tmp_path / f"device_{model}.npz"
        a.run(step_limit=2, checkpoint=path)
        a.close()
        with np.load(path, allow_pickle=False) as f:
            data = {k: f[k] for k in f.files}
        meta = json.loads(data["metadata"].tobytes().decode("utf-8"))
        assert REFERENCE_KEYS <= set(meta) and set(meta) - REFERENCE_KEYS <= {"device"}
        assert meta["format"] == "TORCH checkpoint" and meta["layout"] == "x-fastest"
        c = meta["config"]
        axes = "xyz"[:dim]
        names = ["rho"] + ["momentum_" + ax for ax in axes] + ["energy"]
        want = {"metadata"} | {"state/" + n for n in names}
        if c["reactive"]:
            want |= {"state/lambda"} | ({"state/beta"} if c["kinetic_model"] == "znd" else set())
        assert set(data) == want
        shape = tuple(c["counts_xyz"][::-1])
        state = np.stack([data["state/" + n].reshape(shape) for n in names])
        assert all(data[k].dtype == np.dtype("<f8") for k in data if k != "metadata")
        ref = osession.Session(
            np.ones((dim + 2,) + shape), c["lengths_xyz"], t_end=c["t_end"], cfl=c["cfl"], gamma=c["gamma"],
            boundaries=c["boundaries_xyz"], limiter=c["limiter"], reactive=c["reactive"],
            kinetic_model=c["kinetic_model"], max_steps=c["max_steps"])
        ref.state = state
        if c["reactive"]:
            ref.lam = data["state/lambda"].reshape(shape)
            ref.beta = data["state/beta"].reshape(shape) if c["kinetic_model"] == "znd" else None
        ref.time, ref.steps = meta["time"], meta["steps"]
        ref.initial_integral = np.array(meta["initial_integral"])
        ref.boundary_integral = np.array(meta["boundary_integral"])
        ref.source_integral = np.array(meta["source_integral"])
        assert ref.steps == 2
        np.testing.assert_allclose(ref.integral, meta["current_integral"], rtol=1e-15, atol=1e-300)
        while not ref.completed:
            ref.step()
        assert ref.steps == continuous.steps and ref.time == pytest.approx(continuous.time, rel=1e-14)
        assert_fields(continuous.state, ref.state, what=f"{dim}D {model} resumed by the oracle")
        for name in ("lam", "beta"):
            x, y = getattr(continuous, name), getattr(ref, name)
            assert (x is None) == (y is None)
            if x is not None:
                assert np.max(np.abs(x - y)) <= 1e-12
        # Periodic boundaries: the outflow sums are rounding residues of
        # cancelling face sums, so they are held to the totals' scale.
        scale = np.max(np.abs(continuous.initial_integral))
        np.testing.assert_allclose(continuous.boundary_integral, ref.boundary_integral, rtol=1e-11, atol=1e-12 * scale)
        np.testing.assert_allclose(continuous.source_integral, ref.source_integral, rtol=1e-11, atol=1e-12 * scale)
        continuous.close()
