"""Checkpoint layout and validation on the host (no device).

The datasets and checks follow python/cfdSolver/torchIO.py:66-119,
192-214, 233-391: ``metadata`` is a UTF-8 JSON byte vector,
``state/<component>`` are flat float64 vectors with x fastest, and the
reader rejects malformed records with ValueError before any device state is
made. The session shells here come from the reader itself (host arrays, no
device), so these tests run without a GPU.
"""

import json

import numpy as np
import pytest

from torch_amd import torchIO


def _record(dim=3, reactive=False, model="arrhenius", **meta_over):
    """Datasets of a valid checkpoint built by hand: 6x4x4 (3D), 8x6, 24."""
    counts = [24] if dim == 1 else ([8, 6] if dim == 2 else [6, 4, 4])
    shape = tuple(counts[::-1])
    rng = np.random.default_rng(7)
    state = np.empty((dim + 2, *shape))
    state[0] = 1 + 0.1 * rng.random(shape)
    state[1:-1] = 0.2 * state[0]
    state[-1] = 2.5 + 0.5 * np.sum(state[1:-1] ** 2, axis=0) / state[0]
    config = dict(dimension=dim, counts_xyz=counts, lengths_xyz=[1.0] * dim, boundaries_xyz=["Periodic"] * dim,
                  t_end=0.08, cfl=0.4, gamma=1.4, limiter="VanLeer", reactive=reactive, kinetic_model=model,
                  inlet_primitive=[1.0, 0.0, 100.0, 101325.0], controller="Sound Speed Controlled",
                  variant="Variante 1 - Original", wavespeed="Pressure Based", max_steps=100000,
                  geometry="cartesian", origins_xyz=[0.0] * dim)
    volume = np.prod([1.0 / n for n in counts])
    total = state.reshape(dim + 2, -1).sum(axis=1) * volume
    meta = dict(format="TORCH checkpoint", schema_version=2, solver_version="1.4", geometry="cartesian",
                units="SI", layout="x-fastest", config=config, run_id="run-1", time=0.01, steps=3,
                initial_integral=total.tolist(), boundary_integral=[0.0] * (dim + 2),
                source_integral=[0.0] * (dim + 2), current_integral=total.tolist())
    meta.update(meta_over)
    data = {"metadata": np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)}
    names = ["rho"] + ["momentum_" + a for a in "xyz"[:dim]] + ["energy"]
    for name, row in zip(names, state):
        data["state/" + name] = row.ravel().astype("<f8")
    if reactive:
        data["state/lambda"] = np.full(state[0].size, 0.5)
        if model == "znd":
            data["state/beta"] = np.ones(state[0].size)
    return data, state


def _meta(data):
    return json.loads(bytes(data["metadata"]).decode())


def _with_meta(data, **changes):
    meta = _meta(data)
    meta.update(changes)
    out = dict(data)
    out["metadata"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    return out


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_round_trip_keeps_layout_and_counters(tmp_path, dim):
    data, state = _record(dim)
    sim = torchIO.record(data)
    np.testing.assert_array_equal(sim.state, state)
    assert (sim.time, sim.steps, sim.run_id) == (0.01, 3, "run-1")
    path = tmp_path / "c.npz"
    torchIO.write_checkpoint(path, sim)
    again = torchIO.load_content(path)
    assert set(again) == set(data)
    for name in data:
        if name != "metadata":
            assert again[name].dtype == np.dtype("<f8") and again[name].ndim == 1
            np.testing.assert_array_equal(again[name], data[name])
    # x varies fastest: consecutive entries of a row are neighbours along x.
    np.testing.assert_array_equal(again["state/rho"][: state.shape[-1]], state[0].reshape(-1)[: state.shape[-1]])
    meta = _meta(again)
    assert meta["schema_version"] == 2 and meta["solver_version"] == "1.4" and meta["layout"] == "x-fastest"
    assert meta["device"] == dict(precision="fp64", constants={}, checked=True)
    np.testing.assert_array_equal(torchIO.record(again).state, state)


def test_reactive_fractions_travel(tmp_path):
    data, _ = _record(3, reactive=True, model="znd")
    sim = torchIO.record(data)
    assert sim.lam.shape == (4, 4, 6) and np.all(sim.beta == 1)
    torchIO.write_checkpoint(tmp_path / "z.npz", sim)
    assert {"state/lambda", "state/beta"} <= set(torchIO.load_content(tmp_path / "z.npz"))


def test_older_families_are_accepted_and_upgraded():
    data, _ = _record(1)
    sim = torchIO.record(_with_meta(data, solver_version="1.2"))
    assert _meta(torchIO.content(sim))["solver_version"] == "1.4"
    legacy = _meta(data)
    legacy.update(schema_version=1, solver_version="1.1")
    del legacy["config"]["geometry"], legacy["config"]["origins_xyz"]
    old = dict(data)
    old["metadata"] = np.frombuffer(json.dumps(legacy).encode(), dtype=np.uint8)
    torchIO.record(old)


@pytest.mark.parametrize("change", [
    dict(format="other"), dict(units="cgs"), dict(layout="z-fastest"), dict(schema_version=2, solver_version="1.3"),
    dict(schema_version=1, solver_version="1.1"),  # schema 1 with a geometry key
    dict(time=1.0), dict(time=-1e-3), dict(steps=True), dict(steps=-1), dict(run_id=""),
    dict(boundary_integral=[0.0, 0.0]), dict(current_integral=[1.0] * 5), dict(geometry="spherical"),
])
def test_malformed_metadata_is_rejected(change):
    data, _ = _record(3)
    with pytest.raises(ValueError):
        torchIO.record(_with_meta(data, **change))


def test_malformed_datasets_are_rejected():
    data, _ = _record(3)
    for bad in (
        {**data, "state/rho": data["state/rho"].astype("<f4")},
        {**data, "state/rho": data["state/rho"][:-1]},
        {**data, "state/lambda": np.ones(96)},
        {k: v for k, v in data.items() if k != "state/energy"},
        {k: v for k, v in data.items() if k != "metadata"},
        {**data, "metadata": data["metadata"].astype(np.int16)},
        {**data, "state/rho": -data["state/rho"]},
    ):
        with pytest.raises(ValueError):
            torchIO.record(bad)
    reactive, _ = _record(3, reactive=True)
    with pytest.raises(ValueError):
        torchIO.record({**reactive, "state/lambda": np.full(96, 1.5)})


def test_hdf5_container_needs_h5py(tmp_path):
    try:
        import h5py  # noqa: F401
        pytest.skip("h5py is installed")
    except ImportError:
        pass
    sim = torchIO.record(_record(2)[0])
    with pytest.raises(ValueError, match="h5py"):
        torchIO.write_checkpoint(tmp_path / "c.h5", sim)
    assert not list(tmp_path.iterdir())


def test_interrupted_write_keeps_the_previous_checkpoint(tmp_path, monkeypatch):
    sim = torchIO.record(_record(2)[0])
    path = tmp_path / "c.npz"
    torchIO.write_checkpoint(path, sim)
    before = path.read_bytes()

    def fail(*args, **kwargs):
        raise OSError("disk full")

    monkeypatch.setattr(np, "savez", fail)
    with pytest.raises(OSError):
        torchIO.write_checkpoint(path, sim)
    assert path.read_bytes() == before
    assert [p.name for p in tmp_path.iterdir()] == ["c.npz"]
