"""Batches on the device: every instance ends where the same session ends
stepped alone (bitwise in FP64), whatever else is in the batch, when it was
admitted and which 3D tile the autotuner picked."""

import copy

import numpy as np
import pytest

from oracle.session import Session as OracleSession
from tests.parity import assert_fields

pytestmark = pytest.mark.gpu


def _two_state(dim, counts, rng, gamma):
    """Random planar or spherical two-state Riemann data (primitive layout)."""
    shape = tuple(counts[::-1])
    axes = np.meshgrid(*[(np.arange(n) + 0.5) / n for n in counts[::-1]], indexing="ij")
    xyz = axes[::-1]
    if rng.random() < 0.5:
        inside = xyz[int(rng.integers(dim))] < 0.5
    else:
        r2 = sum((x - 0.5) ** 2 for x in xyz)
        inside = r2 < rng.uniform(0.1, 0.3) ** 2
    prim = np.empty((dim + 2,) + shape)
    for mask in (inside, ~inside):
        rho, p = rng.uniform(0.1, 2.0), rng.uniform(0.05, 5.0)
        v = rng.normal(size=dim)
        v *= rng.uniform(0, 1) / max(np.linalg.norm(v), 1e-12)
        prim[0][mask] = rho
        for k in range(dim):
            prim[1 + k][mask] = v[k]
        prim[-1][mask] = p
    return prim


def _rough(dim, counts, rng):
    """Random cell-by-cell state that drives the first-order fallback."""
    shape = tuple(counts[::-1])
    prim = np.empty((dim + 2,) + shape)
    prim[0] = 10 ** rng.uniform(-2, 1, shape)
    for k in range(1, dim + 1):
        prim[k] = rng.uniform(-3, 3, shape)
    prim[-1] = 10 ** rng.uniform(-3, 1, shape)
    return prim


SPECS = [
    # dim, counts, limiter, boundaries, rough, extra options
    (1, [300], "VanLeer", ["Open"], False, {}),
    (1, [257], "Unos", ["Reflective"], False, {"wavespeed": "Enthalpy Based"}),
    (1, [200], "Superbee", ["Periodic"], False, {"variant": "Variante_2", "controller": "w=0"}),
    (1, [90], "MINBEE", ["Open"], True, {}),
    (2, [40, 33], "MINBEE", ["Open", "Reflective"], False, {}),
    (2, [64, 20], "Albada", ["Periodic", "Inflow"], False, {"inlet_primitive": [0.9, 0.05, 0.3, 0.8]}),
    (2, [16, 18], "Superbee", ["Open", "Open"], True, {}),
    (3, [20, 14, 11], "Superbee", ["Open", "Reflective", "Periodic"], False, {}),
    (3, [70, 9, 37], "VanLeer", ["Reflective"] * 3, False, {}),
    (3, [33, 33, 33], "Unos", ["Open"] * 3, False, {}),
    (3, [9, 8, 10], "Superbee", ["Open"] * 3, True, {}),
]


def _sessions(precision, seed=5, specs=SPECS, t_end=10.0):
    from torch_amd.torchSession import Simulation
    rng = np.random.default_rng(seed)
    out = []
    for dim, counts, limiter, bcs, rough, extra in specs:
        gamma = 1.4 if rng.random() < 0.5 else 1.67
        prim = _rough(dim, counts, rng) if rough else _two_state(dim, counts, rng, gamma)
        lengths = [1.0 + 0.1 * k for k in range(dim)]
        out.append(Simulation(prim, lengths, t_end=t_end, cfl=0.4, gamma=gamma, boundaries=bcs, limiter=limiter,
                              precision=precision, **extra))
    return out


def _alone(sim, steps):
    ref = copy.deepcopy(sim)
    for _ in range(steps):
        if ref.completed:
            break
        ref.step()
    return ref


def _check_same(batch, ident, ref, exact=True):
    info = batch.info(ident)
    assert info["status"] == 0
    assert info["steps"] == ref.steps
    assert info["time"] == ref.time
    assert info["fallbacks"] == ref.fallbacks
    state = batch.state(ident)
    if exact:
        assert np.array_equal(state, ref.state)
        assert np.array_equal(info["boundary_integral"], ref.boundary_integral)
    else:
        assert_fields(state, ref.state, 1e-12)
        np.testing.assert_allclose(info["boundary_integral"], ref.boundary_integral, rtol=1e-12, atol=1e-14)


def test_batch_fp64_matches_sessions_bitwise():
    from torch_amd.torchBatch import Batch
    sims = _sessions("fp64")
    batch = Batch(precision="fp64")
    ids = [batch.add(s) for s in sims]
    steps = 9
    batch.run(steps)
    for s, i in zip(sims, ids):
        _check_same(batch, i, _alone(s, steps))
    # the rough instances took the first-order fallback inside the batch
    assert sum(batch.info(i)["fallbacks"] for i in ids) > 0
    batch.close()


def test_deep_grid_batch_of_one_matches_session_bitwise():
    """A grid 128 planes deep takes the longer default z chunk in a session
    and in a batch alike, so the balance sums fold the same blocks."""
    from torch_amd import Simulation
    from torch_amd.torchBatch import Batch

    rng = np.random.default_rng(11)
    for precision in ("fp64", "fp32"):
        sim = Simulation(_two_state(3, (24, 20, 130), rng, 1.4), [1.0, 1.0, 5.0], t_end=10.0, cfl=0.5,
                         limiter="Superbee", boundaries=["Open", "Reflective", "Open"], precision=precision)
        batch = Batch(precision=precision)
        ident = batch.add(sim)
        batch.run(6)
        _check_same(batch, ident, _alone(sim, 6))
        batch.close()
        sim.close()


def test_batch_fp32_matches_sessions():
    from torch_amd.torchBatch import Batch
    sims = _sessions("fp32", seed=6)
    batch = Batch(precision="fp32")
    ids = [batch.add(s) for s in sims]
    steps = 7
    batch.run(steps)
    for s, i in zip(sims, ids):
        ref = _alone(s, steps)
        info = batch.info(i)
        assert info["steps"] == ref.steps and info["time"] == ref.time
        # Same kernels and same per-cell arithmetic: bitwise in FP32 too.
        assert np.array_equal(batch.state(i), ref.state)
        assert np.array_equal(info["boundary_integral"], ref.boundary_integral)
    batch.close()


def test_batch_instance_matches_oracle():
    """One rough 3D instance with the fallback, batched, against the oracle."""
    from torch_amd.torchBatch import Batch
    from torch_amd.torchSession import Simulation
    rng = np.random.default_rng(3)
    prim = _rough(3, [9, 7, 10], rng)
    sim = Simulation(prim, [1.0, 1.0, 1.0], t_end=10.0, cfl=0.45, limiter="Superbee", precision="fp64")
    ora = OracleSession(prim, [1.0, 1.0, 1.0], t_end=10.0, cfl=0.45, limiter="Superbee")
    batch = Batch(precision="fp64")
    i = batch.add(sim)
    batch.run(5)
    for _ in range(5):
        ora.step()
    info = batch.info(i)
    assert info["fallbacks"] == ora.fallbacks > 0
    assert info["time"] == pytest.approx(ora.time, rel=1e-12)
    assert_fields(batch.state(i), ora.state, 1e-12)
    np.testing.assert_allclose(info["boundary_integral"], ora.boundary_integral, rtol=1e-11, atol=1e-13)
    batch.close()


def test_batch_end_time_and_step_cap():
    from torch_amd.torchBatch import Batch
    sims = _sessions("fp64", seed=8, specs=SPECS[:1] + SPECS[4:5] + SPECS[7:8], t_end=0.004)
    batch = Batch(precision="fp64")
    ids = [batch.add(s) for s in sims]
    capped = batch.add(sims[2], t_end=10.0, max_steps=2)
    batch.run(40)
    for s, i in zip(sims, ids):
        ref = _alone(s, 40)
        assert ref.completed
        info = batch.info(i)
        assert info["completed"] and not info["active"]
        _check_same(batch, i, ref)
    info = batch.info(capped)
    assert info["steps"] == 2 and not info["active"] and not info["completed"]
    batch.close()


def test_continuous_batching_same_alone_and_in_mix():
    """Instances admitted and removed between runs keep their trajectories."""
    from torch_amd.torchBatch import Batch
    sims = _sessions("fp64", seed=9)
    batch = Batch(precision="fp64")
    first = {batch.add(s): (s, 0) for s in sims[:4]}
    batch.run(3)
    later = {batch.add(s): (s, 3) for s in sims[4:8]}
    batch.run(4)
    gone = next(iter(first))
    batch.remove(gone)
    last = {batch.add(s): (s, 7) for s in sims[8:]}
    batch.run(5)
    members = {**first, **later, **last}
    members.pop(gone)
    for ident, (sim, admitted) in members.items():
        _check_same(batch, ident, _alone(sim, 12 - admitted))
    batch.close()


def test_batch_budget_admission():
    from torch_amd.torchBatch import Batch
    sims = _sessions("fp32", seed=10, specs=SPECS[7:10])
    probe = Batch(precision="fp32")
    need = [probe.footprint(s) for s in sims]
    probe.close()
    batch = Batch(precision="fp32", hbm_budget=need[0] + need[1] + need[2] // 2)
    assert batch.add(sims[0]) is not None
    b = batch.add(sims[1])
    assert b is not None
    assert batch.add(sims[2]) is None  # does not fit until something leaves
    batch.remove(b)
    assert batch.add(sims[2]) is not None
    batch.close()


def test_batch_autotune_keeps_results():
    from torch_amd.torchBatch import Batch
    sims = _sessions("fp64", seed=12, specs=SPECS[7:])
    batch = Batch(precision="fp64")
    ids = [batch.add(s) for s in sims]
    batch.run(2)
    (bx, by, zc), ms = batch.tune(reps=2)
    assert (bx, by) in ((64, 4), (64, 8), (32, 8), (32, 16)) and zc in (8, 16, 24, 32, 64) and ms > 0
    batch.run(4)
    for s, i in zip(sims, ids):
        # the tile changes the order of the boundary sums, not the cell arithmetic
        _check_same(batch, i, _alone(s, 6), exact=False)
        assert np.array_equal(batch.state(i), _alone(s, 6).state)
    batch.close()


def test_fp32_batch_autotune_and_wide_tile():
    """The FP32 form's sweep includes its 768-thread 32x24 tile (the FP64
    form has none and refuses it); whatever the tuner picks, every instance
    steps exactly as it does alone on the default tile."""
    from torch_amd.torchBatch import Batch
    with pytest.raises(ValueError):
        Batch(precision="fp64").set_tiles(32, 24, 24)
    from torch_amd import Simulation
    p = np.ones((5, 6, 6, 6))
    p[1:4] = 0.0
    with pytest.raises(ValueError, match="tile"):
        Simulation(p, [1.0, 1.0, 1.0], t_end=1.0, precision="fp64", tile=(32, 24, 24)).step()
    sims = _sessions("fp32", seed=14, specs=SPECS[7:])
    batch = Batch(precision="fp32")
    ids = [batch.add(s) for s in sims]
    batch.run(2)
    (bx, by, zc), ms = batch.tune(reps=2)
    assert (bx, by) in ((64, 4), (64, 8), (32, 8), (32, 16), (32, 24)) and zc in (8, 16, 24, 32, 64) and ms > 0
    batch.run(2)
    batch.set_tiles(32, 24, 24)
    batch.run(2)
    for s, i in zip(sims, ids):
        assert np.array_equal(batch.state(i), _alone(s, 6).state)
    batch.close()


def test_step_kernel_timing_takes_ordinary_steps():
    """The step graph captured with an event pair around the one 3D work
    list advances the instance exactly as the ordinary graph does (bitwise
    against the session alone), later runs replay the ordinary graph, and a
    batch with several 3D work lists is refused."""
    from torch_amd.torchBatch import Batch
    sims = _sessions("fp32", seed=15, specs=SPECS[7:8])
    batch = Batch(precision="fp32")
    i = batch.add(sims[0])
    batch.run(2)
    ms = batch.time_step_kernel(3)
    assert ms > 0
    batch.run(1)
    ref = _alone(sims[0], 6)
    assert batch.info(i)["steps"] == ref.steps == 6
    assert np.array_equal(batch.state(i), ref.state)
    batch.close()
    batch = Batch(precision="fp32")
    for s in _sessions("fp32", seed=16, specs=SPECS[7:9]):  # Superbee and VanLeer: two 3D lists
        batch.add(s)
    with pytest.raises(ValueError, match="one 3D work list"):
        batch.time_step_kernel(2)
    batch.close()


def test_batch_commit_into_session():
    from torch_amd.torchBatch import Batch
    sims = _sessions("fp64", seed=13, specs=SPECS[4:6])
    refs = [_alone(s, 5) for s in sims]
    batch = Batch(precision="fp64")
    ids = [batch.add(s) for s in sims]
    batch.run(5)
    for s, i, ref in zip(sims, ids, refs):
        batch.commit(i, s)
        assert s.steps == ref.steps and s.time == ref.time
        assert np.array_equal(s.state, ref.state)
        s.step()
        ref.step()
        assert np.array_equal(s.state, ref.state)
    batch.close()


IGNITION = {"q_heat": 3e9, "pre_exponential": 1e6}


def _ignition(precision, model, shape=(8, 10, 12), constants=IGNITION, bcs=("Open", "Periodic", "Reflective"),
              hot=1e8):
    """Hot, fast-reacting mixture with a large heat release: the heated
    half-step fails the CFL test every step (one dt halving each)."""
    from torch_amd.torchSession import Simulation
    nz, ny, nx = shape
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx,
                          indexing="ij")
    prim = np.zeros((5,) + shape)
    prim[0] = 1.0
    prim[4] = hot * (1 + 0.5 * np.exp(-((x - 0.3) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2) / 0.05))
    lam0 = 1.0 - 0.2 * x
    return Simulation(prim, [1.0, 1.0, 1.0], t_end=1e-5, reactive=True, kinetic_model=model, lambda_init=lam0,
                      boundaries=list(bcs), precision=precision, constants=constants)


def _check_reactive(batch, ident, ref):
    info = batch.info(ident)
    assert info["status"] == 0
    assert info["steps"] == ref.steps and info["time"] == ref.time
    assert np.array_equal(batch.state(ident), ref.state)
    lam, beta = batch.fractions(ident)
    assert np.array_equal(lam, ref.lam)
    if ref.beta is not None:
        assert np.array_equal(beta, ref.beta)
    assert np.array_equal(info["boundary_integral"], ref.boundary_integral)
    assert np.array_equal(info["source_integral"], ref.source_integral)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_reactive_batch_matches_sessions(precision):
    """Arrhenius and ZND ignition (with dt halvings) mixed with inert work."""
    from torch_amd.torchBatch import Batch
    sims = [_ignition(precision, "arrhenius"), _ignition(precision, "znd", shape=(6, 9, 7)),
            _ignition(precision, "arrhenius", shape=(5, 6, 40), bcs=("Reflective", "Open", "Open"))]
    inert = _sessions(precision, seed=14, specs=SPECS[5:6] + SPECS[9:10])
    steps = 5
    refs = [_alone(s, steps) for s in sims + inert]
    batch = Batch(precision=precision)
    ids = [batch.add(s, max_steps=steps) for s in sims + inert]
    batch.run(steps + 40)  # rejected chemistry attempts take a replay each
    retries = 0
    for s, i, ref in zip(sims + inert, ids, refs):
        if s.config["reactive"]:
            _check_reactive(batch, i, ref)
            retries += batch.info(i)["retries"]
        else:
            _check_same(batch, i, ref)
    assert retries > 0  # the dt-halving path ran inside the batch
    batch.close()


def test_reactive_batch_commit_and_continue():
    from torch_amd.torchBatch import Batch
    sim = _ignition("fp64", "znd")
    ref = _alone(sim, 6)
    batch = Batch(precision="fp64")
    i = batch.add(sim, max_steps=3)
    batch.run(30)
    batch.commit(i, sim)
    for _ in range(3):
        sim.step()
    assert sim.steps == ref.steps and sim.time == ref.time
    assert np.array_equal(sim.state, ref.state) and np.array_equal(sim.lam, ref.lam)
    assert np.array_equal(sim.source_integral, ref.source_integral)
    batch.close()


def test_membership_changes_update_the_graph_in_place():
    """Retiring an instance and admitting another with the same work lists
    (3D, Superbee) rewrites the captured step graph in place instead of
    instantiating it again (torch_batch_graph_counts: one instantiation, one
    update), and every instance still equals its session stepped alone,
    bitwise."""
    from torch_amd.torchBatch import Batch
    a, b = _sessions("fp32", seed=31, specs=[SPECS[7], SPECS[10]])
    (c,) = _sessions("fp32", seed=32, specs=[SPECS[7]])
    ref_a, ref_b, ref_c = _alone(a, 6), _alone(b, 3), _alone(c, 3)
    batch = Batch(precision="fp32")
    ia, ib = batch.add(a, max_steps=6), batch.add(b, max_steps=3)
    batch.run(3)
    _check_same(batch, ib, ref_b)
    batch.remove(ib)
    ic = batch.add(c, max_steps=3)
    batch.run(3)
    _check_same(batch, ia, ref_a)
    _check_same(batch, ic, ref_c)
    assert batch._batch.graph_counts() == (1, 1)
    batch.close()
