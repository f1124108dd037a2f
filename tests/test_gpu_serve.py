"""Serving on the device: requests admitted as they arrive, by HBM budget,
into one continuously batched device run. Each result is the one the same
request gives alone (and, for one of them, the one a session stepped alone
gives), bitwise, since every instance keeps its own arithmetic."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs(count=6):
    from torch_amd.workloads import served_requests
    specs = served_requests(seed=355, count=count, t_run=0.02, znd_fraction=0.34, sizes=(14, 22))
    assert any(s["model"] == "znd" for s in specs)
    return specs


def _serve(specs, arrivals, precision, budget=0):
    from torch_amd.torchServe import Server, ignition_request
    server = Server(precision=precision, hbm_budget=budget, round_steps=3, keep_results=True)
    reqs = server.serve([ignition_request(s, a) for s, a in zip(specs, arrivals)])
    server.close()
    return {r.ident: r for r in reqs}


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_served_results_same_alone_and_in_mix(precision):
    specs = _specs()
    mixed = _serve(specs, [0.0, 0.0, 0.005, 0.01, 0.02, 0.03], precision)
    for spec in specs:
        alone = _serve([spec], [0.0], precision)[spec["index"]]
        got = mixed[spec["index"]]
        assert got.status == 0 and alone.status == 0
        assert got.steps == alone.steps > 0
        for a, b in zip(got.result, alone.result):
            if b is not None:
                assert np.array_equal(a, b)


def test_served_budget_queues_requests():
    """A budget for about two instances at a time: the rest wait, then run."""
    from torch_amd import _abi
    from torch_amd.torchSolver3D import config
    specs = _specs(5)
    need = []
    for s in specs:
        cfg = config([s["n"]] * 3, None, [1.0] * 3, dim=3, reactive=True, model=s["model"], gamma_mode=1,
                     precision="fp32")
        need.append(_abi.DeviceBatch.footprint(cfg, _abi.precision_code("fp32")))
    budget = sorted(need)[-1] + sorted(need)[-2]
    got = _serve(specs, [0.0] * len(specs), "fp32", budget=budget)
    ref = _serve(specs, [0.0] * len(specs), "fp32")
    admitted = sorted(r.admitted for r in got.values())
    assert admitted[-1] > admitted[0]  # some requests waited for room
    for i, r in got.items():
        assert r.status == 0
        assert np.array_equal(r.result[0], ref[i].result[0])


def test_served_request_matches_session():
    """The served path against the single-session step loop."""
    from torch_amd.torchServe import ignition_request
    from torch_amd.torchSession import Simulation
    spec = _specs(2)[1]
    req = ignition_request(spec, 0.0)
    prim, opts = req.build()
    sim = Simulation(prim, [1.0] * 3, t_end=opts["t_end"], reactive=True, kinetic_model=opts["kinetic_model"],
                     lambda_init=opts["lambda_init"], constants=opts["constants"], cfl=opts["cfl"],
                     limiter=opts["limiter"], precision="fp64")
    while not sim.completed:
        sim.step()
    served = _serve([spec], [0.0], "fp64")[spec["index"]]
    assert served.steps == sim.steps
    assert np.array_equal(served.result[0], sim.state)
    assert np.array_equal(served.result[1], sim.lam)
    sim.close()
