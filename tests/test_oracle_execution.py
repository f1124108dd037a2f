"""The oracle's worker path (the CPU baseline's): batched HLLC through the
bounded pool equals the single-batch grid bitwise for every batch size and
worker count (testPerformance.py:23-52, testAdaptive.py:460-480), and the
pool keeps the reference's contract (torchExecution.py:7-68)."""

import threading

import numpy as np
import pytest

from oracle import flux as oflux
from oracle import presets as opresets
from oracle import session as osession
from oracle.execution import Execution, validate_workers


def faces(rng, nv, shape):
    rho = rng.uniform(0.2, 2.0, (2,) + shape)
    vel = rng.uniform(-1.0, 1.0, (2, nv - 2) + shape)
    p = rng.uniform(0.1, 3.0, (2,) + shape)
    out = []
    for s in range(2):
        u = np.empty((nv,) + shape)
        u[0] = rho[s]
        u[1:-1] = rho[s] * vel[s]
        u[-1] = p[s] / 0.4 + 0.5 * rho[s] * np.sum(vel[s] ** 2, axis=0)
        out.append(u)
    return out


@pytest.mark.parametrize("nv", [4, 5])
def test_batched_grid_equals_single_batch_bitwise(nv):
    rng = np.random.default_rng(11)
    shape = (7, 9, 13) if nv == 5 else (9, 13)
    left, right = faces(rng, nv, shape)
    gl = rng.choice([1.4, 1.67], shape)
    for direction in range(nv - 2):
        whole = oflux.hllc_grid(left, right, gl, 1.4, direction=direction)
        for batch in (1, 17, 100, 8192):
            for workers in (1, 4):
                ex = Execution(workers)
                got = oflux.hllc_grid(left, right, gl, 1.4, direction=direction, batch_size=batch, execution=ex)
                ex.close()
                assert np.array_equal(got, whole, equal_nan=True), (direction, batch, workers)


def test_worker_session_equals_single_batch_session():
    prim, opt = opresets.preset("3d_explosion", [12, 10, 8], limiter="Albada", t_end=0.2)
    opt.pop("reactive")
    a = osession.Session(prim, [1.0] * 3, **opt)
    b = osession.Session(prim, [1.0] * 3, workers=3, **opt)
    for _ in range(3):
        a.step()
        b.step()
    assert np.array_equal(a.state, b.state) and a.time == b.time
    np.testing.assert_array_equal(a.boundary_integral, b.boundary_integral)
    b.execution.close()


def test_execution_contract():
    for bad in (0, 65, 1.0, True, "2"):
        with pytest.raises(ValueError):
            validate_workers(bad)
    ex = Execution(4)
    assert list(ex.map(lambda v: v * v, range(50))) == [v * v for v in range(50)]
    live, peak, lock = [0], [0], threading.Lock()

    def slow(v):
        with lock:
            live[0] += 1
            peak[0] = max(peak[0], live[0])
        threading.Event().wait(0.002)
        with lock:
            live[0] -= 1
        if v == 7:
            raise RuntimeError("batch 7 failed")
        return v

    with pytest.raises(RuntimeError, match="batch 7"):
        list(ex.map(slow, range(40)))
    assert peak[0] <= 4 and live[0] == 0
    ex.close()
    assert list(Execution(1).map(str, [None, 1])) == ["None", "1"]
