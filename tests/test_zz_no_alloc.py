"""The allocation-free HLLC entry point (torch_hllc_faces_static) against
torch_hllc_faces through the C ABI: the same kernels, so the same results
bit for bit, on two calls in a row (the second smaller, over the staging the
first left behind), in both precisions and for 4 and 5 components; sizes
beyond its fixed staging are rejected."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def faces(rng, nv, m):
    rho = rng.uniform(0.2, 2.0, (2, m))
    vel = rng.uniform(-1.0, 1.0, (2, nv - 2, m))
    p = rng.uniform(0.1, 3.0, (2, m))
    g = rng.choice([1.4, 1.67], (2, m))
    out = []
    for s in range(2):
        u = np.empty((nv, m))
        u[0] = rho[s]
        u[1:-1] = rho[s] * vel[s]
        u[-1] = p[s] / (g[s] - 1) + 0.5 * rho[s] * np.sum(vel[s] ** 2, axis=0)
        out.append(u)
    return out[0], out[1], g[0], g[1]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nv", [4, 5])
def test_static_faces_match_the_allocating_entry_point(precision, nv):
    from torch_amd import _abi

    ctx = _abi.Context(0)
    try:
        rng = np.random.default_rng(97 + nv + 10 * precision)
        for m in (65536, 3001):
            left, right, gl, gr = faces(rng, nv, m)
            for direction in range(nv - 2):
                want = _abi.hllc_faces(ctx, nv, direction, left, right, gl, gr, precision)
                got = _abi.hllc_faces_static(ctx, nv, direction, left, right, gl, gr, precision)
                assert np.array_equal(got, want), (m, direction)
                assert np.all(np.isfinite(got))
        left, right, gl, gr = faces(rng, nv, 65537)
        with pytest.raises(ValueError, match="at most 65536 faces"):
            _abi.hllc_faces_static(ctx, nv, 0, left, right, gl, gr, precision)
    finally:
        ctx.close()
