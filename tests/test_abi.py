"""CPU checks of the C-ABI boundary: the library builds, loads and exports
every entry point declared in include/torch_hip.h (no device calls)."""

import ctypes
import os
import re

import pytest

from conftest import ROOT
from torch_amd import _abi


def declared():
    names = set()
    for fname in os.listdir(os.path.join(ROOT, "include")):
        if fname.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fname)).read()
            names |= set(re.findall(r"^\s*int\s+(torch_\w+)\s*\(", text, re.M))
    return names


def test_header_and_binding_agree():
    assert declared() == set(_abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_abi.LIBRARY):
        pytest.skip("libtorchhip.so not built (run python -m torch_amd.build)")
    try:
        lib = ctypes.CDLL(_abi.LIBRARY)
    except OSError as exc:  # the HIP runtime itself is absent
        pytest.skip(f"HIP runtime not loadable here: {exc}")
    for name in declared():
        assert hasattr(lib, name), name


def test_configuration_validation_is_host_side():
    from torch_amd.torchSolver3D import config

    with pytest.raises(ValueError):
        config((4, 4, 4), dim=3, limiter="bad")
    with pytest.raises(ValueError):
        _abi.boundary_code("Inflow", 3, 1)
    with pytest.raises(ValueError):
        _abi.precision_code("fp16")
    c = config((8, 6, 4), (0.1, 0.2, 0.3), dim=3, bcs=("Open", "Reflective", "Periodic"), precision="fp32")
    assert (c.n[0], c.n[1], c.n[2]) == (8, 6, 4)
    assert (c.bc[0], c.bc[1], c.bc[2]) == (0, 2, 1)
    assert c.precision == 1


def test_execution_contract():
    from torch_amd import Execution, DeviceExecution, validate_workers
    import copy
    import threading

    for bad in (0, 65, 1.0, True, "2"):
        with pytest.raises(ValueError):
            validate_workers(bad)
    ex = Execution(4)
    assert list(ex.map(lambda v: v * 2, range(20))) == [v * 2 for v in range(20)]
    barrier = threading.Barrier(2, timeout=5)
    assert list(Execution(2).map(lambda v: barrier.wait() * 0 + v, [1, 2])) == [1, 2]
    dup = copy.deepcopy(DeviceExecution(3, precision="fp32"))
    assert dup.workers == 3 and dup.precision == "fp32"
    ex.close()
