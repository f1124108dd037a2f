"""Execution backends: the reference's bounded CPU worker pool and its
MI355X replacement.

Mirrors python/cfdSolver/torchExecution.py:7-68. ``Execution`` keeps the
reference contract for host callables (ordered results, at most ``workers``
live tasks, pending work cancelled and awaited on failure).
``DeviceExecution`` is the drop-in backend for the flux workers: it owns one
HIP device context of libtorchhip, and the step functions that receive it
(``hllc_flux_grid``, ``muscl_hancock_step_2d/3d``, ``Simulation``) evaluate
their faces on the GPU instead of in 8192-face host batches. Its ``map``
still serves host callables exactly like ``Execution.map``.
"""

from collections import deque
from concurrent.futures import ThreadPoolExecutor

from . import _abi


def validate_workers(workers):
    """torchExecution.py:7-14: an int (not bool) in 1..64."""
    if not isinstance(workers, int) or isinstance(workers, bool) or not 1 <= workers <= 64:
        raise ValueError("Workers must be an integer from 1 to 64")
    return workers


class Execution:
    """Lazy thread pool with ordered results and bounded live tasks."""

    def __init__(self, workers=1):
        self.pool = None
        self.workers = validate_workers(workers)

    def map(self, function, items):
        if self.workers == 1:
            yield from map(function, items)
            return
        if self.pool is None:
            self.pool = ThreadPoolExecutor(max_workers=self.workers, thread_name_prefix="torch-flux")
        source = iter(items)
        live = deque()
        done = object()
        try:
            while len(live) < self.workers:
                item = next(source, done)
                if item is done:
                    break
                live.append(self.pool.submit(function, item))
            while live:
                result = live.popleft().result()
                item = next(source, done)
                if item is not done:
                    live.append(self.pool.submit(function, item))
                yield result
        finally:
            # No task may keep using a failed step's arrays.
            for future in live:
                future.cancel()
            for future in live:
                if not future.cancelled():
                    try:
                        future.result()
                    except Exception:
                        pass

    def close(self):
        if self.pool is not None:
            self.pool.shutdown(wait=True, cancel_futures=True)
            self.pool = None

    def __deepcopy__(self, memo):
        return type(self)(self.workers)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceExecution(Execution):
    """GPU flux-worker backend: one libtorchhip context on ``device``.

    ``precision`` ("fp64" or "fp32") selects the kernel build used by the
    step functions that receive this backend.
    """

    def __init__(self, workers=1, *, device=0, precision="fp64"):
        super().__init__(workers)
        self.device = int(device)
        self.precision = str(precision).lower()
        _abi.precision_code(self.precision)
        self._context = None
        # 3D kernel tile (block_x, block_y, zchunk); None uses the tuned default.
        self.tile = None

    @property
    def context(self):
        if self._context is None:
            self._context = _abi.Context(self.device)
        return self._context

    def close(self):
        super().close()
        if self._context is not None:
            self._context.close()
            self._context = None

    def __deepcopy__(self, memo):
        return DeviceExecution(self.workers, device=self.device, precision=self.precision)


_default = None


def default_execution():
    """Process-wide device backend for calls that pass ``execution=None``."""
    global _default
    if _default is None:
        _default = DeviceExecution()
    return _default


def device_of(execution):
    """The DeviceExecution to run on: the given one, or the default."""
    return execution if isinstance(execution, DeviceExecution) else default_execution()
