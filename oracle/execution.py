"""Bounded CPU worker pool for independent flux batches (oracle restatement).

TEST INFRASTRUCTURE ONLY: the CPU baseline's worker path. Restates
python/cfdSolver/torchExecution.py:7-68: workers is an int in 1..64 (bool
and float rejected); map() is plain map for one worker, otherwise a lazily
created thread pool with at most `workers` tasks live, results yielded in
submission order, and on an exception the pending tasks are cancelled and
the running ones awaited. NumPy releases the interpreter lock inside its
ufuncs, so batches run concurrently on the pool's threads.
"""

from collections import deque
from concurrent.futures import ThreadPoolExecutor


def validate_workers(workers):
    if not isinstance(workers, int) or isinstance(workers, bool) or not 1 <= workers <= 64:
        raise ValueError("Workers must be an integer from 1 to 64")
    return workers


class Execution:
    def __init__(self, workers=1):
        self.workers = validate_workers(workers)
        self.pool = None

    def map(self, function, items):
        if self.workers == 1:
            yield from map(function, items)
            return
        if self.pool is None:
            self.pool = ThreadPoolExecutor(max_workers=self.workers)
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
            for task in live:
                task.cancel()
            for task in live:
                if not task.cancelled():
                    try:
                        task.result()
                    except Exception:
                        pass

    def close(self):
        if self.pool is not None:
            self.pool.shutdown(wait=True, cancel_futures=True)
            self.pool = None

    def __del__(self):
        self.close()
