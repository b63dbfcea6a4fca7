"""Thread-count knob of ``freud.parallel`` (freud/parallel.py).

The GPU path does not use host threads; the value is kept for API parity and is
reported by the benchmark's CPU baseline.
"""

import os

_num_threads = os.cpu_count() or 1


def get_num_threads():
    return _num_threads


def set_num_threads(nthreads=None):
    global _num_threads
    if nthreads is None or nthreads == 0:
        nthreads = os.cpu_count() or 1
    if nthreads < 0:
        msg = "nthreads must be positive"
        raise ValueError(msg)
    _num_threads = int(nthreads)


class NumThreads:
    """Context manager setting the thread count temporarily."""

    def __init__(self, N=None):
        self.restore_N = _num_threads
        self.N = N

    def __enter__(self):
        set_num_threads(self.N)
        return self

    def __exit__(self, *args):
        set_num_threads(self.restore_N)
