"""The oracle KIFMM on several host cores -- TEST INFRASTRUCTURE ONLY (the CPU
baseline beside the GPU benchmark).

The reference parallelises its per-leaf loops with numba ``prange``: the S2M
upwardor over leaves (kifmm2d/scalar/helpers.py:159-173) and the evaluation
over targets (L2T and P2P, helpers.py:180-235), with threaded BLAS for the
dense passes.  This module does the same to oracle.kifmm2d.FMM with worker
processes forked at each parallel step (NumPy holds the GIL between array
operations, so threads do not scale here): the leaves of an S2M level and the
targets of an evaluation are split into chunks, each computed by the
oracle's own serial code in a child that sees the parent's arrays, and the
results are combined in the parent.  The tree build stays serial, as the
reference's level-0 partition is (tree.py:191-220 runs one node on one
thread).  Results equal the serial oracle's up to the order of additions.
"""
import multiprocessing as mp

import numpy as np

from .kifmm2d import FMM

_CTX = mp.get_context("fork")
_JOB = {}


def _run(i):
    return _JOB["fn"](i)


def _par_map(fn, n, workers):
    _JOB["fn"] = fn
    try:
        with _CTX.Pool(workers) as pool:
            return pool.map(_run, range(n))
    finally:
        _JOB.clear()


class ParallelFMM(FMM):
    """oracle.kifmm2d.FMM with forked workers for S2M and evaluation."""

    def __init__(self, *args, workers=1, **kw):
        self.workers = int(workers)
        super().__init__(*args, **kw)

    def _s2m(self, lev, boxes, prec, partial, chunk=4_000_000):
        w = self.workers
        if w <= 1 or boxes.size < 4 * w:
            return FMM._s2m(self, lev, boxes, prec, partial, chunk)
        parts = np.array_split(boxes, 4 * w)
        rows = [lev.this_density_ind[b] for b in parts]

        def work(i):
            loc = np.zeros_like(partial)
            FMM._s2m(self, lev, parts[i], prec, loc, chunk)
            return loc[rows[i]]

        for r, v in zip(rows, _par_map(work, len(parts), w)):
            partial[r] += v

    def evaluate_to_points(self, x, y, name, check):
        x = np.asarray(x, dtype=np.float64).ravel()
        y = np.asarray(y, dtype=np.float64).ravel()
        w = self.workers
        if w <= 1 or x.size < 4 * w:
            return FMM.evaluate_to_points(self, x, y, name, check)
        parts = np.array_split(np.arange(x.size), 4 * w)

        def work(i):
            return FMM.evaluate_to_points(self, x[parts[i]], y[parts[i]], name, check)

        return np.concatenate(_par_map(work, len(parts), w), axis=1)
