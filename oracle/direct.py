"""Exact FP64 direct sums -- TEST INFRASTRUCTURE ONLY.

u(t) = sum_s K(s, t) tau_s over all sources, optionally skipping sources whose
coordinates equal the target's exactly (the self-interaction rule of
``flexmm/kifmm2d/scalar/helpers.py:214``).  Chunked over targets so that any
N fits in memory.

Pinned by: tiny cases summed by explicit Python loops
(tests/test_oracle_kernels.py).
"""
import numpy as np


def direct(Kv, sx, sy, tau, tx, ty, check=False, chunk_pairs=20_000_000):
    """Return an (nout, nt) array of sum_s Kv(s, t)[k] * tau_s."""
    sx = np.asarray(sx, dtype=np.float64).ravel()
    sy = np.asarray(sy, dtype=np.float64).ravel()
    tau = np.asarray(tau, dtype=np.float64).ravel()
    tx = np.asarray(tx, dtype=np.float64).ravel()
    ty = np.asarray(ty, dtype=np.float64).ravel()
    out = None
    step = max(1, chunk_pairs // max(1, sx.size))
    for s in range(0, tx.size, step):
        ttx = tx[s:s + step, None]
        tty = ty[s:s + step, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            vals = Kv(sx[None, :], sy[None, :], ttx, tty)
        if not isinstance(vals, tuple):
            vals = (vals,)
        if out is None:
            out = np.zeros((len(vals), tx.size))
        if check:
            keep = ~((sx[None, :] == ttx) & (sy[None, :] == tty))
        for k, v in enumerate(vals):
            w = v * tau[None, :]
            if check:
                w = np.where(keep, w, 0.0)
            out[k, s:s + step] = w.sum(axis=1)
    if out is None:
        out = np.zeros((1, tx.size))
    return out
