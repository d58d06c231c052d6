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


def direct_stokes(sx, sy, fx, fy, tx, ty, check=False, chunk_pairs=20_000_000):
    """(u, v, p) of the 2D Stokeslet summed over all sources, as the reference's
    kernel_apply_single(_check) (flexmm/kifmm2d/vector/stokes_fmm.py:102-184):
    u = (1/4 pi) sum [(-log r + dx^2/r^2) f_x + dx dy/r^2 f_y], v likewise,
    p = (1/2 pi) sum (dx f_x + dy f_y)/r^2; check skips coincident points.

    Pinned by: the Stokes equations (div u = 0, Laplacian u = grad p away from
    the sources) by finite differences, and a tiny explicit-loop case
    (tests/test_oracle_kernels.py)."""
    sx, sy, fx, fy, tx, ty = [np.asarray(a, dtype=np.float64).ravel()
                              for a in (sx, sy, fx, fy, tx, ty)]
    out = np.zeros((3, tx.size))
    step = max(1, chunk_pairs // max(1, sx.size))
    for s in range(0, tx.size, step):
        dx = tx[s:s + step, None] - sx[None, :]
        dy = ty[s:s + step, None] - sy[None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            d2 = dx * dx + dy * dy
            id2 = 1.0 / d2
            logid = 0.5 * np.log(id2)
            uu = (logid + dx * dx * id2) * fx + dx * dy * id2 * fy
            vv = dx * dy * id2 * fx + (logid + dy * dy * id2) * fy
            pp = (dx * fx + dy * fy) * id2
        if check:
            keep = ~((dx == 0) & (dy == 0))
            uu, vv, pp = [np.where(keep, a, 0.0) for a in (uu, vv, pp)]
        out[0, s:s + step] = uu.sum(axis=1) * (0.25 / np.pi)
        out[1, s:s + step] = vv.sum(axis=1) * (0.25 / np.pi)
        out[2, s:s + step] = pp.sum(axis=1) * (0.5 / np.pi)
    return out


def direct_stokes_dipole(sx, sy, fx, fy, sigx, sigy, nx, ny, tx, ty, check=False,
                         chunk_pairs=10_000_000):
    """(u, v, p) of Stokeslets (forces f) plus stresslets (double-layer density
    sigma, normal n) summed over all sources, as the reference's
    kernel_apply_single(_check)_with_dipole
    (flexmm/kifmm2d/vector/stokes_fmm.py:129-235).  With r = t - s:
    stresslet velocity (1/pi) (r.n)(r.sigma) r / |r|^4,
    stresslet pressure (1/pi) [-(n.sigma)/|r|^2 + 2 (r.n)(r.sigma)/|r|^4];
    check skips coincident points.

    Pinned by: the stresslet velocity equals minus the stress of a Stokeslet
    of force sigma contracted with n (stress from finite differences of
    direct_stokes), the Stokes equations of the double layer (div u = 0,
    Laplacian u = grad p) by finite differences, and a tiny explicit-loop case
    (tests/test_oracle_kernels.py)."""
    sx, sy, sigx, sigy, nx, ny, tx, ty = [np.asarray(a, dtype=np.float64).ravel()
                                          for a in (sx, sy, sigx, sigy, nx, ny, tx, ty)]
    out = direct_stokes(sx, sy, fx, fy, tx, ty, check=check, chunk_pairs=chunk_pairs)
    step = max(1, chunk_pairs // max(1, sx.size))
    for s in range(0, tx.size, step):
        dx = tx[s:s + step, None] - sx[None, :]
        dy = ty[s:s + step, None] - sy[None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            id2 = 1.0 / (dx * dx + dy * dy)
            rn = dx * nx + dy * ny
            rs = dx * sigx + dy * sigy
            w = rn * rs * id2 * id2
            uu = w * dx
            vv = w * dy
            pp = -(nx * sigx + ny * sigy) * id2 + 2.0 * w
        if check:
            keep = ~((dx == 0) & (dy == 0))
            uu, vv, pp = [np.where(keep, a, 0.0) for a in (uu, vv, pp)]
        out[0, s:s + step] += uu.sum(axis=1) / np.pi
        out[1, s:s + step] += vv.sum(axis=1) / np.pi
        out[2, s:s + step] += pp.sum(axis=1) / np.pi
    return out


def direct_src(src, sx, sy, tau_rows, extras_rows, tx, ty, check=False, chunk_pairs=20_000_000):
    """(nout, nt) array of sum_s src(s, t, tau_s, extras_s) for a source kernel
    ``src(sx, sy, tx, ty, tau_rows, extras_rows) -> tuple`` already weighted by
    the densities (the upwardor/evaluator kernels of register_upwardor,
    helpers.py:155-174); check skips exactly coincident points.

    Pinned by: tiny explicit-loop cases (tests/test_oracle_kernels.py)."""
    sx, sy, tx, ty = [np.asarray(a, dtype=np.float64).ravel() for a in (sx, sy, tx, ty)]
    tau_rows = np.asarray(tau_rows, dtype=np.float64).reshape(-1, sx.size)
    ex = None if extras_rows is None else np.asarray(extras_rows, dtype=np.float64).reshape(-1, sx.size)
    out = None
    step = max(1, chunk_pairs // max(1, sx.size))
    for s in range(0, tx.size, step):
        ttx = tx[s:s + step, None]
        tty = ty[s:s + step, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            vals = src(sx[None, :], sy[None, :], ttx, tty, tau_rows[:, None, :],
                       None if ex is None else ex[:, None, :])
        if out is None:
            out = np.zeros((len(vals), tx.size))
        if check:
            keep = ~((sx[None, :] == ttx) & (sy[None, :] == tty))
        for k, v in enumerate(vals):
            if check:
                v = np.where(keep, v, 0.0)
            out[k, s:s + step] = v.sum(axis=1)
    if out is None:
        out = np.zeros((1, tx.size))
    return out
