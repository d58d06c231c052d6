"""Per-width FP64 operators of the KIFMM, computed once on the host.

Mirrors ``Precomputation`` of the reference (flexmm/kifmm2d/scalar/helpers.py:
250-317) and packs the result into the device layout of
``fmm2d_set_operators`` (include/flexmm_amd.h):

* equivalent/check circles: p equi-angular points on radii
  r1 = w sqrt(2)/2 and r2 = w (4 - sqrt(2))/2 (helpers.py:5-15);
* A = K(small circle -> large circle); its pseudo-inverse
  V diag(1/s) U^T with singular values below 1e-15 dropped
  (SVD_Solver, helpers.py:240-248) serves both S2L and L2S (helpers.py:263-267);
* M2MC = K(children's small circles -> parent's large circle) (p x 4p,
  children ordered (-,-), (-,+), (+,-), (+,+); helpers.py:270-284), L2LC = M2MC^T;
* the 40 M2L matrices K(small circle shifted by (ox, oy) w -> small circle)
  (helpers.py:288-296), stored dense, and their SVD-compressed form by the
  reference's rule (M2L_Evaluator, helpers.py:17-36): r = the number of
  singular values above 1e-14 (absolute); if 2 r < p the slot is applied as
  M2 (M1 q) with M1 = diag(s_:r) V_:r^T (r x p) and M2 = U_:r (p x r); if
  r = 0 it contributes nothing; otherwise it stays dense.  The device applies
  the compressed product (rank > 0) or the dense matrix (rank -1) per slot.

For the mixed-precision form each M2L matrix is split as M = c 1^T + R with
c_i the mean of row i: R goes to FP32 matrix cores, while c times the source
box's total equivalent charge stays FP64.  Removing the dominant constant mode
keeps the FP32 accumulation error an order of magnitude below the 1e-7 target.

These are setup computations of the plan (once per width, cached), not part
of the evaluation path.
"""
import numpy as np

PINV_TOL = 1e-15   # helpers.py:241
M2L_SVD_TOL = 1e-14   # helpers.py:251
OP_NBLK = 50


def padded_dim(p):
    if p < 4 or p > 64:
        raise ValueError("Nequiv = %d outside [4, 64]" % p)
    return 32 if p <= 32 else 64


def m2l_slots():
    """(ox, oy) of the 40 M2L slots: (ii, jj) over the 7x7 grid row-major,
    skipping the 3x3 centre; offset (ii - 3, jj - 3) (tree.py:431-435)."""
    out = []
    for ii in range(7):
        for jj in range(7):
            if 2 <= ii <= 4 and 2 <= jj <= 4:
                continue
            out.append((ii - 3, jj - 3))
    return out


def circles(width, p):
    theta = np.linspace(0, 2 * np.pi, p, endpoint=False)
    r1 = 0.5 * width * np.sqrt(2)
    r2 = 0.5 * width * (4 - np.sqrt(2))
    return (r1 * np.cos(theta), r1 * np.sin(theta), r2 * np.cos(theta), r2 * np.sin(theta))


def kernel_matrix(K, sx, sy, tx, ty):
    """M[j, i] = K(s_i -> t_j)."""
    return K(sx[None, :], sy[None, :], tx[:, None], ty[:, None])


class LevelOperators:
    """FP64 operators of one box width."""

    def __init__(self, kernel, p, width):
        self.p = p
        self.width = float(width)
        sx, sy, lx, ly = circles(width, p)
        A = kernel_matrix(kernel, sx, sy, lx, ly)
        U, S, VH = np.linalg.svd(A)
        SI = np.where(S < PINV_TOL, 0.0, 1.0 / np.where(S < PINV_TOL, 1.0, S))
        self.pinv1 = SI[:, None] * U.T
        self.pinv2 = VH.T
        self.pinv_rank = int((S >= PINV_TOL).sum())
        ssx, ssy, _, _ = circles(0.5 * width, p)
        q = 0.25 * width
        cex = np.concatenate([ssx - q, ssx - q, ssx + q, ssx + q])
        cey = np.concatenate([ssy - q, ssy + q, ssy - q, ssy + q])
        self.M2MC = kernel_matrix(kernel, cex, cey, lx, ly)
        self.m2l = [kernel_matrix(kernel, sx + ox * width, sy + oy * width, sx, sy)
                    for ox, oy in m2l_slots()]
        self.m2l_svd = [compress_m2l(m, p) for m in self.m2l]
        self.small = (sx, sy)
        self.large = (lx, ly)

    @property
    def m2l_ranks(self):
        """Per slot: the compressed rank r (2 r < p), 0 for a null slot, -1 dense."""
        return [c[0] for c in self.m2l_svd]

    def pack(self):
        """Flat FP64 array in the fmm2d_set_operators layout."""
        p = self.p
        PD = padded_dim(p)
        means = np.zeros((40, PD))
        resid = np.zeros((40, PD, PD), dtype=np.float32)
        for s, m in enumerate(self.m2l):
            c = m.mean(axis=1)
            means[s, :p] = c
            resid[s, :p, :p] = (m - c[:, None]).astype(np.float32)
        rows, cols = residual_fragment_order(PD)
        resid = resid[:, rows, cols]
        blocks = np.zeros((OP_NBLK, PD, PD))
        blocks[0, :p, :p] = self.pinv1
        blocks[1, :p, :p] = self.pinv2
        for c in range(4):
            m = self.M2MC[:, c * p:(c + 1) * p]
            blocks[2 + c, :p, :p] = m
            blocks[6 + c, :p, :p] = m.T
        for s, m in enumerate(self.m2l):
            blocks[10 + s, :p, :p] = m
        circ = np.zeros((2, PD))
        circ[0, :p] = self.large[0]
        circ[1, :p] = self.large[1]
        ranks = np.full(64, -1.0)
        svd = np.zeros((40, PD, PD))
        for s, (r, M1, M2) in enumerate(self.m2l_svd):
            ranks[s] = r
            if r > 0:
                svd[s] = pack_svd_block(M1, M2, PD)
        return np.concatenate([blocks.ravel(), circ.ravel(), means.ravel(),
                               resid.ravel().view(np.float64), ranks, svd.ravel()])


def residual_fragment_order(PD):
    """(rows, cols) such that an FP32 residual block R is stored as
    R[rows, cols]: the operand order of the FP32 MFMA M2L, in which one
    16-byte load per lane feeds four MFMA steps and a wave's load reads 1 KB
    in one piece.  Entry ((rb MP + mp) 64 + lane) 4 + e, MP = PD / 16, holds
    R[16 rb + lane % 16][8 m + 2 (lane // 16) + e % 2] with m = 2 mp + e // 2."""
    MP = PD // 16
    rb, mp, lane, e = np.meshgrid(np.arange(PD // 16), np.arange(MP), np.arange(64),
                                  np.arange(4), indexing="ij")
    rows = 16 * rb + lane % 16
    cols = 8 * (2 * mp + e // 2) + 2 * (lane // 16) + e % 2
    return rows.ravel(), cols.ravel()


def compress_m2l(m, p, tol=M2L_SVD_TOL):
    """helpers.py:17-36: (rank, M1, M2) with M2 @ M1 the compressed matrix when
    2 rank < p; (0, None, None) for a null matrix; (-1, None, None) dense."""
    U, S, VH = np.linalg.svd(m)
    r = int((S > tol).sum())
    if r == 0:
        return 0, None, None
    if 2 * r < p:
        return r, S[:r, None] * VH[:r], U[:, :r].copy()
    return -1, None, None


def pack_svd_block(M1, M2, PD):
    """One [PD][PD] block holding a compressed slot for the device: rows
    0 .. PD/2-1 are M1 (r x p, zero padded); rows PD/2 .. PD-1 hold M2 (p x r)
    two rows of M2 per block row: block[PD/2 + i // 2][(i % 2) PD/2 + j] =
    M2[i, j], so that a lane of the second MFMA stage reads its 4 consecutive
    columns of one M2 row from one place."""
    H = PD // 2
    r, p = M1.shape
    assert r <= H and M2.shape == (p, r)
    b = np.zeros((PD, PD))
    b[:r, :p] = M1
    m2 = np.zeros((PD, H))
    m2[:p, :r] = M2
    b[H:, :] = m2.reshape(H, 2 * H)
    return b
