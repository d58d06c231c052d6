"""CPU restatement of flexmm's black-box (Chebyshev) FMM -- TEST INFRASTRUCTURE ONLY.

Follows ``flexmm/flexmm2d/bbfmm.py`` (BB_FMM, BB_Helper, BB_Precomputation)
and the generic passes of ``flexmm/flexmm2d/fmm.py`` in NumPy FP64, on the
same quadtree as the KIFMM (``oracle.tree``, the reference's ``tree.py``).

* Chebyshev nodes: the p Gauss-Chebyshev points in ascending order
  (bbfmm.py:131-133); node (i, j) of the p x p grid has index i p + j.
* Multipole = Chebyshev moments m[i p + j] = sum_s T_i(x_s) T_j(y_s) tau_s of
  the sources scaled to [-1, 1]^2 (x = (s - centre) 2/w; anterpolate2,
  chebV2 :11-24, source2pMs :173-184); the pseudo-inverse steps of the KIFMM
  are the identity (fmm.py:58-61).
* M2M: M2MC = chebV2(children's nodes in parent coordinates)^T
  blockdiag(VI^T) (bbfmm.py:138-155), children in the order
  (-,-), (-,+), (+,-), (+,+); L2LC = M2MC^T (:157).
* M2L of offset (ix, iy): VI K(nodes w/2 + offset w -> nodes w/2) VI^T
  (bbfmm.py:97-117), K[j, i] = kernel(source i -> target j).
* L2T: sum c[i p + j] T_i(x) T_j(y) (chebeval2d1 :27-32, L2targets :185-199).
* P2P: the kernel over the leaf's colleagues, as the KIFMM.

Differences from the reference: precomputations are cached per (kernel, p,
round(width, 12)); the reference's default helper object (bbfmm.py:202) is
shared by every BB_FMM, whatever its kernel or order.

Pinned by (tests/test_oracle_bbfmm.py): the Chebyshev basics against
numpy.polynomial.chebyshev; M2M and L2L exactness on polynomials (moments of
the same sources about the parent; the parent expansion evaluated in a
child); the M2L of one source-target pair equal to the closed-form bivariate
Chebyshev interpolation of the kernel in both points (Fong & Darve 2009);
the whole FMM against FP64 direct sums with the error falling geometrically
in p; planted mistakes (a transposed M2L, swapped children) caught.
"""
import numpy as np

from .kifmm2d import FMM as _KIFMM, M2LEvaluator, kernel_form, get_bbox, _segment_add
from .tree import Tree


def cheb_nodes(p):
    """bbfmm.py:131-132: np.polynomial.chebyshev.chebgauss(p) nodes, ascending."""
    k = np.arange(1, p + 1)
    return np.cos((2 * k - 1) * np.pi / (2 * p))[::-1].copy()


def chebt(p, x):
    """(n, p) array of T_0..T_{p-1} at x (all_chebt :44-64), by the
    three-term recurrence."""
    x = np.asarray(x, dtype=np.float64).ravel()
    out = np.empty((x.size, p))
    out[:, 0] = 1.0
    if p > 1:
        out[:, 1] = x
    for k in range(2, p):
        out[:, k] = 2 * x * out[:, k - 1] - out[:, k - 2]
    return out


def chebv2(x, y, p):
    """chebV2 (bbfmm.py:14-24): (n, p^2), column i p + j = T_i(x) T_j(y)."""
    xv, yv = chebt(p, x), chebt(p, y)
    return (xv[:, :, None] * yv[:, None, :]).reshape(xv.shape[0], p * p)


def cheb_eval2d(x, y, c, p):
    """sum_{i,j} c[..., i p + j] T_i(x) T_j(y) for arrays of points and
    coefficient rows (chebeval2d1, bbfmm.py:27-32)."""
    return (chebv2(x, y, p) * c).sum(axis=1)


class BBInfo:
    """The width-independent part of BB_Helper.prepare (bbfmm.py:127-166)."""

    def __init__(self, p):
        self.p = p
        nv = cheb_nodes(p)
        nx, ny = np.meshgrid(nv, nv, indexing="ij")
        self.nodex, self.nodey = nx.ravel(), ny.ravel()
        V = chebv2(self.nodex, self.nodey, p)          # == chebvander2d(..., [p-1, p-1])
        self.VI = np.linalg.inv(V)
        sx, sy = self.nodex / 2, self.nodey / 2
        cx = np.concatenate([sx - 0.5, sx - 0.5, sx + 0.5, sx + 0.5])
        cy = np.concatenate([sy - 0.5, sy + 0.5, sy - 0.5, sy + 0.5])
        n = p * p
        W1 = np.zeros((4 * n, 4 * n))
        for i in range(4):
            W1[i * n:(i + 1) * n, i * n:(i + 1) * n] = self.VI.T
        self.M2MC = chebv2(cx, cy, p).T @ W1            # (p^2, 4 p^2)
        self.L2LC = self.M2MC.T


class _Identity:
    """partial_multipole_to_multipole / partial_local_to_local (fmm.py:58-61)."""

    def __call__(self, b):
        return b


class BBPrecomputation:
    """BB_Precomputation (bbfmm.py:95-120) for one width."""

    def __init__(self, width, K, info):
        self.width = width
        self.p = info.p
        self.M2MC = info.M2MC
        self.L2LC = info.L2LC
        self.width_adj = 2.0 / width
        snx, sny = info.nodex * 0.5 * width, info.nodey * 0.5 * width
        self.M2L = np.empty((7, 7), dtype=object)
        self.M2LF = np.empty((7, 7), dtype=object)
        for ix in range(7):
            for iy in range(7):
                if abs(ix - 3) <= 1 and abs(iy - 3) <= 1:
                    continue
                W1 = kernel_form(K, snx + (ix - 3) * width, sny + (iy - 3) * width, snx, sny)
                m = info.VI @ W1 @ info.VI.T
                self.M2L[ix, iy] = m
                self.M2LF[ix, iy] = M2LEvaluator(m)
        self.S2L_Solver = _Identity()
        self.L2S_Solver = _Identity()


_INFO = {}
_PRECOMP = {}


def get_bb_precomputation(K, p, width):
    if p not in _INFO:
        _INFO[p] = BBInfo(p)
    key = (K, p, round(width, 12))
    if key not in _PRECOMP:
        _PRECOMP[key] = BBPrecomputation(width, K, _INFO[p])
    return _PRECOMP[key]


class BBFMM(_KIFMM):
    """BB_FMM (bbfmm.py:200-231) with the generic passes of flexmm2d/fmm.py:33-177.
    evaluate_to_points returns a 1-D potential, as flexmm2d's does (fmm.py:177)."""

    def __init__(self, x, y, K, Ncutoff=50, p=8, bbox=None, widths=()):
        self.x = np.asarray(x, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.K = K
        self.p = p
        self.Nequiv = p * p
        self.Ncutoff = Ncutoff
        self.bbox = get_bbox(self.x, self.y, bbox, widths)
        self.tree = Tree(self.x, self.y, Ncutoff, self.bbox)
        self.precs = [get_bb_precomputation(K, p, lev.width) for lev in self.tree.Levels]
        self.evaluators = {"potential_evaluation": (lambda sx, sy, tx, ty: (K(sx, sy, tx, ty),), 1,
                                                    None)}

    def _s2m(self, lev, boxes, prec, partial, chunk=4_000_000):
        """source2pMs (bbfmm.py:173-184): Chebyshev moments per box."""
        tree = self.tree
        cnt = lev.top_ind[boxes] - lev.bot_ind[boxes]
        owner = np.repeat(np.arange(boxes.size), cnt)
        starts = np.repeat(lev.bot_ind[boxes] - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
        pos = np.arange(owner.size) + starts
        n = self.Nequiv
        step = max(1, chunk // n)
        acc = np.zeros((boxes.size, n))
        for s in range(0, pos.size, step):
            sl = slice(s, s + step)
            pp = pos[sl]
            ob = boxes[owner[sl]]
            sx = (tree.x[pp] - lev.xmid[ob]) * prec.width_adj
            sy = (tree.y[pp] - lev.ymid[ob]) * prec.width_adj
            vals = chebv2(sx, sy, self.p) * self.tau_ordered[pp][:, None]
            _segment_add(acc, owner[sl], vals)
        partial[lev.this_density_ind[boxes]] += acc

    def _l2t(self, x, y, inds, locs, Kv, pot, chunk=4_000_000):
        """L2targets (bbfmm.py:185-199)."""
        tree = self.tree
        for ind in range(tree.levels):
            sel = np.flatnonzero(inds == ind)
            if sel.size == 0 or self.LE[ind] is None:
                continue
            lev = tree.Levels[ind]
            prec = self.precs[ind]
            step = max(1, chunk // self.Nequiv)
            for s in range(0, sel.size, step):
                ss = sel[s:s + step]
                lc = locs[ss]
                tx = (x[ss] - lev.xmid[lc]) * prec.width_adj
                ty = (y[ss] - lev.ymid[lc]) * prec.width_adj
                pot[0, ss] += cheb_eval2d(tx, ty, self.LE[ind][lc], self.p)

    def evaluate_to_points(self, x, y, name="potential_evaluation", check=False):
        return super().evaluate_to_points(x, y, name, check)[0]
