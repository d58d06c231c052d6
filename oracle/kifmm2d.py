"""CPU restatement of flexmm's scalar 2D KIFMM -- TEST INFRASTRUCTURE ONLY.

Follows ``flexmm/kifmm2d/scalar/fmm.py`` and ``helpers.py`` of the reference
(citations inline), in NumPy FP64.  Differences from the reference:

* kernels are vectorised NumPy functions ``K(sx, sy, tx, ty)`` (oracle.kernels)
  instead of numba-jitted scalar functions;
* the bounding box follows the fresh-helper rule of ``Helper.get_bbox``
  (helpers.py:68-88, ``widths`` empty) unless cached widths are passed;
* precomputations are cached per (kernel, p, round(width, 12)) in a module
  dict, the FloatDict rule of float_dict.py:23-24, but never shared between
  FMM instances' per-level surface lists (helpers.py:95-108 overwrites them);
* level 2 is given one partial-local row per node (fmm.py:139 allocates 16).

Pinned by (tests/test_oracle_fmm.py): FP64 direct sums on all targets at
small N (potential and gradient), convergence of the error with p, the
analytic structure of the precomputed operators (closed-form entries,
pseudo-inverse identities, SVD reconstruction), and planted mistakes in the
restatement that those checks must catch.
"""
import numpy as np

from .tree import Tree

SVD_TOL_M2L = 1e-14      # helpers.py:251
SVD_TOL_PINV = 1e-15     # helpers.py:241


def get_level_information(node_width, theta):
    """helpers.py:5-15 (dd = 0)."""
    r1 = 0.5 * node_width * np.sqrt(2)
    r2 = 0.5 * node_width * (4 - np.sqrt(2))
    return (r1 * np.cos(theta), r1 * np.sin(theta),
            r2 * np.cos(theta), r2 * np.sin(theta), r1, r2)


def kernel_form(K, sx, sy, tx, ty):
    """out[j, i] = K(s_i, t_j); helpers.py:129-153 (no self-diagonal case here)."""
    return K(sx[None, :], sy[None, :], tx[:, None], ty[:, None])


class SVDSolver:
    """helpers.py:240-248: x = V diag(1/s) U^T b with s < tol treated as inf."""

    def __init__(self, A, tol=SVD_TOL_PINV):
        self.U, S, self.VH = np.linalg.svd(A)
        S = S.copy()
        S[S < tol] = np.inf
        self.SI = 1.0 / S
        self.rank = int(np.isfinite(S).sum())

    def __call__(self, b):
        mult = self.SI[:, None] if b.ndim > 1 else self.SI
        return self.VH.T.dot(mult * self.U.T.dot(b))


class M2LEvaluator:
    """helpers.py:17-45: SVD compression if 2*rank < p, dense if rank > 0."""

    def __init__(self, m2l, svd_tol=SVD_TOL_M2L):
        self.m2l = m2l
        U, S, VH = np.linalg.svd(m2l)
        cuts = np.where(S > svd_tol)[0]
        self.kind = "null"
        self.rank = 0
        if len(cuts) > 0:
            cutoff = int(cuts[-1] + 1)
            self.rank = cutoff
            if 2 * cutoff < m2l.shape[0]:
                self.M2 = U[:, :cutoff].copy()
                self.M1 = VH[:cutoff] * S[:cutoff][:, None]
                self.kind = "svd"
            else:
                self.kind = "dense"

    def __call__(self, tau):
        if self.kind == "dense":
            return self.m2l @ tau
        if self.kind == "svd":
            return self.M2 @ (self.M1 @ tau)
        return np.zeros((self.m2l.shape[0],) + tau.shape[1:])


class Precomputation:
    """helpers.py:250-317."""

    def __init__(self, width, K, p):
        self.width = width
        theta = np.linspace(0, 2 * np.pi, p, endpoint=False)
        sx, sy, lx, ly, _, _ = get_level_information(width, theta)
        small_to_large = kernel_form(K, sx, sy, lx, ly)
        self.S2L = small_to_large
        self.S2L_Solver = SVDSolver(small_to_large)
        self.L2S_Solver = SVDSolver(kernel_form(K, sx, sy, lx, ly))   # helpers.py:266-267
        ssx, ssy, _, _, _, _ = get_level_information(0.5 * width, theta)
        q = 0.25 * width
        cexs = np.concatenate([ssx - q, ssx - q, ssx + q, ssx + q])      # helpers.py:271-276
        ceys = np.concatenate([ssy - q, ssy + q, ssy - q, ssy + q])      # helpers.py:277-282
        self.M2MC = kernel_form(K, cexs, ceys, lx, ly)                   # (p, 4p)
        self.L2LC = self.M2MC.T                                          # helpers.py:286
        self.M2L = np.empty((7, 7), dtype=object)
        self.M2LF = np.empty((7, 7), dtype=object)
        for ix in range(7):
            for iy in range(7):
                if abs(ix - 3) <= 1 and abs(iy - 3) <= 1:
                    continue
                m = kernel_form(K, sx + (ix - 3) * width, sy + (iy - 3) * width, sx, sy)
                self.M2L[ix, iy] = m
                self.M2LF[ix, iy] = M2LEvaluator(m)
        self.small_x, self.small_y = sx, sy
        self.large_x, self.large_y = lx, ly


_PRECOMP_CACHE = {}


def get_precomputation(K, p, width):
    key = (K, p, round(width, 12))
    if key not in _PRECOMP_CACHE:
        _PRECOMP_CACHE[key] = Precomputation(width, K, p)
    return _PRECOMP_CACHE[key]


def get_bbox(px, py, bbox, widths=()):
    """helpers.py:68-88."""
    if bbox is None:
        bbox = [np.min(px), np.max(px), np.min(py), np.max(py)]
    if len(widths) > 0:
        small_width = np.min(np.asarray(widths, dtype=float))
        bw = max(bbox[1] - bbox[0], bbox[3] - bbox[2])
        width = 2 ** np.ceil(np.log2(bw / small_width)) * small_width
    else:
        width = max(bbox[1] - bbox[0], bbox[3] - bbox[2])
    cx = (bbox[1] + bbox[0]) / 2
    cy = (bbox[3] + bbox[2]) / 2
    return [cx - width / 2, cx + width / 2, cy - width / 2, cy + width / 2]


class FMM:
    """fmm.py:37-199 with the kernel given as (K, K_eval) NumPy functions.

    ``evaluators`` maps a name to a vectorised ``(sx, sy, tx, ty) -> tuple``
    kernel whose outputs are multiplied by the density, the role of
    register_evaluator (helpers.py:176-238).
    """

    def __init__(self, x, y, K, Ncutoff, Nequiv, bbox=None, widths=()):
        self.x = np.asarray(x, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.K = K
        self.Nequiv = Nequiv
        self.Ncutoff = Ncutoff
        self.bbox = get_bbox(self.x, self.y, bbox, widths)
        self.tree = Tree(self.x, self.y, Ncutoff, self.bbox)
        self.precs = [get_precomputation(K, Nequiv, lev.width) for lev in self.tree.Levels]
        self.evaluators = {"potential_evaluation": (lambda sx, sy, tx, ty: (K(sx, sy, tx, ty),), 1,
                                                    None)}

    def register_evaluator(self, Kvec, name, outputs, src=None):
        """Kvec: the L2T kernel applied to the local density (KAS1 of
        helpers.py:176-238).  src (optional): the P2P kernel of the sources,
        ``src(sx, sy, tx, ty, tau_rows, extras_rows) -> tuple`` already
        weighted by the densities (KAS2, e.g. the slp+dlp kernel of
        examples/pure_kifmm_laplace_dipoles.py:32-40); without it P2P is
        Kvec times the single density."""
        self.evaluators[name] = (Kvec, outputs, src)

    # -- upward pass: fmm.py:73-124 -------------------------------------
    def upwards_pass(self, tau, src=None, extras=None):
        """tau: (n,) charges, or (k, n) densities with src, the upwardor's
        kernel ``src(sx, sy, tx, ty, tau_rows, extras_rows) -> value``
        (register_upwardor, helpers.py:155-174; extras (m, n), fmm.py:80-90)."""
        tree = self.tree
        p = self.Nequiv
        tau = np.asarray(tau, dtype=np.float64)
        self.src = src
        if src is None:
            tau = tau.reshape(-1)
            self.tau_ordered = tau[tree.ordv]
            self.tau_rows = self.tau_ordered[None, :]
        else:
            tau = tau.reshape(-1, tree.x.size)
            self.tau_rows = tau[:, tree.ordv]
            self.tau_ordered = self.tau_rows[0]
        self.extras_rows = None if extras is None else \
            np.asarray(extras, dtype=np.float64).reshape(-1, tree.x.size)[:, tree.ordv]
        self.multipoles = [None] * tree.levels
        self.reshaped = [None] * tree.levels
        self.check_potentials = [None] * tree.levels
        for ind in reversed(range(2, tree.levels)):
            lev = tree.Levels[ind]
            prec = self.precs[ind]
            anum = lev.n_allocate_density
            partial = np.zeros((anum, p))
            if ind != tree.levels - 1:
                anc = tree.Levels[ind + 1]
                temp1 = (prec.M2MC @ self.reshaped[ind + 1].T).T
                # distribute, fmm.py:14-18
                ok = anc.parent_density_ind >= 0
                partial[lev.this_density_ind[anc.short_parent_ind[ok]]] = temp1[anc.parent_density_ind[ok]]
            # upwardor (S2M), helpers.py:159-173
            boxes = np.flatnonzero(lev.compute_upwards & (lev.this_density_ind >= 0))
            if boxes.size:
                self._s2m(lev, boxes, prec, partial)
            self.check_potentials[ind] = partial.copy()
            self.multipoles[ind] = prec.S2L_Solver(partial.T).T.copy()
            self.reshaped[ind] = self.multipoles[ind].reshape(anum // 4, 4 * p)

    def _s2m(self, lev, boxes, prec, partial, chunk=4_000_000):
        tree = self.tree
        cnt = lev.top_ind[boxes] - lev.bot_ind[boxes]
        owner = np.repeat(np.arange(boxes.size), cnt)
        starts = np.repeat(lev.bot_ind[boxes] - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
        pos = np.arange(owner.size) + starts
        p = prec.large_x.size
        step = max(1, chunk // p)
        rows = lev.this_density_ind[boxes]
        acc = np.zeros((boxes.size, p))
        for s in range(0, pos.size, step):
            sl = slice(s, s + step)
            pp = pos[sl]
            ob = boxes[owner[sl]]
            sx = tree.x[pp] - lev.xmid[ob]
            sy = tree.y[pp] - lev.ymid[ob]
            if self.src is None:
                vals = self.K(sx[:, None], sy[:, None], prec.large_x[None, :], prec.large_y[None, :])
                vals *= self.tau_ordered[pp][:, None]
            else:
                ex = None if self.extras_rows is None else self.extras_rows[:, pp][:, :, None]
                vals = self.src(sx[:, None], sy[:, None], prec.large_x[None, :],
                                prec.large_y[None, :], self.tau_rows[:, pp][:, :, None], ex)[0]
            _segment_add(acc, owner[sl], vals)
        partial[rows] += acc

    # -- downward pass: fmm.py:125-166 ----------------------------------
    def downwards_pass(self):
        tree = self.tree
        p = self.Nequiv
        L = tree.levels
        self.LE = [None] * L
        self.PLE = [None] * L
        for ind in range(min(2, L)):
            self.LE[ind] = np.zeros((tree.Levels[ind].n_node, p))
        if L > 2:
            self.PLE[2] = np.zeros((tree.Levels[2].n_node, p))
        for ind in range(2, L):
            lev = tree.Levels[ind]
            prec = self.precs[ind]
            dil = tree.interaction_lists[ind]
            Q = self.multipoles[ind]
            for i in range(7):
                for j in range(7):
                    if abs(i - 3) <= 1 and abs(j - 3) <= 1:
                        continue
                    sel = np.flatnonzero(dil[i, j] >= 0)
                    if sel.size == 0 or Q.shape[0] == 0:
                        continue
                    Y = prec.M2LF[i, j](Q.T).T
                    self.PLE[ind][sel] += Y[dil[i, j, sel]]     # add_interactions_prepared, fmm.py:20-27
            self.LE[ind] = prec.L2S_Solver(self.PLE[ind].T).T.copy()
            if ind < L - 1:
                doit = lev.compute_downwards
                ple = (prec.L2LC @ self.LE[ind][doit].T).T
                sorter = np.argsort(lev.children_ind[doit])
                self.PLE[ind + 1] = ple[sorter].reshape(tree.Levels[ind + 1].n_node, p)

    def build_expansions(self, tau, src=None, extras=None):
        """fmm.py:168-170."""
        self.upwards_pass(tau, src, extras)
        self.downwards_pass()

    def source_evaluation(self, x, y):
        return self.evaluate_to_points(x, y, "potential_evaluation", True)

    def target_evaluation(self, x, y):
        return self.evaluate_to_points(x, y, "potential_evaluation", False)

    # -- evaluation: fmm.py:182-199 -------------------------------------
    def evaluate_to_points(self, x, y, name, check):
        tree = self.tree
        Kv, nout, src = self.evaluators[name]
        x = np.asarray(x, dtype=np.float64).ravel()
        y = np.asarray(y, dtype=np.float64).ravel()
        inds, locs = tree.locate_points(x, y)
        pot = np.zeros((nout, x.size))
        self._l2t(x, y, inds, locs, Kv, pot)
        self._p2p(x, y, inds, locs, Kv, pot, check, src)
        return pot

    def _l2t(self, x, y, inds, locs, Kv, pot, chunk=4_000_000):
        # local_expansion_evaluation, helpers.py:219-235
        tree = self.tree
        for ind in range(tree.levels):
            sel = np.flatnonzero(inds == ind)
            if sel.size == 0 or self.LE[ind] is None:
                continue
            lev = tree.Levels[ind]
            prec = self.precs[ind]
            p = prec.large_x.size
            step = max(1, chunk // p)
            for s in range(0, sel.size, step):
                ss = sel[s:s + step]
                lc = locs[ss]
                tx = (x[ss] - lev.xmid[lc])[:, None]
                ty = (y[ss] - lev.ymid[lc])[:, None]
                outs = Kv(prec.large_x[None, :], prec.large_y[None, :], tx, ty)
                le = self.LE[ind][lc]
                for k, o in enumerate(outs):
                    pot[k, ss] += (o * le).sum(axis=1)

    def _p2p(self, x, y, inds, locs, Kv, pot, check, srck=None):
        # neighbor_evaluation, helpers.py:180-217: per target, the sources of
        # the 9 colleagues of its leaf; skip exact coincidences when check.
        tree = self.tree
        sx_all, sy_all, tau = tree.x, tree.y, self.tau_ordered
        key = inds * (1 << 40) + locs
        order = np.argsort(key, kind="stable")
        ks = key[order]
        bounds = np.flatnonzero(np.diff(ks)) + 1
        groups = np.split(order, bounds)
        for g in groups:
            ind = int(inds[g[0]])
            loc = int(locs[g[0]])
            lev = tree.Levels[ind]
            col = lev.colleagues[loc]
            col = col[col >= 0]
            idx = np.concatenate([np.arange(lev.bot_ind[c], lev.top_ind[c]) for c in col]) if col.size else np.empty(0, dtype=np.int64)
            if idx.size == 0:
                continue
            tx = x[g][:, None]
            ty = y[g][:, None]
            sx = sx_all[idx][None, :]
            sy = sy_all[idx][None, :]
            if srck is not None:
                # the sources' own kernel, weighted by their densities (KAS2)
                ex = None if self.extras_rows is None else self.extras_rows[:, idx][:, None, :]
                with np.errstate(divide="ignore", invalid="ignore"):
                    outs = srck(sx, sy, tx, ty, self.tau_rows[:, idx][:, None, :], ex)
                if check:
                    keep = ~((sx == tx) & (sy == ty))
                    outs = [np.where(keep, o, 0.0) for o in outs]
                for k, o in enumerate(outs):
                    pot[k, g] += o.sum(axis=1)
                continue
            with np.errstate(divide="ignore", invalid="ignore"):
                outs = Kv(sx, sy, tx, ty)
            w = np.broadcast_to(tau[idx][None, :], (g.size, idx.size))
            if check:
                keep = ~((sx == tx) & (sy == ty))
                w = np.where(keep, w, 0.0)
                outs = [np.where(keep, o, 0.0) for o in outs]
            for k, o in enumerate(outs):
                pot[k, g] += (o * w).sum(axis=1)


def _segment_add(acc, owner, vals):
    """acc[owner[i]] += vals[i] for sorted owner (contiguous runs)."""
    if owner.size == 0:
        return
    starts = np.concatenate([[0], np.flatnonzero(np.diff(owner)) + 1])
    sums = np.add.reduceat(vals, starts, axis=0)
    acc[owner[starts]] += sums
