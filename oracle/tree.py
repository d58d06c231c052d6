"""CPU restatement of flexmm's adaptive quadtree -- TEST INFRASTRUCTURE ONLY.

Follows ``flexmm/tree.py`` of the reference step by step (citations inline),
vectorised with NumPy.  Deliberate differences, none of which changes any
node set, flag, colleague set, interaction-list set or per-node point set:

* the 4-way in-place partition (tree.py:13-125) is replaced by a stable
  partition by class; the within-node point order therefore differs, which
  is why trees are compared canonically (nodes keyed by (level, ix, iy),
  point *sets* per node);
* refinement stops at depth ``MAX_DEPTH`` (31); the reference has no cap and
  loops forever on more than ``ppl`` coincident points (tree.py:204,219);
* interaction lists are not formed for fake (X-list) nodes whose parent's
  colleague has no children: the reference then reads wrapped indices
  (tree.py:418-421 with ``children_ind == -1``) but the resulting partial
  locals of fake nodes are never used.

Pinned by (tests/test_oracle_tree.py): invariants (leaf capacity, the ranges
partition [0, N), 2:1 corner balance, X-list closure), brute-force O(n^2)
colleague and interaction-list construction from node centres, exact point
membership by the FP64 midpoint recurrence, and the analytic node and M2L
pair counts of a full uniform tree.
"""
import numpy as np

TREE_SEARCH_CROSSOVER = 100   # tree.py:11
MAX_DEPTH = 31                # depth cap (the reference has none)


def _left_pack(cand, mask, width):
    """Pack the True entries of each row of ``cand`` to the left of a -1 row."""
    n = cand.shape[0]
    out = -np.ones((n, width), dtype=np.int64)
    if n == 0:
        return out
    cnt = mask.sum(axis=1)
    if cnt.max(initial=0) > width:
        raise RuntimeError("more than %d colleagues" % width)
    if cand.shape[1] < width:
        pad = width - cand.shape[1]
        cand = np.concatenate([cand, -np.ones((n, pad), dtype=cand.dtype)], axis=1)
        mask = np.concatenate([mask, np.zeros((n, pad), dtype=bool)], axis=1)
    order = np.argsort(~mask, axis=1, kind="stable")
    packed = np.take_along_axis(cand, order, axis=1)[:, :width]
    keep = np.arange(width)[None, :] < cnt[:, None]
    out[keep] = packed[keep]
    return out


class Level:
    """Nodes of one level; tree.py:275-360."""

    def __init__(self, xmin, ymin, width, bot_ind, top_ind, parent_ind, ix, iy):
        self.xmin = xmin
        self.ymin = ymin
        self.width = width
        self.half_width = 0.5 * width
        self.bot_ind = bot_ind
        self.top_ind = top_ind
        self.parent_ind = parent_ind
        self.ix = ix                      # canonical integer box coordinates
        self.iy = iy
        n = xmin.shape[0]
        self.leaf = np.ones(n, dtype=bool)
        self.fake_leaf = np.zeros(n, dtype=bool)
        self.children_ind = -np.ones(n, dtype=np.int64)
        self.basic_computations()

    def basic_computations(self):
        # tree.py:301-308
        self.ns = self.top_ind - self.bot_ind
        self.xmid = self.xmin + self.half_width
        self.ymid = self.ymin + self.half_width
        self.n_node = self.xmin.shape[0]
        self.short_parent_ind = self.parent_ind[::4]

    def tag_colleagues(self, ancestor=None):
        # tree.py:309-335: direct O(n^2) search below the crossover, else
        # children of the parent's colleagues (a node counts as having
        # children iff children_ind >= 0, tree.py:331).
        dist2 = (1.5 * self.width) ** 2
        n = self.n_node
        if n < TREE_SEARCH_CROSSOVER or ancestor is None:
            dx = self.xmid[:, None] - self.xmid[None, :]
            dy = self.ymid[:, None] - self.ymid[None, :]
            mask = dx * dx + dy * dy < dist2
            cand = np.broadcast_to(np.arange(n)[None, :], (n, n))
            self.colleagues = _left_pack(cand, mask, 9)
            return
        pc = ancestor.colleagues[self.parent_ind]                  # (n, 9)
        pch = np.where(pc >= 0, ancestor.children_ind[np.maximum(pc, 0)], -1)
        cand = (pch[:, :, None] + np.arange(4)[None, None, :]).reshape(n, 36)
        valid = np.repeat(pch >= 0, 4, axis=1)
        safe = np.where(valid, cand, 0)
        dx = self.xmid[:, None] - self.xmid[safe]
        dy = self.ymid[:, None] - self.ymid[safe]
        mask = valid & (dx * dx + dy * dy < dist2)
        self.colleagues = _left_pack(cand, mask, 9)

    def get_depths(self, descendant):
        # tree.py:340-343 with numba_get_depths tree.py:362-369
        self.depths = np.zeros(self.n_node, dtype=np.int64)
        if descendant is not None:
            nl = np.flatnonzero(~self.leaf)
            if nl.size:
                ci = self.children_ind[nl]
                cd = descendant.depths[ci[:, None] + np.arange(4)[None, :]]
                self.depths[nl] = cd.max(axis=1) + 1

    def add_nodes(self, xmin, ymin, bot_ind, top_ind, parent_ind, ix, iy, xlist):
        # tree.py:344-355
        m = xmin.shape[0]
        self.xmin = np.concatenate([self.xmin, xmin])
        self.ymin = np.concatenate([self.ymin, ymin])
        self.bot_ind = np.concatenate([self.bot_ind, bot_ind])
        self.top_ind = np.concatenate([self.top_ind, top_ind])
        self.parent_ind = np.concatenate([self.parent_ind, parent_ind])
        self.ix = np.concatenate([self.ix, ix])
        self.iy = np.concatenate([self.iy, iy])
        self.leaf = np.concatenate([self.leaf, np.full(m, not xlist)])
        self.fake_leaf = np.concatenate([self.fake_leaf, np.full(m, xlist)])
        self.children_ind = np.concatenate([self.children_ind, -np.ones(m, dtype=np.int64)])
        self.basic_computations()

    def get_xlist(self):
        # tree.py:356-358 with numba_get_Xlist tree.py:385-397: only nodes that
        # had depths computed (i < n, cj < n) take part.
        n = self.depths.shape[0]
        self.Xlist = np.zeros(self.n_node, dtype=bool)
        col = self.colleagues[:n]
        valid = (col >= 0) & (col < n)
        cd = np.where(valid, self.depths[np.where(valid, col, 0)], 0)
        lower = valid & (self.depths[:n, None] - cd < 0)
        self.Xlist[:n] = self.leaf[:n] & lower.any(axis=1)


def divide_and_reorder(x, y, ordv, tosplit, level, children_start, xlist):
    """tree.py:127-189 (partition by class 2*(x>xmid)+(y>ymid), tree.py:32-34).

    Returns the arrays of the 4*nsplit new nodes; updates ``level.leaf``
    (unless ``xlist``) and ``level.children_ind`` in place, and reorders
    x, y, ordv within each split node.
    """
    sp = np.flatnonzero(tosplit)
    ns = sp.size
    hw = level.half_width
    if ns == 0:
        e = np.empty(0)
        ei = np.empty(0, dtype=np.int64)
        return e, e, ei, ei, ei, ei, ei
    bot = level.bot_ind[sp]
    top = level.top_ind[sp]
    cnt = top - bot
    owner = np.repeat(np.arange(ns), cnt)
    starts = np.repeat(bot - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
    pos = np.arange(owner.size) + starts
    cls = 2 * (x[pos] > level.xmid[sp][owner]).astype(np.int64) + \
        (y[pos] > level.ymid[sp][owner]).astype(np.int64)
    order = np.lexsort((cls, owner))
    src = pos[order]
    x[pos] = x[src]
    y[pos] = y[src]
    ordv[pos] = ordv[src]
    per = np.bincount(owner * 4 + cls, minlength=4 * ns).reshape(ns, 4)
    nns = np.concatenate([np.zeros((ns, 1), dtype=np.int64), np.cumsum(per, axis=1)], axis=1)
    new_bot = (bot[:, None] + nns[:, :4]).ravel()
    new_top = (bot[:, None] + nns[:, 1:]).ravel()
    xm = level.xmid[sp]
    ym = level.ymid[sp]
    new_xmin = np.stack([xm - hw, xm - hw, xm, xm], axis=1).ravel()   # tree.py:169-172
    new_ymin = np.stack([ym - hw, ym, ym - hw, ym], axis=1).ravel()   # tree.py:173-176
    parent = np.repeat(sp, 4)
    j = np.tile(np.arange(4), ns)
    new_ix = 2 * np.repeat(level.ix[sp], 4) + (j >> 1)
    new_iy = 2 * np.repeat(level.iy[sp], 4) + (j & 1)
    if not xlist:
        level.leaf[sp] = False                                         # tree.py:185-186
    level.children_ind[sp] = children_start + 4 * np.arange(ns)       # tree.py:189
    return new_xmin, new_ymin, new_bot, new_top, parent, new_ix, new_iy


class Tree:
    """tree.py:447-589 (and locate_points tree.py:703-727)."""

    def __init__(self, x, y, ppl, bbox):
        self.x = np.array(x, dtype=np.float64, copy=True)
        self.y = np.array(y, dtype=np.float64, copy=True)
        self.points_per_leaf = ppl
        self.bbox = bbox
        self.xmin, self.xmax, self.ymin, self.ymax = [float(b) for b in bbox]  # tree.py:473-476
        self.N = self.x.shape[0]
        self.ordv = np.arange(self.N)
        width = self.xmax - self.xmin                                   # tree.py:484
        z = np.zeros(1, dtype=np.int64)
        level0 = Level(np.array([self.xmin]), np.array([self.ymin]), width,
                       z.copy(), np.array([self.N]), np.array([-1]), z.copy(), z.copy())
        self.Levels = [level0]
        if self.N > ppl:                                                # tree.py:490-497
            cur = level0
            keep_going = True
            while keep_going:
                depth = len(self.Levels) - 1
                to_split = (cur.ns > ppl) & (depth < MAX_DEPTH)
                out = divide_and_reorder(self.x, self.y, self.ordv, to_split, cur, 0, False)
                new = Level(out[0], out[1], cur.half_width, out[2], out[3], out[4], out[5], out[6])
                self.Levels.append(new)
                cur = new
                keep_going = bool(np.any(new.ns > ppl)) and depth + 1 < MAX_DEPTH
        self.levels = len(self.Levels)
        self.tag_colleagues()
        self.gather_depths()
        self.level_restrict()
        self.split_xlist()
        self.post_process()
        self.build_interaction_lists()

    def tag_colleagues(self):
        # tree.py:519-525
        for ind, lev in enumerate(self.Levels):
            lev.tag_colleagues(None if ind == 0 else self.Levels[ind - 1])

    def gather_depths(self):
        # tree.py:578-581
        for ind in reversed(range(self.levels)):
            desc = None if ind == self.levels - 1 else self.Levels[ind + 1]
            self.Levels[ind].get_depths(desc)

    def _split(self, ind, mask, xlist):
        # split_bad_leaves, tree.py:258-273
        lev = self.Levels[ind]
        desc = self.Levels[ind + 1]
        out = divide_and_reorder(self.x, self.y, self.ordv, mask, lev, desc.n_node, xlist)
        desc.add_nodes(*out, xlist)
        desc.tag_colleagues(lev)

    def level_restrict(self):
        # tree.py:529-541 with numba_get_bads tree.py:371-383
        new_nodes = 1
        while new_nodes > 0:
            new_nodes = 0
            for ind in range(1, self.levels - 2):
                lev = self.Levels[ind]
                col = lev.colleagues
                valid = col >= 0
                cd = np.where(valid, lev.depths[np.where(valid, col, 0)], 0)
                diff = lev.depths[:, None] - cd
                bads = lev.leaf & (valid & ((diff > 1) | (diff < -1))).any(axis=1)
                nb = int(bads.sum())
                self._split(ind, bads, False)
                self.gather_depths()
                new_nodes += nb

    def split_xlist(self):
        # tree.py:542-553
        for ind in range(self.levels):
            lev = self.Levels[ind]
            if ind == 0 or ind == self.levels - 1:
                lev.Xlist = np.zeros(lev.n_node, dtype=bool)
            else:
                lev.get_xlist()
                if lev.Xlist.any():
                    self._split(ind, lev.Xlist, True)

    def post_process(self):
        # tree.py:554-577
        for lev in self.Levels:
            ns = lev.top_ind - lev.bot_ind
            lev.not_leaf = ~lev.leaf
            lev.compute_upwards = ((lev.leaf & ~lev.Xlist) | lev.fake_leaf) & (ns > 0)
            lev.compute_downwards = (lev.not_leaf | lev.Xlist) & ~lev.fake_leaf
            lev.has_source = ns > 0
            if lev is not self.Levels[0]:
                sib = lev.has_source.reshape(-1, 4).any(axis=1)
                lev.sib_has_source = sib
                lev.n_sib_has_source = int(sib.sum())
                lev.allocate_density = np.repeat(sib, 4)
                lev.n_allocate_density = int(lev.allocate_density.sum())
                lev.parent_density_ind = -np.ones(lev.n_node // 4, dtype=np.int64)
                lev.parent_density_ind[sib] = np.arange(lev.n_sib_has_source)
                lev.this_density_ind = -np.ones(lev.n_node, dtype=np.int64)
                lev.this_density_ind[lev.allocate_density] = np.arange(lev.n_allocate_density)
            else:
                lev.this_density_ind = -np.ones(lev.n_node, dtype=np.int64)
                lev.n_allocate_density = 0

    def build_interaction_lists(self):
        # tree.py:582-589 with build_interaction_list tree.py:399-435:
        # dilists[dx+3, dy+3, i] = multipole row of the child of a parent's
        # colleague at offset (dx, dy), for the 40 offsets of the 7x7 ring.
        self.interaction_lists = [None] * self.levels
        for ind in range(2, self.levels):
            lev = self.Levels[ind]
            par = self.Levels[ind - 1]
            n = lev.n_node
            dil = -np.ones((7, 7, n), dtype=np.int64)
            pc = par.colleagues[lev.parent_ind]                       # (n, 9)
            notself = (pc >= 0) & (pc != lev.parent_ind[:, None])
            pch = np.where(notself, par.children_ind[np.maximum(pc, 0)], -1)
            ok = notself & (pch >= 0)
            cand = (pch[:, :, None] + np.arange(4)[None, None, :]).reshape(n, 36)
            ok = np.repeat(ok, 4, axis=1)
            safe = np.where(ok, cand, 0)
            xd = np.rint((lev.xmid[safe] - lev.xmid[:, None]) / lev.width).astype(np.int64)
            yd = np.rint((lev.ymid[safe] - lev.ymid[:, None]) / lev.width).astype(np.int64)
            di = lev.this_density_ind[safe]
            ring = (np.abs(xd) <= 3) & (np.abs(yd) <= 3) & ~((np.abs(xd) <= 1) & (np.abs(yd) <= 1))
            keep = ok & (di >= 0) & ring
            rows, cols = np.nonzero(keep)
            dil[xd[rows, cols] + 3, yd[rows, cols] + 3, rows] = di[rows, cols]
            self.interaction_lists[ind] = dil

    def locate_points(self, x, y):
        # tree.py:703-727: descend while not leaf, child += 2*(x>xmid) + (y>ymid)
        x = np.asarray(x, dtype=np.float64).ravel()
        y = np.asarray(y, dtype=np.float64).ravel()
        inds = np.zeros(x.size, dtype=np.int64)
        locs = np.zeros(x.size, dtype=np.int64)
        active = np.arange(x.size)
        ind = 0
        while active.size:
            lev = self.Levels[ind]
            loc = locs[active]
            stop = lev.leaf[loc]
            done = active[stop]
            inds[done] = ind
            active = active[~stop]
            loc = loc[~stop]
            if active.size == 0:
                break
            c = lev.children_ind[loc] + 2 * (x[active] > lev.xmid[loc]) + (y[active] > lev.ymid[loc])
            locs[active] = c
            ind += 1
        return inds, locs

    # ------------------------------------------------------------------
    # canonical view used by the parity tests
    def canonical(self):
        """Per-level dictionaries keyed by (ix, iy).

        Each value holds the reference flags, the sorted original indices of
        the node's points, its colleagues as a set of (ix, iy), whether it
        has a multipole row, and (levels >= 2, non-fake nodes) its M2L
        partners as a set of (dx, dy, ix, iy).
        """
        out = []
        for ind, lev in enumerate(self.Levels):
            d = {}
            rowmap = {int(r): j for j, r in enumerate(lev.this_density_ind) if r >= 0}
            for i in range(lev.n_node):
                key = (int(lev.ix[i]), int(lev.iy[i]))
                col = lev.colleagues[i]
                col = col[col >= 0]
                rec = dict(
                    leaf=bool(lev.leaf[i]), fake=bool(lev.fake_leaf[i]),
                    xlist=bool(lev.Xlist[i]),
                    compute_upwards=bool(lev.compute_upwards[i]),
                    compute_downwards=bool(lev.compute_downwards[i]),
                    points=tuple(np.sort(self.ordv[lev.bot_ind[i]:lev.top_ind[i]]).tolist()),
                    colleagues=frozenset((int(lev.ix[c]), int(lev.iy[c])) for c in col),
                    has_density=bool(lev.this_density_ind[i] >= 0),
                    xmid=float(lev.xmid[i]), ymid=float(lev.ymid[i]),
                )
                if ind >= 2 and not lev.fake_leaf[i]:
                    dil = self.interaction_lists[ind][:, :, i]
                    parts = set()
                    for a in range(7):
                        for b in range(7):
                            r = dil[a, b]
                            if r >= 0:
                                j = rowmap[int(r)]
                                parts.add((a - 3, b - 3, int(lev.ix[j]), int(lev.iy[j])))
                    rec["m2l"] = frozenset(parts)
                d[key] = rec
            out.append(d)
        return out
