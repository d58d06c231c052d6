"""Pin the oracle quadtree (restatement of flexmm/tree.py) independently of
its own code: invariants, O(n^2) brute force from node geometry, closed-form
counts of a full uniform tree, and planted mistakes that these checks catch."""
import numpy as np
import pytest

from oracle import tree as otree
from tests import helpers as H


def _trees():
    x, y, _ = H.uniform(6000, 31)
    yield "uniform", x, y, 40, [0, 1, 0, 1]
    x, y, _ = H.clustered(8000, 4, 0.004, 32)
    yield "clustered", x, y, 25, [0, 1, 0, 1]


def cell_bounds(tree, lev, i):
    w = lev.width
    return lev.xmid[i] - 0.5 * w, lev.xmid[i] + 0.5 * w, lev.ymid[i] - 0.5 * w, lev.ymid[i] + 0.5 * w


def check_invariants(t, ncut):
    L = t.levels
    for d, lev in enumerate(t.Levels):
        n = lev.n_node
        # (1) every node's points lie in its cell (ties go to the low side)
        for i in range(n):
            b, e = lev.bot_ind[i], lev.top_ind[i]
            if e > b:
                x0, x1, y0, y1 = cell_bounds(t, lev, i)
                px, py = t.x[b:e], t.y[b:e]
                tol = 1e-12 * lev.width
                assert np.all(px >= x0 - tol) and np.all(px <= x1 + tol)
                assert np.all(py >= y0 - tol) and np.all(py <= y1 + tol)
        # (2) children partition the parent's range in child order
        if d + 1 < L:
            child = t.Levels[d + 1]
            for i in np.flatnonzero(lev.children_ind >= 0):
                c = lev.children_ind[i]
                assert child.bot_ind[c] == lev.bot_ind[i]
                assert child.top_ind[c + 3] == lev.top_ind[i]
                for j in range(3):
                    assert child.top_ind[c + j] == child.bot_ind[c + j + 1]
        # (3) real leaves hold at most ncut points
        real_leaf = lev.leaf & ~lev.fake_leaf
        assert np.all(lev.ns[real_leaf & ~lev.Xlist] <= ncut) or d == otree.MAX_DEPTH
    # (4) 2:1 corner balance: a real leaf has no same-level colleague with a grandchild
    depth = {}
    for d in reversed(range(L)):
        lev = t.Levels[d]
        dd = np.zeros(lev.n_node, dtype=int)
        if d + 1 < L:
            nxt = depth[d + 1]
            for i in range(lev.n_node):
                c = lev.children_ind[i]
                if c >= 0 and not lev.Xlist[i]:
                    dd[i] = 1 + max(nxt[c:c + 4])
        depth[d] = dd
    for d in range(1, L):
        lev = t.Levels[d]
        for i in range(lev.n_node):
            if lev.leaf[i] and not lev.fake_leaf[i]:
                for c in lev.colleagues[i]:
                    if c >= 0 and not lev.fake_leaf[c]:
                        assert depth[d][c] <= 1, (d, i, c)
    # (5) X-list closure: every colleague of a real non-leaf node has children
    for d in range(L - 1):
        lev = t.Levels[d]
        for i in range(lev.n_node):
            if not lev.leaf[i] and not lev.fake_leaf[i]:
                for c in lev.colleagues[i]:
                    if c >= 0:
                        assert lev.children_ind[c] >= 0, (d, i, c)


def brute_colleagues(lev):
    dx = lev.xmid[:, None] - lev.xmid[None, :]
    dy = lev.ymid[:, None] - lev.ymid[None, :]
    return (dx * dx + dy * dy) < (1.5 * lev.width) ** 2


def brute_m2l(t, d):
    """Partners of each real node at level d: same-level nodes with a
    multipole row, not adjacent to it, whose parent is adjacent to (and not)
    its parent; returned as sets of node indices."""
    lev = t.Levels[d]
    par = t.Levels[d - 1]
    adj = brute_colleagues(lev)
    padj = brute_colleagues(par)
    out = {}
    for i in range(lev.n_node):
        if lev.fake_leaf[i]:
            continue
        pi = lev.parent_ind[i]
        s = set()
        for j in range(lev.n_node):
            pj = lev.parent_ind[j]
            if lev.this_density_ind[j] >= 0 and not adj[i, j] and padj[pi, pj] and pi != pj:
                s.add(j)
        out[i] = s
    return out


@pytest.mark.parametrize("name,x,y,ncut,bbox", list(_trees()))
def test_invariants(name, x, y, ncut, bbox):
    t = otree.Tree(x, y, ncut, bbox)
    check_invariants(t, ncut)


@pytest.mark.parametrize("name,x,y,ncut,bbox", list(_trees()))
def test_colleagues_brute_force(name, x, y, ncut, bbox):
    t = otree.Tree(x, y, ncut, bbox)
    for lev in t.Levels:
        adj = brute_colleagues(lev)
        for i in range(lev.n_node):
            got = set(int(c) for c in lev.colleagues[i] if c >= 0)
            assert got == set(np.flatnonzero(adj[i]).tolist())


@pytest.mark.parametrize("name,x,y,ncut,bbox", list(_trees()))
def test_interaction_lists_brute_force(name, x, y, ncut, bbox):
    t = otree.Tree(x, y, ncut, bbox)
    for d in range(2, t.levels):
        lev = t.Levels[d]
        rowmap = {int(r): j for j, r in enumerate(lev.this_density_ind) if r >= 0}
        ref = brute_m2l(t, d)
        dil = t.interaction_lists[d]
        for i, s in ref.items():
            got = set(rowmap[int(r)] for r in dil[:, :, i].ravel() if r >= 0)
            assert got == s, (d, i)


def test_full_uniform_counts():
    # 4^6 boxes x 64 points: every level-5 box splits, every level-6 box is a leaf
    x, y, _ = H.uniform(4 ** 6 * 64, 41)
    t = otree.Tree(x, y, 150, [0, 1, 0, 1])
    assert [l.n_node for l in t.Levels] == [4 ** d for d in range(7)]
    # M2L pairs of a full level d: sum over boxes of in-domain partners in the
    # 6x6 window of the parent's neighbourhood minus the 3x3; closed form per
    # axis: window positions a = 6 per child minus edge losses.
    for d in range(2, 7):
        n = 2 ** d
        i = np.arange(n)
        # number of in-domain window columns per coordinate, for a box at i:
        lo = 2 * (i // 2) - 2
        hi = 2 * (i // 2) + 3
        win = np.minimum(hi, n - 1) - np.maximum(lo, 0) + 1
        near = np.minimum(i + 1, n - 1) - np.maximum(i - 1, 0) + 1
        total = int((win[:, None] * win[None, :] - near[:, None] * near[None, :]).sum())
        got = int((t.interaction_lists[d] >= 0).sum())
        assert got == total, (d, got, total)


def test_c2_pair_counts_closed_form():
    """The closed-form counts quoted for a full 8-level tree (1M uniform,
    Ncutoff 100): M2L pairs per level 2..7."""
    expect = {7: 431676, 6: 105276, 5: 25020, 4: 5628, 3: 1116, 2: 156}
    for d, e in expect.items():
        n = 2 ** d
        i = np.arange(n)
        lo = 2 * (i // 2) - 2
        hi = 2 * (i // 2) + 3
        win = np.minimum(hi, n - 1) - np.maximum(lo, 0) + 1
        near = np.minimum(i + 1, n - 1) - np.maximum(i - 1, 0) + 1
        assert int((win[:, None] * win[None, :] - near[:, None] * near[None, :]).sum()) == e


# --- planted mistakes: each must be caught by one of the checks above -------

def check_ties_low(t):
    """A point on a cell's lower edge belongs to the lower neighbour (x > xmid
    is strict, tree.py:32-33), so it is never on the lower edge of its own
    cell unless that edge is the root box edge."""
    for d, lev in enumerate(t.Levels):
        for i in range(lev.n_node):
            b, e = lev.bot_ind[i], lev.top_ind[i]
            if e > b:
                x0, x1, y0, y1 = cell_bounds(t, lev, i)
                assert np.all((t.x[b:e] > x0) | (x0 <= 0.0))
                assert np.all((t.y[b:e] > y0) | (y0 <= 0.0))


def _tie_points():
    g = np.arange(0, 16) / 16.0
    gx, gy = np.meshgrid(g, g, indexing="ij")
    return gx.ravel(), gy.ravel()


def test_ties_go_low():
    x, y = _tie_points()
    check_ties_low(otree.Tree(x, y, 4, [0, 1, 0, 1]))


def test_planted_tie_rule_is_caught(monkeypatch):
    """Classifying with >= instead of > puts mid-line points on the high side."""
    x, y = _tie_points()
    orig = otree.divide_and_reorder

    def wrong(x_, y_, ordv, tosplit, level, children_start, xlist):
        xm, ym = level.xmid.copy(), level.ymid.copy()
        level.xmid = np.nextafter(xm, -np.inf)   # x > nextafter(xmid) == x >= xmid
        level.ymid = np.nextafter(ym, -np.inf)
        try:
            return orig(x_, y_, ordv, tosplit, level, children_start, xlist)
        finally:
            level.xmid, level.ymid = xm, ym

    monkeypatch.setattr(otree, "divide_and_reorder", wrong)
    t = otree.Tree(x, y, 4, [0, 1, 0, 1])
    with pytest.raises(AssertionError):
        check_ties_low(t)


def test_planted_missing_balance_is_caught(monkeypatch):
    x, y, _ = H.clustered(8000, 4, 0.004, 32)
    monkeypatch.setattr(otree.Tree, "level_restrict", lambda self: None)
    t = otree.Tree(x, y, 25, [0, 1, 0, 1])
    with pytest.raises(AssertionError):
        check_invariants(t, 25)


def test_planted_wrong_colleague_radius_is_caught(monkeypatch):
    x, y, _ = H.uniform(6000, 31)
    orig = otree.Level.tag_colleagues

    def wrong(self, ancestor=None):
        self.width *= 1.5   # radius 2.25 w: picks up next-nearest boxes
        try:
            orig(self, ancestor)
        finally:
            self.width /= 1.5

    monkeypatch.setattr(otree.Level, "tag_colleagues", wrong)
    with pytest.raises(RuntimeError):   # more than 9 colleagues
        otree.Tree(x, y, 40, [0, 1, 0, 1])


def test_fast_tree_comparison_accepts_equal_and_catches_planted_differences():
    """The vectorised full-size tree comparison used by the GPU tests at the
    headline sizes: it accepts the oracle's own tree laid out as device arrays
    (X-list and balance nodes included) and rejects a moved point, a dropped
    colleague, a dropped M2L partner and a flipped flag."""
    from tests import helpers as H
    x, y, _ = H.clustered(6000, 4, 0.004, 17)
    t = otree.Tree(x, y, 20, [0, 1, 0, 1])
    assert any(lev.fake_leaf.any() for lev in t.Levels)
    dev = H.OracleAsDevice(t)
    assert H.compare_trees_fast(dev, t) == sum(lev.n_node for lev in t.Levels)
    for name, plant in (
            ("perm", lambda a: a.__setitem__(slice(None), np.roll(a, 1))),
            ("colleagues", lambda a: a.__setitem__(int(np.flatnonzero(a >= 0)[-1]), -1)),
            ("m2l", lambda a: a.__setitem__(int(np.flatnonzero(a >= 0)[7]), -1)),
            ("flags", lambda a: a.__setitem__(5, a[5] ^ H.NF_UP))):
        bad = H.OracleAsDevice(t)
        plant(bad._a[name])
        with pytest.raises(AssertionError):
            H.compare_trees_fast(bad, t)
