"""Quadtree parity: the device tree against the oracle restatement of
flexmm/tree.py, compared canonically (nodes keyed by (level, ix, iy); flags,
point sets, colleagues, M2L partners, and the FP64 node centres bit for bit)."""
import numpy as np
import pytest

from oracle import tree as otree
from oracle import kifmm2d as okifmm
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _build(x, y, ncut, p=32, bbox=None):
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    the_help.clear()
    return FMM(x, y, kernels.laplace, ncut, p, bbox=bbox)


def _oracle_tree(x, y, ncut, bbox):
    return otree.Tree(x, y, ncut, bbox)


CASES = {
    "uniform_20k": lambda: (*H.uniform(20000, 11)[:2], 64, None),
    "uniform_100k_c1": lambda: (*H.uniform(100000, 1)[:2], 64, None),
    "clustered_30k": lambda: (*H.clustered(30000, 5, 0.01, 3)[:2], 50, [0, 1, 0, 1]),
    "clustered_deep": lambda: (*H.clustered(40000, 3, 0.001, 4)[:2], 20, [0, 1, 0, 1]),
    # 21 and 23 levels: the 2:1 balance runs both its bitmap depths (<= 13)
    # and its sorted-candidate depths below them
    "clustered_21_levels": lambda: (*H.clustered(30000, 3, 2e-5, 8)[:2], 20, [0, 1, 0, 1]),
    "clustered_23_levels": lambda: (*H.clustered(24000, 4, 3e-6, 10)[:2], 25, [0, 1, 0, 1]),
    "tiny_root_leaf": lambda: (*H.uniform(50, 5)[:2], 64, None),
    "small_two_levels": lambda: (*H.uniform(100, 6)[:2], 30, None),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_tree_matches_oracle(case):
    x, y, ncut, bbox = CASES[case]()
    print("case", case, "n", x.size, flush=True)
    f = _build(x, y, ncut, bbox=bbox)
    ora = _oracle_tree(x, y, ncut, f.bbox)
    assert f.bbox == okifmm.get_bbox(x, y, bbox)
    n = H.compare_canonical(H.gpu_canonical(f), ora.canonical())
    assert n == sum(l.n_node for l in ora.Levels)


def test_tree_dyadic_ties():
    """Points exactly on cell mid-lines go to the low child (x > xmid is strict)."""
    g = np.arange(0, 64) / 64.0
    x, y = np.meshgrid(g, g, indexing="ij")
    x = np.concatenate([x.ravel(), [0.5, 0.25, 0.75]])
    y = np.concatenate([y.ravel(), [0.5, 0.5, 0.125]])
    f = _build(x, y, 8, bbox=[0, 1, 0, 1])
    ora = _oracle_tree(x, y, 8, f.bbox)
    H.compare_canonical(H.gpu_canonical(f), ora.canonical())


def test_tree_duplicates_below_cutoff():
    x, y, _ = H.uniform(5000, 7)
    x = np.concatenate([x, np.full(10, x[0])])
    y = np.concatenate([y, np.full(10, y[0])])
    f = _build(x, y, 32, bbox=None)
    ora = _oracle_tree(x, y, 32, f.bbox)
    H.compare_canonical(H.gpu_canonical(f), ora.canonical())


def test_tree_isolated_points_finish_alone_in_a_sort_round():
    """Points alone in their depth-4 cell while every other depth-4 cell holds
    more than Ncutoff: the first radix round of the key sort (depths 1-4)
    finishes one-element segments only.  The sorted permutation still holds
    every point once, and the tree equals the oracle's."""
    rng = np.random.default_rng(31)
    n = 60000
    # the central block [0.25, 0.75]^2 is exactly 8 x 8 depth-4 cells
    x = np.concatenate([0.25 + 0.5 * rng.random(n), [0.01, 0.97, 0.02]])
    y = np.concatenate([0.25 + 0.5 * rng.random(n), [0.02, 0.015, 0.99]])
    f = _build(x, y, 64, bbox=[0, 1, 0, 1])
    perm = f.tree_array("perm").cpu().numpy().astype(np.int64)
    assert np.array_equal(np.sort(perm), np.arange(x.size))
    ora = _oracle_tree(x, y, 64, f.bbox)
    assert H.compare_trees_fast(f, ora) == sum(l.n_node for l in ora.Levels)


def test_tree_uniform_counts_full():
    """A full uniform tree: levels 0..L-1 hold 4^d nodes and the M2L lists hold
    the closed-form number of pairs."""
    x, y, _ = H.uniform(262144, 21)
    # 256 points per level-5 box on average (all split), 64 per level-6 leaf
    f = _build(x, y, 150, bbox=[0, 1, 0, 1])
    t = f.tree
    assert t.level_nodes == [4 ** d for d in range(t.levels)]
    m2l = f.tree_array("m2l").cpu().numpy().reshape(-1, 40)
    # per level: sum over boxes of the ring partners that lie inside the domain
    for d in range(2, t.levels):
        n = 2 ** d
        g0 = t.node_off[d]
        cnt = int((m2l[g0:g0 + 4 ** d] >= 0).sum())
        ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        expect = 0
        for ox in range(-3, 4):
            for oy in range(-3, 4):
                if abs(ox) <= 1 and abs(oy) <= 1:
                    continue
                cx, cy = ix & 1, iy & 1
                ok = (ox >= -2 - cx) & (ox <= 3 - cx) & (oy >= -2 - cy) & (oy <= 3 - cy)
                ok &= (ix + ox >= 0) & (ix + ox < n) & (iy + oy >= 0) & (iy + oy < n)
                expect += int(ok.sum())
        assert cnt == expect, (d, cnt, expect)


# complete trees (uniform points): depth D from the depth-(D+1) histogram
# (D = Dh - 1) or from the depth-D one (D = Dh); and cases the closed-form
# build must refuse (a cell of depth D-1 at or under Ncutoff, an empty leaf,
# clustered points)
FULL_CASES = {
    "uniform_1M_ncut100": lambda: (*H.uniform(1_000_000, 2)[:2], 100, [0, 1, 0, 1]),
    "uniform_100k_ncut64": lambda: (*H.uniform(100_000, 1)[:2], 64, None),
    "uniform_300k_ncut40": lambda: (*H.uniform(300_000, 12)[:2], 40, [0, 1, 0, 1]),
    "uniform_50k_ncut200": lambda: (*H.uniform(50_000, 13)[:2], 200, [0, 1, 0, 1]),
    "grid_65536_ncut16": lambda: (*[a.copy() for a in H.grid(256)], 16, [0, 1, 0, 1]),
    "uniform_with_hole": lambda: (*_with_hole(200_000, 14), 64, [0, 1, 0, 1]),
    "clustered_100k": lambda: (*H.clustered(100_000, 10, 0.01, 15)[:2], 64, [0, 1, 0, 1]),
}


def _with_hole(n, seed):
    x, y, _ = H.uniform(n, seed)
    keep = ~((np.abs(x - 0.5) < 0.02) & (np.abs(y - 0.5) < 0.02))
    return x[keep].copy(), y[keep].copy()


@pytest.mark.parametrize("case", sorted(FULL_CASES))
def test_complete_tree_build_equals_general_build(case):
    """The closed-form build of complete trees (FMM2D_TREE_AUTO) gives the
    general build's tree array for array, bit for bit (keys, permutation,
    sorted points, every node array), and both equal the oracle's tree."""
    import torch
    from flexmm_amd import _lib
    x, y, ncut, bbox = FULL_CASES[case]()
    print("complete tree", case, x.size, flush=True)
    f = _build(x, y, ncut, bbox=bbox)
    names = ["keys", "perm", "sx", "sy", "code", "level", "bot", "top", "parent", "child",
             "flags", "colleagues", "xmid", "ymid", "row", "m2l", "evalid"]
    auto = {k: f.tree_array(k).cpu() for k in names}
    info_auto = (f.tree.levels, f.tree.n_nodes, f.tree.n_rows, f.tree.n_leaves, f.tree.n_s2m)
    _lib.check(f.lib.fmm2d_set_tree_mode(f.handle, _lib.TREE_GENERAL))
    try:
        f.build_tree()
        gen = {k: f.tree_array(k).cpu() for k in names}
        info_gen = (f.tree.levels, f.tree.n_nodes, f.tree.n_rows, f.tree.n_leaves, f.tree.n_s2m)
    finally:
        _lib.check(f.lib.fmm2d_set_tree_mode(f.handle, _lib.TREE_AUTO))
    assert info_auto == info_gen, (info_auto, info_gen)
    for k in names[4:]:
        a, b = auto[k], gen[k]
        if k == "m2l":   # levels 0 and 1 carry no lists
            start = f.tree.node_off[2] * 40 if f.tree.levels > 2 else a.numel()
            a, b = a[start:], b[start:]
        assert torch.equal(a, b), "%s: array %s differs" % (case, k)
    # points: the same set in every leaf (the general build's order inside a
    # leaf follows its radix scatter's atomics), each with its own key and
    # coordinates
    bot, top = auto["bot"].numpy(), auto["top"].numpy()
    leaf = (auto["flags"].numpy() & H.NF_LEAF) != 0
    owner = H._point_owner(np.arange(x.size), bot[leaf], top[leaf], x.size)
    assert (owner >= 0).all()
    for d in (auto, gen):
        perm = d["perm"].numpy().astype(np.int64)
        o = np.lexsort((perm, owner))
        d["order"] = o
        assert np.array_equal(d["sx"].numpy(), x[perm]) and np.array_equal(d["sy"].numpy(), y[perm])
    for k in ("perm", "keys"):
        assert np.array_equal(auto[k].numpy()[auto["order"]], gen[k].numpy()[gen["order"]]), k
    f.build_tree()
    n = H.compare_trees_fast(f, _oracle_tree(x, y, ncut, f.bbox))
    print(case, "levels", f.tree.levels, "nodes", n)


def test_tree_sequence_on_one_plan():
    """Consecutive builds on one plan's tree: a complete tree of the previous
    build's depth is launched before the host reads the split counts (its
    kernels gated on the device by those counts); a build of another depth, an
    adaptive build and a build after it take their own paths.  Every tree
    equals the oracle's."""
    seq = [
        ("uniform_100k_ncut64", *H.uniform(100_000, 1)[:2], 64, None),
        ("uniform_100k_ncut64 again", *H.uniform(100_000, 1)[:2], 64, None),
        ("uniform_300k_ncut40", *H.uniform(300_000, 12)[:2], 40, [0, 1, 0, 1]),
        ("clustered_30k", *H.clustered(30000, 5, 0.01, 3)[:2], 50, [0, 1, 0, 1]),
        ("uniform_100k_ncut64 after", *H.uniform(100_000, 1)[:2], 64, None),
        ("uniform_50k_ncut200", *H.uniform(50_000, 13)[:2], 200, [0, 1, 0, 1]),
    ]
    plans = set()
    for name, x, y, ncut, bbox in seq:
        print("build", name, flush=True)
        f = _build(x, y, ncut, bbox=bbox)
        plans.add(f.handle.value)
        ora = _oracle_tree(x, y, ncut, f.bbox)
        n = H.compare_canonical(H.gpu_canonical(f), ora.canonical())
        assert n == sum(l.n_node for l in ora.Levels), name
        f.close()
    assert len(plans) == 1   # every build reused the one pooled plan
