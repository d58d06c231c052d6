"""Shared test helpers: seeded workloads and the canonical tree view."""
import numpy as np


def uniform(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random(n)
    y = rng.random(n)
    tau = rng.standard_normal(n)
    return x, y, tau


def clustered(n, nclusters, sigma, seed, lo=0.1, hi=0.9):
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, (nclusters, 2))
    per = n // nclusters
    pts = c[np.repeat(np.arange(nclusters), per)] + sigma * rng.standard_normal((per * nclusters, 2))
    pts = np.clip(pts, 0.0, 1.0)
    tau = rng.uniform(-1, 1, per * nclusters)
    return pts[:, 0].copy(), pts[:, 1].copy(), tau


def grid(m):
    g = np.linspace(0, 1, m)
    tx, ty = np.meshgrid(g, g, indexing="ij")
    return tx.ravel(), ty.ravel()


def compact_bits(v):
    v = v.astype(np.uint64) & np.uint64(0x5555555555555555)
    for sh, mask in ((1, 0x3333333333333333), (2, 0x0F0F0F0F0F0F0F0F), (4, 0x00FF00FF00FF00FF),
                     (8, 0x0000FFFF0000FFFF), (16, 0x00000000FFFFFFFF)):
        v = (v | (v >> np.uint64(sh))) & np.uint64(mask)
    return v.astype(np.int64)


def demorton(code):
    code = code.astype(np.uint64)
    return compact_bits(code >> np.uint64(1)), compact_bits(code)


NF_LEAF, NF_FAKE, NF_XLIST, NF_UP, NF_DOWN = 1, 2, 4, 8, 16


def m2l_slots():
    out = []
    for ii in range(7):
        for jj in range(7):
            if 2 <= ii <= 4 and 2 <= jj <= 4:
                continue
            out.append((ii - 3, jj - 3))
    return out


def gpu_canonical(fmm):
    """Canonical per-level view of a flexmm_amd FMM's device tree, in the
    format of oracle.tree.Tree.canonical()."""
    t = fmm.tree
    get = lambda n: fmm.tree_array(n).cpu().numpy()
    code = get("code").view(np.uint64)
    level = get("level")
    flags = get("flags")
    bot = get("bot")
    top = get("top")
    perm = get("perm").astype(np.int64)
    coll = get("colleagues").reshape(-1, 9)
    row = get("row")
    m2l = get("m2l").reshape(-1, 40)
    xmid = get("xmid")
    ymid = get("ymid")
    ix, iy = demorton(code)
    rownode = {int(r): g for g, r in enumerate(row) if r >= 0}
    slots = m2l_slots()
    out = []
    for d in range(t.levels):
        dd = {}
        g0 = t.node_off[d]
        for g in range(g0, g0 + t.level_nodes[d]):
            f = int(flags[g])
            col = coll[g]
            rec = dict(
                leaf=bool(f & NF_LEAF), fake=bool(f & NF_FAKE), xlist=bool(f & NF_XLIST),
                compute_upwards=bool(f & NF_UP), compute_downwards=bool(f & NF_DOWN),
                points=tuple(np.sort(perm[bot[g]:top[g]]).tolist()),
                colleagues=frozenset((int(ix[c]), int(iy[c])) for c in col if c >= 0),
                has_density=bool(row[g] >= 0), xmid=float(xmid[g]), ymid=float(ymid[g]),
            )
            assert level[g] == d
            if d >= 2 and not rec["fake"]:
                parts = set()
                for s, r in enumerate(m2l[g]):
                    if r >= 0:
                        j = rownode[int(r)]
                        parts.add((slots[s][0], slots[s][1], int(ix[j]), int(iy[j])))
                rec["m2l"] = frozenset(parts)
            dd[(int(ix[g]), int(iy[g]))] = rec
        out.append(dd)
    return out


def compare_canonical(gpu, ora):
    """Assert two canonical trees are identical; returns the node count."""
    assert len(gpu) == len(ora), "levels %d vs %d" % (len(gpu), len(ora))
    n = 0
    for d, (a, b) in enumerate(zip(gpu, ora)):
        assert set(a) == set(b), "level %d node sets differ: %d vs %d" % (d, len(a), len(b))
        for key in b:
            ra, rb = a[key], b[key]
            for f in ("leaf", "fake", "xlist", "compute_upwards", "compute_downwards",
                      "points", "colleagues", "xmid", "ymid"):
                assert ra[f] == rb[f], "level %d node %s field %s: %r vs %r" % (d, key, f, ra[f], rb[f])
            if d >= 2:
                assert ra["has_density"] == rb["has_density"], (d, key)
            if "m2l" in rb:
                assert ra["m2l"] == rb["m2l"], "level %d node %s m2l differs" % (d, key)
            n += 1
    return n


def rel_max_err(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b))))
