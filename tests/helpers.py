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


# ---------------------------------------------------------------------------
# Full-size trees: the canonical comparison in vectorised form (the
# dictionary view above walks every node's point list in Python).
def _pairs_equal(a, b):
    a = np.unique(np.asarray(a, dtype=np.int64).reshape(-1, np.shape(a)[-1]), axis=0)
    b = np.unique(np.asarray(b, dtype=np.int64).reshape(-1, np.shape(b)[-1]), axis=0)
    return a.shape == b.shape and bool((a == b).all())


def _point_owner(perm, bot, top, npts):
    """owner[p] = index of the node whose range [bot, top) holds point p (-1: none)."""
    owner = -np.ones(npts, dtype=np.int64)
    lens = (top - bot).astype(np.int64)
    if lens.sum() == 0:
        return owner
    starts = np.repeat(bot.astype(np.int64) - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens)
    pos = starts + np.arange(int(lens.sum()))
    owner[perm[pos]] = np.repeat(np.arange(len(bot)), lens)
    return owner


def compare_trees_fast(fmm, otree):
    """Assert that a flexmm_amd FMM's device tree equals an oracle.tree.Tree
    canonically: per level, the node set keyed by (ix, iy), every node's flags,
    midpoints and point set, its colleagues and its M2L partners.  Returns the
    number of nodes compared."""
    t = fmm.tree
    get = lambda n: fmm.tree_array(n).cpu().numpy()
    code = get("code").view(np.uint64)
    flags = get("flags")
    bot, top = get("bot").astype(np.int64), get("top").astype(np.int64)
    perm = get("perm").astype(np.int64)
    coll = get("colleagues").reshape(-1, 9)
    row = get("row")
    m2l = get("m2l").reshape(-1, 40)
    xmid, ymid = get("xmid"), get("ymid")
    gix, giy = demorton(code)
    gkey = (gix << 32) | giy
    rownode = -np.ones(int(row.max()) + 2 if row.size else 1, dtype=np.int64)
    rownode[row[row >= 0]] = np.flatnonzero(row >= 0)
    slots = np.array(m2l_slots(), dtype=np.int64)
    npts = perm.size
    assert t.levels == len(otree.Levels), "levels %d vs %d" % (t.levels, len(otree.Levels))
    total = 0
    for d, lev in enumerate(otree.Levels):
        g = np.arange(t.node_off[d], t.node_off[d] + t.level_nodes[d])
        ok = lev.ix.astype(np.int64) << 32 | lev.iy.astype(np.int64)
        gs, os_ = g[np.argsort(gkey[g], kind="stable")], np.argsort(ok, kind="stable")
        assert np.array_equal(gkey[gs], ok[os_]), "level %d node sets differ" % d
        f = flags[gs]
        for bit, name in ((NF_LEAF, "leaf"), (NF_FAKE, "fake_leaf"), (NF_XLIST, "Xlist"),
                          (NF_UP, "compute_upwards"), (NF_DOWN, "compute_downwards")):
            assert np.array_equal((f & bit) != 0, getattr(lev, name)[os_]), (d, name)
        assert np.array_equal(xmid[gs], lev.xmid[os_]) and np.array_equal(ymid[gs], lev.ymid[os_]), d
        if d >= 2:
            assert np.array_equal(row[gs] >= 0, lev.this_density_ind[os_] >= 0), (d, "density")
        # point sets: the sorted position of the node that holds each point
        assert np.array_equal(_point_owner(perm, bot[gs], top[gs], npts),
                              _point_owner(otree.ordv, lev.bot_ind[os_], lev.top_ind[os_], npts)), \
            "level %d point sets differ" % d
        # colleagues as (node key, colleague key)
        gc = coll[gs]
        gi = np.repeat(gkey[gs][:, None], 9, axis=1)
        gp = np.stack([gi[gc >= 0], gkey[gc[gc >= 0]]], axis=1)
        oc = lev.colleagues[os_]
        oi = np.repeat(ok[os_][:, None], oc.shape[1], axis=1)
        op = np.stack([oi[oc >= 0], ok[oc[oc >= 0]]], axis=1)
        assert _pairs_equal(gp, op), "level %d colleagues differ" % d
        # M2L partners as (node key, dx, dy, partner key), non-fake nodes
        if d >= 2:
            real = (f & NF_FAKE) == 0
            gm = m2l[gs][real]
            kk = np.repeat(gkey[gs][real][:, None], 40, axis=1)
            sl = np.broadcast_to(np.arange(40)[None, :], gm.shape)
            sel = gm >= 0
            gpart = gkey[rownode[gm[sel]]]
            gq = np.stack([kk[sel], slots[sl[sel], 0], slots[sl[sel], 1], gpart], axis=1)
            dil = otree.interaction_lists[d]                          # (7, 7, n)
            orow = {int(r): j for j, r in enumerate(lev.this_density_ind) if r >= 0}
            omap = -np.ones(int(lev.this_density_ind.max()) + 2, dtype=np.int64)
            for r, j in orow.items():
                omap[r] = j
            a, b, i = np.nonzero(dil >= 0)
            keep = ~lev.fake_leaf[i]
            a, b, i = a[keep], b[keep], i[keep]
            oq = np.stack([ok[i], a - 3, b - 3, ok[omap[dil[a, b, i]]]], axis=1)
            assert _pairs_equal(gq, oq), "level %d M2L lists differ" % d
        total += g.size
    return total


# ---------------------------------------------------------------------------
# FP64 direct sums on the GPU with torch, for sampled targets of full-size
# workloads (the NumPy sums of oracle/direct.py, against which each test pins
# these on a subset, take minutes there).
def torch_direct_laplace(torch, sx, sy, tau, tx, ty, check=False, chunk=16):
    """u(t) = sum_s -log(|t - s|^2)/(4 pi) tau_s."""
    out = []
    for s in range(0, tx.numel(), chunk):
        dx = tx[s:s + chunk, None] - sx[None, :]
        dy = ty[s:s + chunk, None] - sy[None, :]
        d2 = dx * dx + dy * dy
        v = -torch.log(torch.where(d2 > 0, d2, 1.0)) * tau[None, :]
        if check:
            v = torch.where(d2 > 0, v, 0.0)
        out.append(v.sum(dim=1))
    return torch.cat(out) / (4 * np.pi)


def torch_direct_stokes(torch, sx, sy, fx, fy, tx, ty, check=False, chunk=16):
    """(u, v) of the 2D Stokeslet, (1/4 pi) sum (-log r I + r r^T / r^2) f."""
    ru, rv = [], []
    for s in range(0, tx.numel(), chunk):
        dx = tx[s:s + chunk, None] - sx[None, :]
        dy = ty[s:s + chunk, None] - sy[None, :]
        d2 = dx * dx + dy * dy
        keep = d2 > 0
        id2 = torch.where(keep, 1.0 / torch.where(keep, d2, 1.0), 0.0)
        lg = -0.5 * torch.log(torch.where(keep, d2, 1.0))
        if not check:
            lg = torch.where(keep, lg, float("inf"))
        ru.append(((lg + dx * dx * id2) * fx[None, :] + dx * dy * id2 * fy[None, :]).sum(1))
        rv.append((dx * dy * id2 * fx[None, :] + (lg + dy * dy * id2) * fy[None, :]).sum(1))
    return torch.cat(ru) / (4 * np.pi), torch.cat(rv) / (4 * np.pi)


def torch_direct_mh_adjoint(torch, k, sx, sy, tau, tx, ty, g, chunk=8):
    """For L = sum_t g_t^2/2-type adjoints: with weights g at the targets,
    return (sum_t g_t K0(k r)/(2 pi), tau_s sum_t g_t dK/dx_s, tau_s sum_t g_t
    dK/dy_s) at the sources s, r = |s - t|."""
    a, b, c = [], [], []
    for s in range(0, sx.numel(), chunk):
        dx = sx[s:s + chunk, None] - tx[None, :]
        dy = sy[s:s + chunk, None] - ty[None, :]
        r = torch.sqrt(dx * dx + dy * dy)
        k0 = torch.special.modified_bessel_k0(k * r)
        k1 = torch.special.modified_bessel_k1(k * r)
        a.append((k0 * g[None, :]).sum(1))
        w = -k * k1 / r * g[None, :]
        b.append((w * dx).sum(1))
        c.append((w * dy).sum(1))
    sc = 1.0 / (2 * np.pi)
    return torch.cat(a) * sc, tau * torch.cat(b) * sc, tau * torch.cat(c) * sc


def morton(ix, iy):
    """Interleave (ix, iy) as the device's node code: x bits odd, y bits even."""
    def spread(v):
        v = v.astype(np.uint64) & np.uint64(0x00000000FFFFFFFF)
        for sh, mask in ((16, 0x0000FFFF0000FFFF), (8, 0x00FF00FF00FF00FF),
                         (4, 0x0F0F0F0F0F0F0F0F), (2, 0x3333333333333333),
                         (1, 0x5555555555555555)):
            v = (v | (v << np.uint64(sh))) & np.uint64(mask)
        return v
    return (spread(ix) << np.uint64(1)) | spread(iy)


class OracleAsDevice:
    """An oracle.tree.Tree laid out as the device tree arrays (global node
    index over all levels), to check the comparisons themselves on the CPU."""

    def __init__(self, otree):
        L = len(otree.Levels)
        sizes = [lev.n_node for lev in otree.Levels]
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        roff = np.concatenate([[0], np.cumsum([lev.n_allocate_density for lev in otree.Levels])])
        slots = m2l_slots()

        class _T:
            pass
        self.tree = _T()
        self.tree.levels = L
        self.tree.node_off = list(off[:-1])
        self.tree.level_nodes = sizes
        a = {}
        cat = lambda f: np.concatenate([f(d, lev) for d, lev in enumerate(otree.Levels)])
        a["code"] = cat(lambda d, lev: morton(lev.ix, lev.iy)).view(np.int64)
        a["flags"] = cat(lambda d, lev: (lev.leaf * NF_LEAF + lev.fake_leaf * NF_FAKE
                                          + lev.Xlist * NF_XLIST + lev.compute_upwards * NF_UP
                                          + lev.compute_downwards * NF_DOWN).astype(np.int32))
        a["bot"] = cat(lambda d, lev: lev.bot_ind).astype(np.int32)
        a["top"] = cat(lambda d, lev: lev.top_ind).astype(np.int32)
        a["perm"] = otree.ordv.astype(np.int32)
        a["xmid"] = cat(lambda d, lev: lev.xmid)
        a["ymid"] = cat(lambda d, lev: lev.ymid)

        def coll(d, lev):
            c = -np.ones((lev.n_node, 9), dtype=np.int64)
            cc = lev.colleagues
            c[:, :cc.shape[1]] = np.where(cc >= 0, cc + off[d], -1)
            return c
        a["colleagues"] = cat(coll).astype(np.int32).ravel()
        a["row"] = cat(lambda d, lev: np.where(lev.this_density_ind >= 0,
                                               lev.this_density_ind + roff[d], -1)).astype(np.int32)

        def m2l(d, lev):
            m = -np.ones((lev.n_node, 40), dtype=np.int64)
            if d >= 2:
                dil = otree.interaction_lists[d]
                for s, (dx, dy) in enumerate(slots):
                    r = dil[dx + 3, dy + 3]
                    m[:, s] = np.where(r >= 0, r + roff[d], -1)
            return m
        a["m2l"] = cat(m2l).astype(np.int32).ravel()
        self._a = a

    def tree_array(self, name):
        import torch
        return torch.from_numpy(self._a[name].copy())
