"""The product's host precomputation (flexmm_amd.kifmm2d.precompute) against
the oracle's restatement of helpers.py:250-317, and the packed device layout."""
import numpy as np
import pytest

from flexmm_amd import kernels
from flexmm_amd.kifmm2d import precompute as pc
from oracle import kernels as okern
from oracle import kifmm2d as okifmm


@pytest.mark.parametrize("p,w", [(32, 1.0), (60, 1.0 / 128), (64, 0.3)])
def test_operators_match_oracle(p, w):
    ops = pc.LevelOperators(kernels.laplace, p, w)
    ora = okifmm.Precomputation(w, okern.laplace, p)
    np.testing.assert_allclose(ops.M2MC, ora.M2MC, rtol=1e-15, atol=0)
    slots = pc.m2l_slots()
    assert len(slots) == 40
    for s, (ox, oy) in enumerate(slots):
        np.testing.assert_allclose(ops.m2l[s], ora.M2L[ox + 3, oy + 3], rtol=1e-15, atol=0)
    # the pseudo-inverse factors apply the same operator
    b = np.random.default_rng(p).standard_normal((p, 5))
    np.testing.assert_allclose(ops.pinv2 @ (ops.pinv1 @ b), ora.S2L_Solver(b), rtol=1e-9, atol=1e-9 * np.abs(ora.S2L_Solver(b)).max())
    np.testing.assert_allclose(ops.large[0], ora.large_x, rtol=0, atol=0)


def test_pack_layout():
    p = 20
    ops = pc.LevelOperators(kernels.laplace, p, 0.5)
    flat = ops.pack()
    PD = pc.padded_dim(p)
    assert PD == 32
    nb = pc.OP_NBLK * PD * PD
    blocks = flat[:nb].reshape(pc.OP_NBLK, PD, PD)
    circ = flat[nb:nb + 2 * PD].reshape(2, PD)
    means = flat[nb + 2 * PD:nb + 42 * PD].reshape(40, PD)
    r32 = nb + 42 * PD
    frag = flat[r32:r32 + 20 * PD * PD].view(np.float32).reshape(40, PD * PD)
    # MFMA operand order: entry ((rb MP + mp) 64 + lane) 4 + e holds row
    # 16 rb + lane % 16, column 8 m + 2 (lane // 16) + e % 2, m = 2 mp + e // 2
    MP = PD // 16
    resid = np.full((40, PD, PD), np.nan, dtype=np.float32)
    for rb in range(PD // 16):
        for mp in range(MP):
            for lane in range(64):
                for e in range(4):
                    m = 2 * mp + e // 2
                    resid[:, 16 * rb + lane % 16, 8 * m + 2 * (lane // 16) + e % 2] = \
                        frag[:, ((rb * MP + mp) * 64 + lane) * 4 + e]
    assert not np.isnan(resid).any()   # every entry stored once
    ranks = flat[r32 + 20 * PD * PD:r32 + 20 * PD * PD + 64]
    assert list(ranks[:40]) == ops.m2l_ranks and np.all(ranks[40:] == -1)
    for s in (0, 13, 39):
        m = ops.m2l[s]
        np.testing.assert_allclose(means[s, :p], m.mean(axis=1), rtol=1e-15)
        np.testing.assert_allclose(resid[s, :p, :p] + means[s, :p, None], m, rtol=1e-6)
        assert np.all(resid[s, p:, :] == 0) and np.all(resid[s, :, p:] == 0)
    np.testing.assert_array_equal(blocks[0, :p, :p], ops.pinv1)
    np.testing.assert_array_equal(blocks[1, :p, :p], ops.pinv2)
    for c in range(4):
        np.testing.assert_array_equal(blocks[2 + c, :p, :p], ops.M2MC[:, c * p:(c + 1) * p])
        np.testing.assert_array_equal(blocks[6 + c, :p, :p], ops.M2MC[:, c * p:(c + 1) * p].T)
    np.testing.assert_array_equal(blocks[10 + 7, :p, :p], ops.m2l[7])
    assert np.all(blocks[:, p:, :] == 0) and np.all(blocks[:, :, p:] == 0)
    np.testing.assert_array_equal(circ[0, :p], ops.large[0])
    assert np.all(circ[:, p:] == 0)
    assert flat.size == 50 * PD * PD + 2 * PD + 40 * PD + 20 * PD * PD + 64 + 40 * PD * PD


@pytest.mark.parametrize("p,w", [(60, 1.0), (60, 1.0 / 128), (64, 0.25), (24, 0.5)])
def test_compressed_m2l_matches_reference_rule(p, w):
    """The SVD-compressed M2L slots: ranks, kinds and products equal to the
    oracle's M2LEvaluator (helpers.py:17-45), and the packed block, read with
    the device kernel's two-stage index arithmetic (stage 1 rows of M1, stage 2
    block[PD/2 + i/2][(i%2) PD/2 + 4 j + kq]), applies M2 (M1 q)."""
    ops = pc.LevelOperators(kernels.laplace, p, w)
    ora = okifmm.Precomputation(w, okern.laplace, p)
    PD = pc.padded_dim(p)
    H = PD // 2
    flat = ops.pack()
    off = 50 * PD * PD + 2 * PD + 40 * PD + 20 * PD * PD
    ranks = flat[off:off + 64]
    svd = flat[off + 64:off + 64 + 40 * PD * PD].reshape(40, PD, PD)
    q = np.random.default_rng(p).standard_normal(p)
    ncomp = 0
    for s, (ox, oy) in enumerate(pc.m2l_slots()):
        ev = ora.M2LF[ox + 3, oy + 3]
        r = int(ranks[s])
        assert r == (ev.rank if ev.kind == "svd" else (0 if ev.kind == "null" else -1))
        if r <= 0:
            continue
        ncomp += 1
        blk = svd[s]
        T = np.zeros(PD)
        T[:H] = blk[:H, :p] @ q                        # stage 1 over the padded rows
        assert np.all(T[r:] == 0)
        out = np.zeros(PD)
        for j in range((r + 3) // 4):                   # stage 2
            for kq in range(4):
                k = 4 * j + kq
                for row in range(PD):
                    out[row] += blk[H + row // 2, (row % 2) * H + k] * T[k]
        ref = ev(q)
        np.testing.assert_allclose(out[:p], ref, rtol=0, atol=1e-13 * np.abs(ref).max())
        assert np.all(out[p:] == 0)
        np.testing.assert_allclose(out[:p], ev.m2l @ q, rtol=0,
                                   atol=1e-12 * np.abs(ev.m2l @ q).max())
    if p >= 60:
        assert ncomp >= 28


def test_padded_dim_limits():
    assert pc.padded_dim(4) == 32 and pc.padded_dim(33) == 64 and pc.padded_dim(64) == 64
    with pytest.raises(ValueError):
        pc.padded_dim(65)


@pytest.mark.parametrize("order,w", [(4, 1.0), (8, 1.0 / 64), (5, 0.3)])
def test_chebyshev_operators_match_oracle(order, w):
    """Black-box operators (flexmm_amd.flexmm2d.ChebOperators) against the
    oracle's BB_Precomputation restatement, and their packed layout: identity
    pseudo-inverse blocks and 2/w as the first circle entry."""
    from flexmm_amd.flexmm2d import bbfmm as bb
    from oracle import bbfmm as obb
    K = kernels.laplace
    ops = bb.ChebOperators(K, order, w)
    ora = obb.BBPrecomputation(w, okern.laplace, obb.BBInfo(order))
    n = order * order
    assert ops.p == n
    np.testing.assert_allclose(ops.M2MC, ora.M2MC, rtol=0, atol=1e-13)
    for s, (ox, oy) in enumerate(pc.m2l_slots()):
        ref = ora.M2L[ox + 3, oy + 3]
        np.testing.assert_allclose(ops.m2l[s], ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    flat = ops.pack()
    PD = pc.padded_dim(n)
    nb = pc.OP_NBLK * PD * PD
    blocks = flat[:nb].reshape(pc.OP_NBLK, PD, PD)
    np.testing.assert_array_equal(blocks[0, :n, :n], np.eye(n))
    np.testing.assert_array_equal(blocks[1, :n, :n], np.eye(n))
    np.testing.assert_array_equal(blocks[6, :n, :n], ops.M2MC[:, :n].T)
    circ = flat[nb:nb + 2 * PD].reshape(2, PD)
    assert circ[0, 0] == 2.0 / w == ora.width_adj
