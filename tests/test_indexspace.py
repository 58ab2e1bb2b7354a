"""Host-side pair scheduling (Rocket's index spaces) against brute force and
the oracle's restatement of Util.calculateTriangleSize."""
import numpy as np
import pytest

from oracle import prnu_oracle as O
from rocket_amd import indexspace as ix


@pytest.mark.parametrize("rb,re,cb,ce", [(0, 10, 0, 10), (0, 4, 6, 10), (3, 9, 2, 7), (5, 5, 0, 3), (0, 1000, 0, 1000)])
@pytest.mark.parametrize("diag", [False, True])
def test_triangle_size_matches_oracle(rb, re, cb, ce, diag):
    assert ix.calculate_triangle_size(rb, re, cb, ce, diag) == O.triangle_size_range(rb, re, cb, ce, diag)


def brute(n, diag):
    return {(i, j) for i in range(n) for j in range(n) if i < j or (diag and i == j)}


@pytest.mark.parametrize("kind,param", [("hilbert", 1), ("hilbert", 64), ("tiling", 1), ("tiling", 7), ("tiling", 50)])
@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 64, 101])
@pytest.mark.parametrize("diag", [False, True])
def test_all_pairs_exactly_once(kind, param, n, diag):
    pairs = ix.all_pairs(n, kind, param, diag)
    got = [tuple(p) for p in pairs]
    assert len(got) == len(set(got)) == O.triangle_size_range(0, n, 0, n, diag)
    assert set(got) == brute(n, diag)


def test_hilbert_order_is_local():
    # on a power-of-two square the curve moves one cell at a time except where
    # it skips over the lower triangle
    pairs = ix.all_pairs(64, "hilbert", 1, True)
    steps = np.abs(np.diff(pairs, axis=0)).sum(axis=1)
    assert np.mean(steps == 1) > 0.95


def test_divide_recursion_covers_space():
    keys = [f"img{k}" for k in range(37)]
    root = ix.HilbertCurveIndexSpace(keys, min_split_size=16, include_diagonal=False)
    assert root.size() == 37 * 36 // 2
    seen, stack, steps = [], [root], 0
    while stack:
        sp = stack.pop()
        d = ix.IndexSpaceDivision()
        sp.divide(d)
        assert sum(c.size() for c in d.get_children()) + len(d.get_entries()) == sp.size()
        stack.extend(d.get_children())
        seen.extend(d.get_entries())
        steps += 1
    assert steps > 1
    assert len(seen) == len(set(seen)) == 37 * 36 // 2
    assert all(keys.index(a) < keys.index(b) for a, b in seen)


def test_tiling_divide_tiles_then_entries():
    keys = list(range(10))
    root = ix.TilingIndexSpace(4, keys, include_diagonal=False)
    d = ix.IndexSpaceDivision()
    root.divide(d)
    assert not d.get_entries()
    # 3 x 3 tiles of the 10 x 10 matrix, minus the 3 below the diagonal
    assert len(d.get_children()) == 6
    total = 0
    for c in d.get_children():
        dd = ix.IndexSpaceDivision()
        c.divide(dd)
        assert not dd.get_children()
        total += len(dd.get_entries())
    assert total == 45


@pytest.mark.parametrize("count,tile", [(1024, 256), (1000, 256), (700, 256), (5000, 256), (3, 2)])
def test_tile_blocks_cover_the_triangle_once(count, tile):
    """Index space over tiles: every pair i < j lies in exactly one block."""
    blocks = ix.tile_blocks(count, tile)
    cov = np.zeros((count, count), dtype=np.int32)
    for r0, rn, c0, cn in blocks:
        assert 0 < rn <= tile and 0 < cn <= tile
        if (r0, rn) == (c0, cn):
            for i in range(rn):
                cov[r0 + i, r0 + i + 1:r0 + rn] += 1
        else:
            assert c0 >= r0 + rn                       # strictly above the diagonal
            cov[r0:r0 + rn, c0:c0 + cn] += 1
    iu = np.triu_indices(count, 1)
    assert (cov[iu] == 1).all() and cov.sum() == len(iu[0])


# Golden curve orders, derived by hand from the rule table of
# HilbertCurveIndexSpace.java:10-20 ({p, q, next rule} per quadrant, p the
# left (row) half, q the right (column) half) and the recursion of
# launchSplits (:51-98): rule A visits quadrants (0,0)->D, (0,1)->A,
# (1,1)->A, (1,0)->B; B: (1,1)->C, (0,1)->B, (0,0)->B, (1,0)->A;
# C: (1,1)->B, (1,0)->C, (0,0)->C, (0,1)->D; D: (0,0)->A, (1,0)->D,
# (1,1)->D, (0,1)->C.  Quadrants (and cells) with no pair i < j (i <= j with
# the diagonal) are skipped.
HILBERT_8_NO_DIAG = [
    # rows 0-3 x cols 0-3, rule D: (0-1 x 0-1, A), (2-3 x 2-3, D), (0-1 x 2-3, C)
    (0, 1), (2, 3), (1, 3), (1, 2), (0, 2), (0, 3),
    # rows 0-3 x cols 4-7, rule A: (0-1 x 4-5, D), (0-1 x 6-7, A), (2-3 x 6-7, A), (2-3 x 4-5, B)
    (0, 4), (1, 4), (1, 5), (0, 5), (0, 6), (0, 7), (1, 7), (1, 6),
    (2, 6), (2, 7), (3, 7), (3, 6), (3, 5), (2, 5), (2, 4), (3, 4),
    # rows 4-7 x cols 4-7, rule A: (4-5 x 4-5, D), (4-5 x 6-7, A), (6-7 x 6-7, A)
    (4, 5), (4, 6), (4, 7), (5, 7), (5, 6), (6, 7),
]
HILBERT_4_DIAG = [(0, 0), (1, 1), (0, 1), (0, 2), (0, 3), (1, 3), (1, 2), (2, 2), (2, 3), (3, 3)]


def test_hilbert_golden_order():
    assert [tuple(p) for p in ix.all_pairs(8, "hilbert", 1, False)] == HILBERT_8_NO_DIAG
    assert [tuple(p) for p in ix.all_pairs(4, "hilbert", 1, True)] == HILBERT_4_DIAG
