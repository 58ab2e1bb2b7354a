"""Host check of the store Gram's work split (no GPU): a mirror of the
device mapping item_waves (rocket_amd/csrc/store.hip) must cover every
128 x 64 wave block that meets the upper triangle of `count` patterns
exactly once (every 16 x 16 sub-tile that meets it, at least once), in
ceil(Q (Q + 1) / 8) items per slab for Q = ceil(count / 128) quarters of
128 patterns (4 quarter-pair units per item, the least possible), each
wave's A and B rows within the quarters its item loads into LDS."""
import numpy as np
import pytest


def tri_index(i, j, nt):
    return i * nt - i * (i - 1) // 2 + (j - i)


def item_waves(u, count, wave):
    """(q[4], nq, a_off, b_off, out tile (I, J) or None, block row/col offset)."""
    Q = (count + 127) >> 7
    m = Q >> 1
    wr, wc = wave >> 2, wave & 3
    toff = m * (m - 1) // 2
    tiles = lambda x, y: [256 * x, 256 * x + 128, 256 * y, 256 * y + 128]   # noqa: E731
    if u < toff:
        pairs = [(i, j) for i in range(m) for j in range(i + 1, m)]
        i, j = pairs[u]
        return tiles(i, j), 4, wr * 128, 256 + wc * 64, (i, j), (wr * 128, wc * 64)
    p = u - toff
    if Q % 2 == 0:
        d0 = 8 * p // 6
        dl = min(m - 1, (8 * p + 7) // 6)
        q = tiles(d0, d0 + 1 if dl > d0 else d0)
        nq = 4 if dl > d0 else 2
        e = 8 * p + wave
        if e >= 6 * m:
            return q, nq, 0, 0, None, None
        d, l = divmod(e, 6)
        bwr = 0 if l < 4 else 1
        bwc = l if l < 4 else l - 2
        base = (d - d0) * 256
        return q, nq, base + bwr * 128, base + bwc * 64, (d, d), (bwr * 128, bwc * 64)
    k, h = wave >> 1, wave & 1
    if p < m:
        d = p
        q = [256 * d, 256 * d + 128, 256 * m, 256 * m]
        sa = 1 if k == 3 else 0
        sb = [0, 1, 2, 2][k]
        tile = (d, d if k < 2 else m)
        return q, 3, sa * 128, sb * 128 + h * 64, tile, (sa * 128, (128 if k == 1 else 0) + h * 64)
    j0 = 4 * (p - m)
    q = [256 * min(j0 + i, m) + (128 if j0 + i < m else 0) for i in range(4)]
    nq = min(4, m + 1 - j0)
    j = j0 + k
    if j > m:
        return q, nq, 0, 0, None, None
    off = (128, 128 + h * 64) if j < m else (0, h * 64)
    return q, nq, k * 128, k * 128 + h * 64, (j, j), off


def items_per_slab(count):
    Q = (count + 127) >> 7
    m = Q >> 1
    return m * (m - 1) // 2 + (m + (m + 4) // 4 if Q % 2 else (6 * m + 7) // 8)


@pytest.mark.parametrize("count", [1, 100, 128, 129, 256, 300, 384, 596, 640, 1000, 1024, 1100, 1500, 2000,
                                   2100, 2200, 2400, 3000, 4500])
def test_items_cover_the_triangle_once(count):
    Q = (count + 127) >> 7
    nt = (count + 255) // 256
    items = items_per_slab(count)
    assert items == -(-Q * (Q + 1) // 8)           # 4 units per item, the least possible
    blocks = np.zeros((nt * 2, nt * 4), dtype=np.int32)      # 128 x 64 wave blocks of the grid
    for u in range(items):
        for wave in range(8):
            q, nq, a_off, b_off, tile, off = item_waves(u, count, wave)
            if tile is None:
                continue
            I, J = tile
            assert a_off % 128 == 0 and b_off % 64 == 0
            qa, qb = a_off // 128, b_off // 128
            assert qa < nq and qb < nq                           # rows the item loads (nq quarters)
            # the LDS quarters hold the pattern rows of the block written
            assert q[qa] == 256 * I + off[0]
            assert q[qb] + b_off % 128 == 256 * J + off[1]
            if I == J:
                assert off[0] <= off[1] + 63                      # meets the upper triangle
            bi, bj = I * 2 + off[0] // 128, J * 4 + off[1] // 64
            blocks[bi, bj] += 1
    bi, bj = np.indices(blocks.shape)
    # a block is needed when it holds a pattern pair i <= j, i, j < count
    want = ((bi * 128 <= bj * 64 + 63) & (bi * 128 < count) & (bj * 64 < count)).astype(np.int32)
    need_cols = (np.arange(nt * 4) * 64 < count)
    # blocks of a needed quarter pair come in twos (the 128 columns of quarter qb): the second half of
    # the last quarter may lie past count and is computed anyway
    want_pairs = want.copy()
    for r in range(want.shape[0]):
        for c in range(0, want.shape[1], 2):
            if want[r, c] or want[r, c + 1]:
                want_pairs[r, c] = want_pairs[r, c + 1] = 1
    assert np.all(blocks <= 1)
    assert np.array_equal(blocks * need_cols[None, :], want)
    assert np.array_equal(blocks, want_pairs)


def test_odd_quarters_save_items():
    # 596 patterns: 5 quarters; the tile-pair packing took 6 items per slab, the quarter packing 4
    assert items_per_slab(596) == 4
    assert items_per_slab(1024) == 9 and items_per_slab(2000) == 34 and items_per_slab(1500) == 20
