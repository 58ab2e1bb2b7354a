"""Host check of the store Gram's work split (no GPU): a mirror of the
device mapping item_waves (rocket_amd/csrc/store.hip) must cover every
128 x 64 wave block that meets the upper triangle of the tile grid exactly
once (every 16 x 16 sub-tile that meets it, at least once), with
nt (nt - 1) / 2 off-diagonal tiles and ceil(6 nt / 8) diagonal packs per
slab, each pack's rows within the two tiles its LDS halves hold."""
import numpy as np
import pytest


def item_waves(u, nt, wave):
    """(rows_x, rows_y, a_off, b_off, out tile (I, J) or None, block row/col offset)."""
    wr, wc = wave >> 2, wave & 3
    toff = nt * (nt - 1) // 2
    if u < toff:
        pairs = [(i, j) for i in range(nt) for j in range(i + 1, nt)]
        i, j = pairs[u]
        return i, j, wr * 128, 256 + wc * 64, (i, j), (wr * 128, wc * 64)
    p = u - toff
    d0 = 8 * p // 6
    dl = min(nt - 1, (8 * p + 7) // 6)
    rows_y = d0 + 1 if dl > d0 else -1
    e = 8 * p + wave
    if e >= 6 * nt:
        return d0, rows_y, 0, 0, None, None
    d, l = divmod(e, 6)
    bwr = 0 if l < 4 else 1
    bwc = l if l < 4 else l - 2
    base = (d - d0) * 256
    return d0, rows_y, base + bwr * 128, base + bwc * 64, (d, d), (bwr * 128, bwc * 64)


@pytest.mark.parametrize("nt", [1, 2, 3, 4, 5, 7, 8, 9])
def test_diagonal_packs_cover_the_triangle_once(nt):
    items = nt * (nt - 1) // 2 + (6 * nt + 7) // 8
    blocks = np.zeros((nt * 2, nt * 4), dtype=np.int32)      # 128 x 64 wave blocks of the grid
    for u in range(items):
        for wave in range(8):
            rows_x, rows_y, a_off, b_off, tile, off = item_waves(u, nt, wave)
            if tile is None:
                continue
            I, J = tile
            lds_tile = lambda row: rows_x if row < 256 else rows_y     # noqa: E731
            assert lds_tile(a_off) == I and lds_tile(b_off) == J    # rows in LDS are the tiles written
            blocks[I * 2 + off[0] // 128, J * 4 + off[1] // 64] += 1
    bi, bj = np.indices(blocks.shape)
    # a block meets the upper triangle when its first row is at most its last column
    want = (bi * 128 <= bj * 64 + 63).astype(np.int32)
    assert np.array_equal(blocks, want)
