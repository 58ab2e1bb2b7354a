"""GPU tests of the pattern spill tier (prnu_store_save / prnu_store_load,
rocket_amd/spill.py): slot images round-trip a store's slots and statistics
bit for bit; the block-streamed all-pairs NCC of a host tier (ragged last
block, files round trip included) matches the FP64 oracle (split-BF16:
|dNCC| <= 1e-5; FP16: its rounding bound) and equals the resident store's
Gram bit for bit when both use the same K slabs; images of another store
geometry, truncated images and truncated files are refused."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import prnu_oracle as O  # noqa: E402
from rocket_amd import _lib, csi, spill  # noqa: E402
from test_gpu_groups import fp16_bound  # noqa: E402

pytestmark = pytest.mark.gpu

VALS = {"split_bf16": 32, "fp32": 32, "fp16": 64, "fp8": 128}


def patterns(count, n, cams, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    k = torch.randn((cams, n), generator=g, device="cuda")
    p = torch.randn((count, n), generator=g, device="cuda")
    p += 0.05 * k[torch.arange(count, device="cuda") % cams]
    return p - p.mean(dim=1, keepdim=True)


def same_slabs(gram, fmt, n, S):
    """Fix the store Gram's K split at S slabs; return the equal slab length for plans."""
    nkb = (n + VALS[fmt] - 1) // VALS[fmt]
    cfg = (ctypes.c_int32 * 3)(S, 1, 0)
    _lib.check(_lib.load().prnu_gram_set_config(gram._ctx.handle, _lib.PRECISIONS[fmt], cfg))
    return (nkb + S - 1) // S * VALS[fmt]


@pytest.mark.parametrize("fmt", ["split_bf16", "fp16", "fp8"])
def test_slot_image_round_trip(fmt):
    n, count = 70 * 53, 90
    p = patterns(count, n, 3, seed=11)
    a = csi.PatternStore(n, count, fmt)
    a.put(p)
    img = torch.empty(a.image_bytes(40), dtype=torch.uint8, pin_memory=True)
    a.save(25, 40, img)
    assert bytes(img[:8].numpy()) == b"PRNUSLOT"
    b = csi.PatternStore(n, 300, fmt)
    b.put(p)                                    # slots 0..89 as in a
    b.load(200, img)                            # slots 25..64 of a at 200..239
    for x, y in zip(a.stats(40, 25), b.stats(40, 200)):
        assert np.array_equal(x, y)
    gram = csi.GramNCC(8, 8)
    out = torch.zeros((40, 40), dtype=torch.float64, device="cuda")
    slab = same_slabs(gram, fmt, n, 5)
    ga = a.gram(gram, count).cpu().numpy()
    plan = csi.GramPlan(gram, blocks=[(b, 200, 40, 200, 40, out)], slab_values=slab)
    plan.run()
    iu = np.triu_indices(40)
    assert np.array_equal(out.cpu().numpy()[iu], ga[25:65, 25:65][iu])
    plan.close()
    a.close()
    b.close()
    gram.close()


@pytest.mark.parametrize("fmt,device_blocks", [("fp16", 3), ("fp16", 4), ("split_bf16", 5)])
def test_host_tier_ncc_equals_resident_gram(fmt, device_blocks, tmp_path):
    n, count, block = 96 * 80, 300, 128          # blocks of 128, 128, 44
    p = patterns(count, n, 4, seed=7)
    gram = csi.GramNCC(8, 8)
    slab = same_slabs(gram, fmt, n, 6)
    st = csi.PatternStore(n, count, fmt)
    st.put(p)
    ref = st.gram(gram, count).cpu().numpy()
    st.close()
    tier = spill.HostPatternTier(n, count, block, fmt)
    for b in range(tier.nblocks):
        r0, r1 = tier.block_range(b)
        tier.put_block(b, p[r0:r1])
    info = {}
    got = tier.ncc(gram, slab_values=slab, device_blocks=device_blocks, stats=info)
    R = min(device_blocks - 2, 3)
    assert info["block_pairs"] == 6
    assert info["block_loads"] == {1: 6, 2: 4, 3: 3}[R]
    iu = np.triu_indices(count)
    assert np.array_equal(got[iu], ref[iu])
    assert np.array_equal(got, got.T)
    # the file tier: write, read back into a new tier, same result
    tier.save_files(str(tmp_path / "tier"))
    back = spill.HostPatternTier(n, count, block, fmt)
    back.load_files(str(tmp_path / "tier"))
    assert np.array_equal(back.ncc(gram, slab_values=slab, device_blocks=device_blocks), got)
    tier.close()
    back.close()
    gram.close()


def test_load_refuses_other_geometry():
    p = patterns(10, 64 * 64, 1, seed=3)
    a = csi.PatternStore(64 * 64, 10, "fp16")
    a.put(p)
    img = torch.empty(a.image_bytes(10), dtype=torch.uint8, pin_memory=True)
    a.save(0, 10, img)
    for other in (csi.PatternStore(64 * 65, 10, "fp16"), csi.PatternStore(64 * 64, 10, "split_bf16"),
                  csi.PatternStore(64 * 64, 9, "fp16")):
        with pytest.raises(_lib.PrnuError):
            other.load(0, img)
        other.close()
    junk = torch.zeros(a.image_bytes(10), dtype=torch.uint8)
    with pytest.raises(_lib.PrnuError):
        a.load(0, junk)
    a.close()


@pytest.mark.parametrize("fmt", ["split_bf16", "fp16"])
def test_host_tier_ncc_matches_oracle(fmt, tmp_path):
    """All pairs of 300 patterns streamed in blocks of 128 (ragged last block)
    through a device store of 3 blocks (one resident row block, so every
    later block streams), and again after a round trip through files,
    against the FP64 oracle NCC of the FP32 patterns."""
    n, count, block = 96 * 80, 300, 128
    p = patterns(count, n, 4, seed=17)
    want = O.ncc_matrix(p.cpu().numpy())
    off = ~np.eye(count, dtype=bool)
    gram = csi.GramNCC(8, 8)
    tier = spill.HostPatternTier(n, count, block, fmt)
    for b in range(tier.nblocks):
        r0, r1 = tier.block_range(b)
        tier.put_block(b, p[r0:r1])

    def check(got):
        assert np.array_equal(got, got.T)
        if fmt == "split_bf16":
            assert np.abs(got - want).max() <= 1e-5
        else:
            b_off, b_diag = fp16_bound(n)
            assert np.abs(got - want)[off].max() <= b_off
            assert np.abs(np.diag(got) - np.diag(want)).max() <= b_diag

    check(tier.ncc(gram, device_blocks=3))
    paths = tier.save_files(str(tmp_path / "tier"))
    back = spill.HostPatternTier(n, count, block, fmt)
    back.load_files(str(tmp_path / "tier"))
    check(back.ncc(gram, device_blocks=4))
    # a truncated block file is refused before any copy
    data = np.fromfile(paths[1], dtype=np.uint8)
    data[:-1].tofile(paths[1])
    with pytest.raises(ValueError):
        spill.HostPatternTier(n, count, block, fmt).load_files(str(tmp_path / "tier"))
    tier.close()
    back.close()
    gram.close()


def test_load_refuses_short_buffer():
    """prnu_store_load reads no byte past the buffer it is given: an image
    whose buffer is shorter than its header's count describes is refused."""
    p = patterns(12, 40 * 40, 1, seed=5)
    a = csi.PatternStore(40 * 40, 12, "split_bf16")
    a.put(p)
    img = torch.empty(a.image_bytes(12), dtype=torch.uint8, pin_memory=True)
    a.save(0, 12, img)
    b = csi.PatternStore(40 * 40, 12, "split_bf16")
    with pytest.raises(_lib.PrnuError):
        b.load(0, img[:-1])
    with pytest.raises(_lib.PrnuError):
        b.load(0, img[:32])                     # shorter than the header
    b.load(0, img)
    for x, y in zip(a.stats(12), b.stats(12)):
        assert np.array_equal(x, y)
    a.close()
    b.close()
