"""GPU tests of the device-cache residency path (rocket_amd/residency.py):
all-pairs NCC of pattern blocks scheduled leaf by leaf through the slot cache
(blocks filled on a copy stream on a miss, compared on the compute stream)
against the FP64 oracle NCC (split-BF16: |dNCC| <= 1e-5) and bit for bit
against the resident store's Gram with the same K slabs, in Hilbert and
tiling leaf order, with caches of 2 to 7 block slots, and with the cache's
hits, misses and evictions those of the same leaf order run through the
oracle's restatement of the reference's cache; blocks put from FP32 patterns
give the same bits as blocks loaded from host slot images; a later set that
shares blocks with an earlier one finds them resident."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import cache_oracle as CO  # noqa: E402
from oracle import prnu_oracle as O  # noqa: E402
from rocket_amd import csi, indexspace, residency, spill  # noqa: E402
from test_gpu_spill import patterns, same_slabs  # noqa: E402
from test_slotcache import simulate_schedule  # noqa: E402

pytestmark = pytest.mark.gpu

N, COUNT, BLOCK = 72 * 64, 230, 40                # 6 blocks: 5 of 40, one of 30


@pytest.fixture(scope="module")
def setup():
    p = patterns(COUNT, N, 4, seed=23)
    gram = csi.GramNCC(8, 8)
    slab = same_slabs(gram, "split_bf16", N, 5)
    st = csi.PatternStore(N, COUNT, "split_bf16")
    st.put(p)
    resident = st.gram(gram, COUNT).cpu().numpy()
    st.close()
    iu = np.triu_indices(COUNT)
    resident[(iu[1], iu[0])] = resident[iu]        # the store Gram fills the upper triangle
    tier = spill.HostPatternTier(N, COUNT, BLOCK, "split_bf16")
    for b in range(tier.nblocks):
        r0, r1 = tier.block_range(b)
        tier.put_block(b, p[r0:r1])
    want = O.ncc_matrix(p.cpu().numpy())
    yield p, gram, slab, resident, tier, want
    tier.close()
    gram.close()


def expected_stats(keys, order, nslots, window, oracle=None):
    o = oracle or CO.SlotCacheOracle(nslots)
    h0, m0, e0 = o.hits, o.misses, o.evictions
    pairs = indexspace.all_pairs(len(keys), order, 1, include_diagonal=True)
    loads = simulate_schedule(o, pairs, list(keys), window)
    return {"hits": o.hits - h0, "misses": o.misses - m0, "evictions": o.evictions - e0, "block_loads": loads}


@pytest.mark.parametrize("order,nslots,window", [("hilbert", 2, 1), ("hilbert", 3, 2), ("tiling", 3, 3),
                                                 ("tiling", 2, 2), ("hilbert", 7, 3)])
def test_cached_gram_matches_oracle_and_resident(setup, order, nslots, window):
    p, gram, slab, resident, tier, want = setup
    print(f"cached gram {order} slots={nslots} window={window}", flush=True)
    cg = residency.CachedBlockGram(gram, N, BLOCK, "split_bf16", nslots=nslots, window=window, slab_values=slab)
    got, st = cg.ncc(residency.TierSource(tier), range(tier.nblocks), order=order)
    assert np.array_equal(got, got.T)
    assert np.abs(got - want).max() <= 1e-5
    assert np.array_equal(got, resident)
    P = tier.nblocks
    assert st["leaves"] == P * (P + 1) // 2
    exp = expected_stats(range(P), order, nslots, window)
    assert {k: st[k] for k in exp} == exp
    sizes = [int(tier.images[k].numel()) for k in range(P)]
    assert min(sizes) * st["block_loads"] <= st["bytes_to_device"] <= max(sizes) * st["block_loads"]
    if nslots >= P:
        assert st["misses"] == P and st["evictions"] == 0
    cg.close()


def test_pattern_source_equals_tier_source(setup):
    """Blocks put from FP32 patterns on the device give the bits that blocks
    loaded from the host slot images give."""
    p, gram, slab, resident, tier, want = setup
    counts = {b: tier.block_range(b)[1] - tier.block_range(b)[0] for b in range(tier.nblocks)}

    def block(b):
        r0, r1 = tier.block_range(b)
        return p[r0:r1]

    cg = residency.CachedBlockGram(gram, N, BLOCK, "split_bf16", nslots=3, window=2, slab_values=slab)
    got, st = cg.ncc(residency.PatternSource(block, counts), range(tier.nblocks))
    assert np.array_equal(got, resident)
    assert st["bytes_to_device"] == 0
    cg.close()


def test_arriving_sets_reuse_resident_blocks(setup):
    """A second set sharing blocks with the first finds them in the cache (the
    same counts as the oracle's cache given both calls); its matrix is the
    resident Gram's rows and columns of those blocks."""
    p, gram, slab, resident, tier, want = setup
    cg = residency.CachedBlockGram(gram, N, BLOCK, "split_bf16", nslots=4, window=2, slab_values=slab)
    src = residency.TierSource(tier)
    o = CO.SlotCacheOracle(4)
    _, st1 = cg.ncc(src, [0, 1, 2, 3])
    assert {k: st1[k] for k in ("hits", "misses", "evictions", "block_loads")} == \
        expected_stats([0, 1, 2, 3], "hilbert", 4, 2, o)
    assert st1["misses"] == 4 and st1["evictions"] == 0
    keys = [2, 3, 5]
    got, st2 = cg.ncc(src, keys)
    exp = expected_stats(keys, "hilbert", 4, 2, o)
    assert {k: st2[k] for k in exp} == exp
    assert st2["hits"] >= 7 and st2["hits"] + st2["misses"] == 9   # 3 diagonal leaves + 3 of two keys
    idx = np.concatenate([np.arange(*tier.block_range(k)) for k in keys])
    assert np.array_equal(got, resident[np.ix_(idx, idx)])
    assert cg.cache.stats()["entries"] == 4
    cg.close()


def test_cache_sized_from_a_byte_budget(setup):
    p, gram, slab, resident, tier, want = setup
    probe = residency.CachedBlockGram(gram, N, BLOCK, "split_bf16", nslots=2, slab_values=slab)
    eb = probe.entry_bytes
    probe.close()
    cg = residency.CachedBlockGram(gram, N, BLOCK, "split_bf16", requested=3 * eb + 1, slab_values=slab)
    assert cg.nslots == 3
    got, _ = cg.ncc(residency.TierSource(tier), range(tier.nblocks))
    assert np.array_equal(got, resident)
    cg.close()


def test_failed_call_gives_back_its_entries(setup):
    """A block source that fails mid-call: the call raises, every entry it
    held is released (the unfilled one INVALID, refilled later), and the next
    call on the same cache gives the resident Gram's bits."""
    p, gram, slab, resident, tier, want = setup
    counts = {b: tier.block_range(b)[1] - tier.block_range(b)[0] for b in range(tier.nblocks)}

    def block(b):
        if b == 4:
            raise RuntimeError("block 4 unavailable")
        r0, r1 = tier.block_range(b)
        return p[r0:r1]

    cg = residency.CachedBlockGram(gram, N, BLOCK, "split_bf16", nslots=3, window=2, slab_values=slab)
    with pytest.raises(RuntimeError, match="block 4 unavailable"):
        cg.ncc(residency.PatternSource(block, counts), range(tier.nblocks))
    st = cg.cache.stats()
    assert st["waiting"] == 0 and st["unreferenced"] == st["entries"]   # nothing still referenced
    got, _ = cg.ncc(residency.TierSource(tier), range(tier.nblocks))
    assert np.array_equal(got, resident)
    cg.close()
