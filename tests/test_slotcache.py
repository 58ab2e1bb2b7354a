"""The device-cache residency policy: the oracle's restatement of
R/cache/AbstractCache.java + DeviceCache.java pinned against textbook LRU and
hand-worked scenarios from the Java, and the native slot cache
(prnu_slotcache_*, host-only C++) checked against the oracle operation by
operation on random sequences.  CPU only: the cache makes no device call."""
import random
from collections import OrderedDict

import pytest

from oracle import cache_oracle as C
from rocket_amd import _lib
from rocket_amd.residency import SlotCache


# ---------------------------------------------------------------- oracle pins
def textbook_lru(stream, capacity):
    """Misses and evicted keys of a plain LRU cache (OrderedDict)."""
    d, misses, evicted = OrderedDict(), 0, []
    for k in stream:
        if k in d:
            d.move_to_end(k)
            continue
        misses += 1
        if len(d) == capacity:
            evicted.append(d.popitem(last=False)[0])
        d[k] = True
    return misses, evicted


def use_one(cache, key):
    """Acquire one key, fill it on a miss, release it (one correlation's
    worth of DeviceWorker use with a single pattern)."""
    t = cache.acquire([key])
    got = cache.poll(t)
    assert got is not None and len(got) == 1
    _, _, writer = got[0]
    if writer:
        cache.ready(key, 1)
    cache.release(key)
    return writer


@pytest.mark.parametrize("capacity,universe,seed", [(2, 5, 0), (3, 8, 1), (5, 12, 2), (8, 9, 3)])
def test_oracle_single_key_use_is_textbook_lru(capacity, universe, seed):
    rng = random.Random(seed)
    stream = [rng.randrange(universe) for _ in range(400)]
    o = C.SlotCacheOracle(capacity)
    misses = sum(use_one(o, k) for k in stream)
    want_misses, want_evicted = textbook_lru(stream, capacity)
    assert misses == want_misses == o.misses
    assert o.evicted == want_evicted
    assert o.hits == len(stream) - misses


def test_oracle_buffer_stack():
    """The last allocated buffer is handed out first; a freed one is reused first."""
    o = C.SlotCacheOracle(3)
    slots = []
    for k in "abc":
        slots.append(o.poll(o.acquire([k]))[0][1])
        o.ready(k, 0)
        o.release(k)
    assert slots == [2, 1, 0]
    got = o.poll(o.acquire(["d"]))                 # evicts a (least recently released)
    assert o.evicted == ["a"] and got == [("d", 2, True)]


def test_oracle_writer_hand_over():
    """A writer that releases without filling passes the entry to the first
    pending reader, which becomes the writer (AbstractCache.java:199-208)."""
    o = C.SlotCacheOracle(2)
    t1 = o.acquire(["A"])
    assert o.poll(t1) == [("A", 1, True)]
    t2 = o.acquire(["A"])
    assert o.poll(t2) is None                       # waits for the writer
    o.release("A")                                  # t1 gives up unfilled
    assert o.poll(t2) == [("A", 1, True)]           # t2 must fill it
    o.ready("A", 7)
    assert o.meta("A") == 7
    t3 = o.acquire(["A"])
    assert o.poll(t3) == [("A", 1, False)]          # a hit
    assert (o.hits, o.misses) == (1, 2)


def test_oracle_ready_wakes_readers():
    o = C.SlotCacheOracle(2)
    o.poll(o.acquire(["A"]))
    t2, t3 = o.acquire(["A"]), o.acquire(["A"])
    assert o.poll(t2) is None and o.poll(t3) is None
    o.ready("A", 1)
    assert o.poll(t2) == [("A", 1, False)] and o.poll(t3) == [("A", 1, False)]
    assert o.resident()["A"] == (1, C.READONLY, 3)


def test_oracle_invalid_entry_is_evicted_first():
    """A released unfilled entry goes to the front of the LRU queue
    (AbstractCache.java:225-229)."""
    o = C.SlotCacheOracle(2)
    o.poll(o.acquire(["A"]))
    o.ready("A", 1)
    o.release("A")                                  # lru [A]
    sb = o.poll(o.acquire(["B"]))[0][1]
    o.release("B")                                  # unfilled -> INVALID, lru [B, A]
    assert o.lru_order() == ["B", "A"]
    assert o.resident()["B"] == (sb, C.INVALID, 0)
    assert o.poll(o.acquire(["C"])) == [("C", sb, True)]
    assert o.evicted == ["B"] and "A" in o.resident()


def test_oracle_head_of_line_blocking():
    """A transaction that does not fit waits at the head and blocks later
    ones, even ones whose keys are resident (AbstractCache.java:319-348)."""
    o = C.SlotCacheOracle(2)
    got = o.poll(o.acquire(["A", "B"]))
    assert [g[2] for g in got] == [True, True]
    sa, sb = got[0][1], got[1][1]
    o.ready("A", 1)
    o.ready("B", 1)
    t2 = o.acquire(["C"])
    t3 = o.acquire(["A"])
    assert o.poll(t2) is None and o.poll(t3) is None and o.waiting() == 2
    o.release("A")                                  # A unreferenced: t2 evicts it
    assert o.poll(t2) == [("C", sa, True)]
    assert o.poll(t3) is None                       # A is gone; no slot until B is released
    o.release("B")
    assert o.poll(t3) == [("A", sb, True)]
    assert o.evicted == ["A", "B"]


def test_oracle_partial_transaction_holds_its_entries():
    o = C.SlotCacheOracle(2)
    o.poll(o.acquire(["X"]))
    t = o.acquire(["A", "B"])                       # A gets the last slot, B must wait
    assert o.poll(t) is None
    assert o.resident()["A"][2] == 1
    o.ready("X", 1)
    o.release("X")
    got = o.poll(t)
    assert [g[0] for g in got] == ["A", "B"] and all(g[2] for g in got)


def test_oracle_errors_and_cleanup():
    o = C.SlotCacheOracle(2)
    with pytest.raises(ValueError):
        o.release("A")
    o.poll(o.acquire(["A"]))
    with pytest.raises(ValueError):
        o.meta("A")                                 # still being written
    o.ready("A", 3)
    with pytest.raises(ValueError):
        o.ready("A", 3)
    with pytest.raises(ValueError):
        o.cleanup()                                 # A is referenced
    o.release("A")
    with pytest.raises(ValueError):
        o.release("A")
    o.cleanup()
    assert o.resident() == {} and len(o.free) == 2


def test_oracle_duplicates_removed():
    assert C.remove_duplicates([3, 1, 3, None, 2, 1]) == [3, 1, 2]
    o = C.SlotCacheOracle(2)
    got = o.poll(o.acquire([5, 5, 5]))
    assert len(got) == 1 and o.resident()[5][2] == 1


def test_sizing_worked_examples():
    # one more buffer while allocated + entry < total
    assert C.device_cache_slots(5, 10) == 1
    assert C.device_cache_slots(5, 11) == 2
    assert C.device_cache_slots(5, 5) == 0
    assert C.device_cache_slots(5, 26) == 5
    # total: request, or free + (negative) margin, at most 5000 entries
    assert C.device_cache_total(100, 10 ** 9, 10) == 100
    assert C.device_cache_total(-300, 1000, 10) == 700
    assert C.device_cache_total(0, 10 ** 9, 10) == 50000
    for eb, req, free in [(5, 10, 0), (5, 11, 0), (7, -20, 1000), (1 << 20, 0, 1 << 34), (3, 2, 0), (4, -10, 5)]:
        want = C.device_cache_slots(eb, C.device_cache_total(req, free, eb))
        assert SlotCache.slots_for(eb, req, free) == want


# ------------------------------------------------- native cache vs the oracle
def snapshot_native(c: SlotCache, universe):
    res = {}
    for k in universe:
        e = c.entry(k)
        if e is not None:
            res[k] = e
    return res


def run_differential(nslots, universe, nops, seed, mix=(0.3, 0.55, 0.7, 0.95)):
    """`mix`: cumulative odds of acquire, poll, fill, release (the rest:
    try_acquire)."""
    rng = random.Random(seed)
    p_acq, p_poll, p_fill, p_rel = mix
    o = C.SlotCacheOracle(nslots)
    n = SlotCache(nslots)
    keys = list(range(universe))
    open_txns = []          # submitted, not yet complete
    held = []               # (key, writer at completion) references we own
    for _ in range(nops):
        r = rng.random()
        if r < p_acq or (not open_txns and not held):
            # at most nslots keys: a larger transaction never completes (it
            # waits at the head for a slot it already holds), in the
            # reference as here
            ks = [rng.choice(keys) for _ in range(rng.randint(1, min(3, nslots)))]
            assert o.acquire(ks) == n.acquire(ks)
            open_txns.append(o.next_id - 1)
        elif r < p_poll and open_txns:
            t = rng.choice(open_txns)
            a, b = o.poll(t), n.poll(t)
            assert a == b
            if a is not None:
                open_txns.remove(t)
                for k, _, w in a:
                    if w and rng.random() < 0.7:      # most writers fill at once
                        o.ready(k, k + 100)
                        n.ready(k, k + 100)
                        w = False
                    held.append((k, w))
        elif r < p_fill and held:
            i = rng.randrange(len(held))
            k, w = held[i]
            if w and o.resident()[k][1] == C.WRITING:
                m = rng.randrange(100)
                o.ready(k, m)
                n.ready(k, m)
                held[i] = (k, False)
            else:
                assert o.meta(k) == n.meta(k)
        elif r < p_rel and held:
            k, _ = held.pop(rng.randrange(len(held)))
            o.release(k)
            n.release(k)
        else:
            k = rng.choice(keys)
            a, b = o.try_acquire(k), n.try_acquire(k)
            assert a == b
            if a is not None:
                held.append((k, False))
        # the whole state, after every operation
        want = {k: (s, st, rc) for k, (s, st, rc) in o.resident().items()}
        assert snapshot_native(n, keys) == want
        assert n.lru_order() == o.lru_order()
        st = n.stats()
        assert (st["hits"], st["misses"], st["evictions"]) == (o.hits, o.misses, o.evictions)
        assert st["waiting"] == o.waiting() and st["free_slots"] == len(o.free)
        assert len(want) <= nslots and len({v[0] for v in want.values()}) == len(want)
    # drain: complete what can complete, release everything, clean up
    while held or open_txns:
        for k, _ in held:
            o.release(k)
            n.release(k)
        held = []
        for t in list(open_txns):
            a, b = o.poll(t), n.poll(t)
            assert a == b
            if a is not None:
                open_txns.remove(t)
                held.extend((k, w) for k, _, w in a)
        assert held or not open_txns, f"transactions {open_txns} cannot complete with nothing held"
    o.cleanup()
    n.cleanup()
    assert n.stats()["entries"] == 0 and n.stats()["free_slots"] == nslots
    n.close()
    return o


@pytest.mark.parametrize("nslots,universe,seed", [(2, 4, 0), (3, 6, 1), (4, 10, 2), (6, 8, 3), (2, 12, 4),
                                                  (5, 5, 5)])
@pytest.mark.parametrize("mix", [(0.3, 0.55, 0.7, 0.95), (0.15, 0.4, 0.45, 0.97)], ids=["crowded", "relaxed"])
def test_native_cache_equals_oracle(nslots, universe, seed, mix):
    o = run_differential(nslots, universe, 1500, seed, mix)
    assert o.evictions > 0 or universe <= nslots
    assert o.misses > 0 and o.hits > 0


def test_native_single_key_use_is_textbook_lru():
    rng = random.Random(9)
    stream = [rng.randrange(7) for _ in range(500)]
    n = SlotCache(3)
    misses = sum(use_one(n, k) for k in stream)
    assert misses == textbook_lru(stream, 3)[0] == n.stats()["misses"]
    n.close()


def test_native_errors():
    with pytest.raises(_lib.PrnuInvalidArgument):
        SlotCache(1)
    n = SlotCache(2)
    with pytest.raises(_lib.PrnuInvalidArgument):
        n.release(4)
    t = n.acquire([4])
    assert n.poll(t) == [(4, 1, True)]
    with pytest.raises(_lib.PrnuInvalidArgument):
        n.poll(t)                                   # already retired
    with pytest.raises(_lib.PrnuInvalidArgument):
        n.meta(4)
    with pytest.raises(_lib.PrnuInvalidArgument):
        n.cleanup()
    n.ready(4, 2)
    with pytest.raises(_lib.PrnuInvalidArgument):
        n.ready(4, 2)
    n.release(4)
    n.cleanup()
    assert n.acquire([]) > 0
    n.close()


# ------------------------------------------- the block scheduler's cache use
def simulate_schedule(cache, pairs, keys, window):
    """The cache calls CachedBlockGram.ncc makes for a leaf order (at most
    `window` leaves in flight, the oldest retired first, also while a
    transaction waits), on any cache with acquire/poll/ready/release; returns
    the block loads (writers filled)."""
    from collections import deque
    inflight, loads = deque(), 0

    def retire():
        for k in inflight.popleft():
            cache.release(k)

    for a, b in pairs:
        if len(inflight) == window:
            retire()
        t = cache.acquire([keys[int(a)], keys[int(b)]])
        got = cache.poll(t)
        while got is None:
            retire()
            got = cache.poll(t)
        for k, _, w in got:
            if w:
                cache.ready(k, 1)
                loads += 1
        inflight.append([k for k, _, _ in got])
    while inflight:
        retire()
    return loads


@pytest.mark.parametrize("P,S,window", [(16, 4, 1), (32, 8, 1), (32, 8, 3), (64, 16, 2)])
def test_hilbert_order_reloads_fewer_blocks(P, S, window):
    """Leaves of the block triangle in Hilbert order against the tiling order
    with unit tiles (row after row): the curve keeps its working set in a
    few slots.  The native cache and the oracle agree on every count."""
    from rocket_amd import indexspace
    loads = {}
    for order in ("hilbert", "tiling"):
        pairs = indexspace.all_pairs(P, order, 1, include_diagonal=True)
        o = C.SlotCacheOracle(S)
        n = SlotCache(S)
        lo = simulate_schedule(o, pairs, list(range(P)), window)
        ln = simulate_schedule(n, pairs, list(range(P)), window)
        st = n.stats()
        assert ln == lo and (st["hits"], st["misses"], st["evictions"]) == (o.hits, o.misses, o.evictions)
        assert lo == o.misses
        n.close()
        loads[order] = lo
    assert loads["hilbert"] < loads["tiling"]
