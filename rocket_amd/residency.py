"""Device-cache residency: pattern blocks kept in the slots of one device
store by Rocket's cache policy, and all-pairs NCC scheduled through it.

:class:`SlotCache` binds the native policy (``prnu_slotcache_*``, the
behaviour of ``R/cache/AbstractCache.java:19-409`` over the buffer stack of
``R/cache/DeviceCache.java:19-62``).  :class:`CachedBlockGram` is the device
worker (``R/scheduler/DeviceWorker.java:202-320``) at the granularity this
GPU wants: a cache entry is a *block* of patterns (a contiguous range of
store slots, so each leaf is one rectangular Gram launch), and a leaf of the
index space over the blocks (``indexspace.all_pairs``, Hilbert or tiling
order) is one transaction of two keys.  A miss is filled on a copy stream
(from a host slot image, or from FP32 patterns on the device) and downgraded
to a reader; the leaf's Gram runs on the compute stream once both blocks are
in; its entries are released when its result has reached the host.  At most
``window`` leaves are in flight (the reference's task semaphore), and a slot
is refilled only after the last launch that read it (an event per slot).

The cache object outlives a call: a later set of blocks that shares keys with
an earlier one finds them resident (hits), which is the reuse policy for an
arriving stream of pattern sets.

(``R/`` = ``src/main/java/nl/esciencecenter/rocket/``.)
"""
from __future__ import annotations

import ctypes
import time
from collections import deque

import numpy as np

from . import _lib, csi, indexspace

STATUS = {0: "WRITING", 1: "READONLY", 2: "INVALID"}


class SlotCache:
    """The residency policy over `nslots` slots (``prnu_slotcache_*``)."""

    def __init__(self, nslots: int):
        h = ctypes.c_void_p()
        _lib.check(_lib.load().prnu_slotcache_create(int(nslots), ctypes.byref(h)), "prnu_slotcache_create")
        self.handle = h
        self.nslots = int(nslots)

    @staticmethod
    def slots_for(entry_bytes: int, requested: int, free_bytes: int | None = None) -> int:
        """Slots of `entry_bytes` for a requested size, or for the free device
        memory less a margin when `requested` <= 0 (the device cache's sizing:
        ``R/RocketLauncher.java:391-398``, ``DeviceCache.java:27-31``)."""
        if free_bytes is None:
            free, total = ctypes.c_size_t(), ctypes.c_size_t()
            _lib.call("prnu_mem_info", ctypes.byref(free), ctypes.byref(total))
            free_bytes = free.value
        n = ctypes.c_int64()
        _lib.call("prnu_slotcache_slots_for", int(entry_bytes), int(requested), int(free_bytes), ctypes.byref(n))
        return n.value

    def acquire(self, keys) -> int:
        arr = np.ascontiguousarray(np.asarray(list(keys), dtype=np.int64))
        t = ctypes.c_int64()
        _lib.call("prnu_slotcache_acquire", self.handle, arr.ctypes.data if arr.size else None, int(arr.size),
                  ctypes.byref(t))
        return t.value

    def poll(self, txn: int):
        """[(key, slot, writer)] of a completed transaction (once), else None."""
        done, n = ctypes.c_int(), ctypes.c_int()
        _lib.call("prnu_slotcache_poll", self.handle, int(txn), ctypes.byref(done), None, None, None, 0,
                  ctypes.byref(n))
        if not done.value:
            return None
        keys = np.zeros(max(1, n.value), dtype=np.int64)
        slots = np.zeros(max(1, n.value), dtype=np.int64)
        writer = np.zeros(max(1, n.value), dtype=np.int32)
        _lib.call("prnu_slotcache_poll", self.handle, int(txn), ctypes.byref(done), keys.ctypes.data,
                  slots.ctypes.data, writer.ctypes.data, int(keys.size), ctypes.byref(n))
        return [(int(keys[i]), int(slots[i]), bool(writer[i])) for i in range(n.value)]

    def ready(self, key: int, meta: int) -> None:
        _lib.call("prnu_slotcache_ready", self.handle, int(key), int(meta))

    def release(self, key: int) -> None:
        _lib.call("prnu_slotcache_release", self.handle, int(key))

    def try_acquire(self, key: int):
        s = ctypes.c_int64()
        _lib.call("prnu_slotcache_try_acquire", self.handle, int(key), ctypes.byref(s))
        return None if s.value < 0 else s.value

    def meta(self, key: int) -> int:
        m = ctypes.c_int64()
        _lib.call("prnu_slotcache_meta", self.handle, int(key), ctypes.byref(m))
        return m.value

    def entry(self, key: int):
        """(slot, status name, refcount) of a resident key, else None."""
        s, st, rc = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int64()
        _lib.call("prnu_slotcache_entry", self.handle, int(key), ctypes.byref(s), ctypes.byref(st), ctypes.byref(rc))
        return None if st.value < 0 else (s.value, STATUS[st.value], rc.value)

    def lru_order(self) -> list:
        n = ctypes.c_int64()
        _lib.call("prnu_slotcache_lru", self.handle, None, 0, ctypes.byref(n))
        out = np.zeros(max(1, n.value), dtype=np.int64)
        _lib.call("prnu_slotcache_lru", self.handle, out.ctypes.data, int(out.size), ctypes.byref(n))
        return [int(k) for k in out[: n.value]]

    def stats(self) -> dict:
        s = np.zeros(8, dtype=np.int64)
        _lib.call("prnu_slotcache_stats", self.handle, s.ctypes.data)
        names = ("hits", "misses", "evictions", "entries", "free_slots", "waiting", "slots", "unreferenced")
        return {k: int(v) for k, v in zip(names, s)}

    def cleanup(self) -> None:
        _lib.call("prnu_slotcache_cleanup", self.handle)

    def close(self) -> None:
        if self.handle:
            _lib.load().prnu_slotcache_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TierSource:
    """Blocks of a :class:`spill.HostPatternTier`: key b is the tier's block b,
    filled by one copy of its slot image (the host-cache hit path,
    ``DeviceWorker.java:165-176``)."""

    def __init__(self, tier):
        self.tier = tier

    def count(self, key: int) -> int:
        r0, r1 = self.tier.block_range(int(key))
        return r1 - r0

    def nbytes(self, key: int) -> int:
        return int(self.tier.images[int(key)].numel())

    def load(self, key: int, store, first_slot: int) -> None:
        store.load(first_slot, self.tier.images[int(key)], sync=False)


class PatternSource:
    """Blocks produced on the device: ``fn(key)`` returns the block's FP32
    CUDA patterns (rows of n values), e.g. from extraction of its frames (the
    input path of ``DeviceWorker.java:122-158``), put into the slots with
    their statistics."""

    def __init__(self, fn, counts: dict):
        self.fn, self.counts = fn, dict(counts)

    def count(self, key: int) -> int:
        return int(self.counts[int(key)])

    def nbytes(self, key: int) -> int:
        return 0

    def load(self, key: int, store, first_slot: int) -> None:
        p = self.fn(int(key))
        if p.shape[0] != self.count(key):
            raise ValueError(f"block {key}: {p.shape[0]} patterns, expected {self.count(key)}")
        store.put(p, first_slot, sync=False)


class CachedBlockGram:
    """All-pairs NCC of sets of pattern blocks through a device cache of
    `nslots` block slots of `block` patterns of `n` values (store `format`).

    `nslots` None sizes the cache as the device cache does: `requested`
    bytes, or the free device memory less `-requested` bytes when
    `requested` <= 0; the default leaves 500 MiB free, the reference's
    default (``R/RocketLauncherArgs.java:14-15``)."""

    def __init__(self, gram: "csi.GramNCC", n: int, block: int, format: str = "split_bf16",
                 nslots: int | None = None, requested: int = -500 * 1024 * 1024, window: int = 3,
                 slab_values: int = 0):
        import torch
        if block < 1 or window < 1:
            raise ValueError("block and window must be positive")
        self.gram, self.n, self.block, self.format = gram, int(n), int(block), format
        self.window, self.slab_values = int(window), int(slab_values)
        probe = csi.PatternStore(self.n, self.block, format)
        self.entry_bytes = probe.image_bytes(self.block) - 64
        probe.close()
        if nslots is None:
            nslots = SlotCache.slots_for(self.entry_bytes, requested)
        self.cache = SlotCache(nslots)
        self.nslots = self.cache.nslots
        self.store = csi.PatternStore(self.n, self.nslots * self.block, format)
        self.compute, self.copy = torch.cuda.Stream(), torch.cuda.Stream()
        self.loaded = [torch.cuda.Event() for _ in range(self.nslots)]
        self.used = [torch.cuda.Event() for _ in range(self.nslots)]
        B, W = self.block, self.window
        self.dbuf = [torch.empty((B, B), dtype=torch.float64, device="cuda") for _ in range(W)]
        self.hbuf = [torch.empty(B * B, dtype=torch.float64, pin_memory=True) for _ in range(W)]
        self.done = [torch.cuda.Event() for _ in range(W)]
        self.plans: dict = {}
        self.totals = {"leaves": 0, "block_loads": 0, "bytes_to_device": 0}

    def warm(self, count: int) -> None:
        """Build, and run once, the launch of every ordered slot pair for
        blocks of `count` patterns, so that later calls on such blocks create
        no plan (the slots' contents are whatever they hold)."""
        import torch
        B = self.block
        with torch.cuda.stream(self.compute):
            for sa in range(self.nslots):
                for sb in range(self.nslots):
                    for k in range(self.window):
                        pk = (sa, count, sb, count, k)
                        if pk not in self.plans:
                            self.plans[pk] = csi.GramPlan(self.gram, blocks=[(self.store, sa * B, count, sb * B, count,
                                                                              self.dbuf[k][:count, :count])],
                                                          slab_values=self.slab_values)
                        self.plans[pk].run(use_graph=True, sync=False)
        self.compute.synchronize()

    def reset(self) -> None:
        """Evict every block (AbstractCache.cleanup): the next call starts cold."""
        self.cache.cleanup()

    # one leaf in flight: (buffer, keys to release, result placement)
    def _retire(self, inflight: deque, result: np.ndarray) -> None:
        k, keys, (r0, c0, rn, cn) = inflight.popleft()
        self.done[k].synchronize()
        result[r0:r0 + rn, c0:c0 + cn] = self.hbuf[k][:rn * cn].view(rn, cn).numpy()
        for key in keys:
            self.cache.release(key)

    def ncc(self, source, keys, order: str = "hilbert", param: int = 1):
        """NCC of every pair of patterns in the blocks `keys` (in that order)
        as a symmetric float64 matrix, and the call's statistics (cache hits,
        misses and evictions, block loads, bytes copied to the device,
        leaves, seconds).  If the call fails (a block source raising, say),
        it first gives back every entry it holds, as the reference's worker
        releases its entries whatever the outcome (DeviceWorker.java:
        285-293): a writer released unfilled leaves its entry INVALID, to be
        filled by the next call that asks for it."""
        import torch
        keys = [int(k) for k in keys]
        if len(set(keys)) != len(keys):
            raise ValueError("block keys must be distinct")
        counts = [source.count(k) for k in keys]
        if any(c < 1 or c > self.block for c in counts):
            raise ValueError(f"block sizes must be in [1, {self.block}]")
        offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        total = int(offs[-1])
        result = np.zeros((total, total), dtype=np.float64)
        before = self.cache.stats()
        leaves = indexspace.all_pairs(len(keys), order, param, include_diagonal=True)
        inflight: deque = deque()
        held: list = []                      # keys of the leaf being set up, until it is in flight
        walk = {"loads": 0, "moved": 0}
        t0 = time.perf_counter()
        try:
            self._walk(source, keys, counts, offs, leaves, inflight, held, result, walk)
            while inflight:
                self._retire(inflight, result)
        except BaseException:
            torch.cuda.synchronize()
            while inflight:
                _, ks, _ = inflight.popleft()
                for key in ks:
                    self.cache.release(key)
            for key in held:
                self.cache.release(key)
            raise
        seconds = time.perf_counter() - t0
        il = np.tril_indices(total, -1)
        result[il] = result.T[il]
        after = self.cache.stats()
        stats = {k: after[k] - before[k] for k in ("hits", "misses", "evictions")}
        stats.update({"block_loads": walk["loads"], "bytes_to_device": walk["moved"], "leaves": len(leaves),
                      "seconds": seconds, "slots": self.nslots, "block": self.block})
        self.totals["leaves"] += len(leaves)
        self.totals["block_loads"] += walk["loads"]
        self.totals["bytes_to_device"] += walk["moved"]
        return result, stats

    def _walk(self, source, keys, counts, offs, leaves, inflight: deque, held: list, result: np.ndarray,
              walk: dict) -> None:
        """The leaves in order: acquire, fill misses, launch, keep `window`
        leaves in flight.  `held` names the keys of the leaf being set up."""
        import torch
        B = self.block
        for step, (a, b) in enumerate(leaves):
            a, b = int(a), int(b)
            ka, kb = keys[a], keys[b]
            if len(inflight) == self.window:             # the worker's task semaphore
                self._retire(inflight, result)
            txn = self.cache.acquire([ka, kb])
            got = self.cache.poll(txn)
            while got is None:                           # every slot referenced: finish older leaves
                if not inflight:
                    raise RuntimeError("slot cache transaction cannot complete with no leaf in flight")
                self._retire(inflight, result)
                got = self.cache.poll(txn)
            held[:] = [kk for kk, _, _ in got]
            slot = {}
            for key, s, writer in got:
                slot[key] = s
                if writer:
                    self.copy.wait_event(self.used[s])
                    with torch.cuda.stream(self.copy):
                        source.load(key, self.store, s * B)
                    self.loaded[s].record(self.copy)
                    self.cache.ready(key, source.count(key))
                    walk["loads"] += 1
                    walk["moved"] += source.nbytes(key)
            sa, sb = slot[ka], slot[kb]
            na, nb = counts[a], counts[b]
            k = step % self.window
            self.compute.wait_event(self.loaded[sa])
            if sb != sa:
                self.compute.wait_event(self.loaded[sb])
            pk = (sa, na, sb, nb, k)
            with torch.cuda.stream(self.compute):
                if pk not in self.plans:
                    self.plans[pk] = csi.GramPlan(self.gram, blocks=[(self.store, sa * B, na, sb * B, nb,
                                                                      self.dbuf[k][:na, :nb])],
                                                  slab_values=self.slab_values)
                self.plans[pk].run(use_graph=True, sync=False)
                self.hbuf[k][:na * nb].view(na, nb).copy_(self.dbuf[k][:na, :nb], non_blocking=True)
            self.used[sa].record(self.compute)
            self.used[sb].record(self.compute)
            self.done[k].record(self.compute)
            inflight.append((k, list(held), (int(offs[a]), int(offs[b]), na, nb)))
            held.clear()

    def close(self) -> None:
        import torch
        torch.cuda.synchronize()
        for p in self.plans.values():
            p.close()
        self.plans = {}
        if self.cache.handle:
            self.cache.cleanup()
            self.cache.close()
        self.store.close()
