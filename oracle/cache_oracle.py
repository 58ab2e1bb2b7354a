"""CPU restatement of Rocket's device-cache residency policy (TEST
INFRASTRUCTURE ONLY).

This module is the checker of the native slot cache (``rocket_amd/csrc/
slotcache.cpp``), never the product: only ``tests/`` import it.

It restates, single-threaded, the behaviour of (``R/`` =
``src/main/java/nl/esciencecenter/rocket/``):

* ``R/cache/AbstractCache.java:19-409`` -- entries with a reference count and
  a status (WRITING / READONLY / INVALID / DELETED, ``:23-36``), pending
  readers of an entry being written, multi-key transactions served in FIFO
  order by ``makeProgress`` (``:311-349``), least-recently-used eviction of
  unreferenced entries (``evictOne``, ``:351-378``), ``cleanup`` (``:380-389``)
  and ``removeDuplicates`` (``:391-408``);
* ``R/util/LRUQueue.java:9-108`` -- the queue of unreferenced keys: a new or
  invalid entry goes to the front (evicted first, ``AbstractCache.java:52,
  225-229``), a released readable entry to the back;
* ``R/cache/DeviceCache.java:19-44,54-62`` -- the buffers are allocated one
  after another while the allocated bytes plus one more buffer stay below the
  requested total, kept on a stack (``ArrayDeque.push`` / ``poll``: the last
  allocated buffer is handed out first and a freed buffer is reused first);
* ``R/RocketLauncher.java:371,391-398`` -- the total: the requested size if it
  is positive, otherwise the free device memory plus that (negative) margin,
  capped at 5000 entries;
* ``R/scheduler/DeviceWorker.java:202-233`` -- an entry that comes back from
  a completed transaction in the WRITING state is a miss (the caller fills it
  and downgrades it to a reader); any other is a hit.

Buffers are numbered 0, 1, ... in allocation order (the slot of the native
cache).  Transactions are numbered 1, 2, ... in submission order; ``poll``
returns a completed transaction's entries once, as the reference's future
delivers them once.

Pinning (``tests/test_slotcache.py``): with one key per transaction, each
filled and released before the next, the policy is textbook LRU, so the
misses and the order of evictions equal those of an independent
``OrderedDict`` LRU of the same capacity on random key streams; hand-worked
scenarios from the Java fix the writer hand-over (a writer that releases
without filling passes the entry to the first pending reader, or leaves it
INVALID at the front of the queue), head-of-line blocking of a transaction
that does not fit, and the buffer stack; invariants (no referenced entry is
ever evicted, no two live entries share a slot, the entry count never
exceeds the slots) hold over random operation sequences; the sizing rule
against worked examples.  No function here is left unpinned.
"""
from __future__ import annotations

from collections import deque

WRITING, READONLY, INVALID, DELETED = "WRITING", "READONLY", "INVALID", "DELETED"

MAX_ENTRIES = 5000          # R/RocketLauncher.java:371


def device_cache_slots(entry_bytes: int, total_bytes: int) -> int:
    """Buffers DeviceCache allocates for a total of `total_bytes`: one more
    while (allocated + entry_bytes) < total (DeviceCache.java:27-31), i.e. the
    largest n with n * entry_bytes < total (exact allocations assumed)."""
    if entry_bytes <= 0:
        raise ValueError("entry size must be positive")
    n = 0
    while n * entry_bytes + entry_bytes < total_bytes:
        n += 1
    return n


def device_cache_total(requested: int, free_bytes: int, entry_bytes: int) -> int:
    """RocketLauncher.java:391-398: requested > 0 -> requested, else free +
    requested (a margin left free), then at most MAX_ENTRIES entries."""
    total = requested if requested > 0 else free_bytes + requested
    return min(total, MAX_ENTRIES * entry_bytes)


class _LRU:
    """LRUQueue: keys in a doubly linked order; duplicates and missing keys are
    errors (LRUQueue.java:29-31,50-52,92-97)."""

    def __init__(self):
        self.keys: list = []

    def add_first(self, key):
        if key in self.keys:
            raise ValueError("queue cannot contain duplicate entries")
        self.keys.insert(0, key)

    def add_last(self, key):
        if key in self.keys:
            raise ValueError("queue cannot contain duplicate entries")
        self.keys.append(key)

    def remove(self, key):
        if key not in self.keys:
            raise ValueError("key not found")
        self.keys.remove(key)

    def remove_first(self):
        return self.keys.pop(0) if self.keys else None


class _Entry:
    def __init__(self, key, slot, lru: _LRU):
        self.key, self.slot = key, slot
        self.meta = None
        self.refcount = 0
        self.status = INVALID
        self.pending: deque = deque()       # transactions waiting for the writer
        lru.add_first(key)                  # AbstractCache.java:52


class _Txn:
    def __init__(self, tid: int, keys: list):
        self.id = tid
        self.keys = keys
        self.entries: list = [None] * len(keys)
        self.pending = len(keys)
        self.complete = len(keys) == 0

    def trigger(self):
        self.pending -= 1
        if self.pending == 0:
            self.complete = True


def remove_duplicates(keys) -> list:
    """First occurrences in order, None dropped (AbstractCache.java:391-408)."""
    out = []
    for k in keys:
        if k is not None and k not in out:
            out.append(k)
    return out


class SlotCacheOracle:
    """AbstractCache + DeviceCache over `nslots` buffers numbered 0..n-1."""

    def __init__(self, nslots: int):
        self.free = deque()
        for s in range(nslots):             # buffers.push(alloc) in order
            self.free.appendleft(s)
        self.nslots = nslots
        self.entries: dict = {}
        self.lru = _LRU()
        self.queue: deque = deque()         # pending transactions
        self.txns: dict = {}
        self.next_id = 1
        self.hits = self.misses = self.evictions = 0
        self.evicted: list = []             # keys in eviction order

    # --- AbstractCache.Entry ----------------------------------------------
    def _acquire_entry(self, e: _Entry, t: _Txn):
        """Entry.acquire (AbstractCache.java:137-177)."""
        waiter = t
        if e.status == WRITING:
            e.pending.append(t)
            waiter = None
        elif e.status == READONLY:
            pass
        elif e.status == INVALID:
            e.status = WRITING
        else:
            raise RuntimeError("internal error, cache entry has invalid state")
        e.refcount += 1
        if e.refcount == 1:
            self.lru.remove(e.key)
        if waiter is not None:
            waiter.trigger()

    def ready(self, key, meta) -> None:
        """Entry.downgradeToReader (AbstractCache.java:103-121)."""
        e = self.entries.get(key)
        if e is None or e.status != WRITING:
            raise ValueError("cannot downgrade cache entry, entry is not writable")
        e.meta = meta
        e.status = READONLY
        waiting, e.pending = e.pending, deque()
        while waiting:
            waiting.popleft().trigger()

    def release(self, key) -> None:
        """Entry.release (AbstractCache.java:179-232)."""
        e = self.entries.get(key)
        if e is None or e.refcount <= 0:
            raise ValueError("cannot release cache entry since refcount is zero")
        waiter = None
        if e.status == WRITING:
            waiter = e.pending.popleft() if e.pending else None
            if waiter is None:
                e.status = INVALID
        elif e.status != READONLY:
            raise RuntimeError("internal error, cache entry has invalid state")
        e.refcount -= 1
        if e.refcount == 0:
            if e.status == READONLY:
                self.lru.add_last(key)
            else:
                self.lru.add_first(key)
            self._make_progress()
        if waiter is not None:
            waiter.trigger()

    def meta(self, key):
        """Entry.size (AbstractCache.java:76-84)."""
        e = self.entries.get(key)
        if e is None or e.status != READONLY:
            raise ValueError("cannot obtain size data while entry is writable")
        return e.meta

    # --- AbstractCache ----------------------------------------------------
    def try_acquire(self, key):
        """tryAcquireImmediately (AbstractCache.java:288-295, Entry :123-135):
        the slot of a readable entry, now referenced, else None."""
        e = self.entries.get(key)
        if e is None or e.status != READONLY:
            return None
        e.refcount += 1
        if e.refcount == 1:
            self.lru.remove(key)
        return e.slot

    def acquire(self, keys) -> int:
        """acquireAllAsync (AbstractCache.java:297-303): a transaction id."""
        t = _Txn(self.next_id, remove_duplicates(keys))
        self.next_id += 1
        self.txns[t.id] = t
        self.queue.append(t)
        self._make_progress()
        return t.id

    def poll(self, tid: int):
        """The completed transaction's [(key, slot, writer)] once (writer: the
        entry is WRITING, so the caller fills it -- DeviceWorker.java:208-226),
        or None while it waits."""
        t = self.txns.get(tid)
        if t is None:
            raise ValueError(f"unknown transaction {tid}")
        if not t.complete:
            return None
        del self.txns[tid]
        out = []
        for k, e in zip(t.keys, t.entries):
            w = e.status == WRITING
            self.misses += w
            self.hits += not w
            out.append((k, e.slot, w))
        return out

    def _make_progress(self):
        """AbstractCache.java:311-349."""
        done = False
        while not done and self.queue:
            t = self.queue.popleft()
            success = True
            for i, key in enumerate(t.keys):
                if t.entries[i] is not None:
                    continue
                e = self.entries.get(key)
                if e is not None:
                    t.entries[i] = e
                    self._acquire_entry(e, t)
                    continue
                if self.free:                       # createBuffer: buffers.poll()
                    e = _Entry(key, self.free.popleft(), self.lru)
                    self.entries[key] = e
                    t.entries[i] = e
                    self._acquire_entry(e, t)
                    continue
                success = False
            if not success:
                self.queue.appendleft(t)
                done = not self._evict_one()

    def _evict_one(self) -> bool:
        """AbstractCache.java:351-378 with DeviceCache.destroyBuffer."""
        key = self.lru.remove_first()
        if key is None:
            return False
        e = self.entries.pop(key)
        if e.refcount != 0 or e.status == DELETED or e.pending:
            raise RuntimeError("internal error, cannot delete invalid cache entry")
        e.status = DELETED
        self.free.appendleft(e.slot)                # buffers.push(buffer)
        self.evictions += 1
        self.evicted.append(key)
        return True

    def cleanup(self) -> None:
        """AbstractCache.java:380-389."""
        while self._evict_one():
            pass
        if self.lru.keys or self.entries:
            raise ValueError(f"failed to clean up cache, {len(self.entries)} entries are still in use!")

    # --- inspection -------------------------------------------------------
    def resident(self) -> dict:
        """key -> (slot, status, refcount) of every live entry."""
        return {k: (e.slot, e.status, e.refcount) for k, e in self.entries.items()}

    def lru_order(self) -> list:
        return list(self.lru.keys)

    def waiting(self) -> int:
        return len(self.queue)
