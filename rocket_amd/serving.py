"""Camera-identification serving with continuous batching (no reference
counterpart: Rocket only runs offline all-pairs jobs).

A :class:`FingerprintStore` keeps camera fingerprints resident in HBM
(``prnu_fpstore_*``).  A :class:`Server` answers queries -- one 8-bit image
each, "which cameras does this image come from?" -- with the top-k stored
fingerprints by NCC:

* requests are admitted into in-flight batches ("slots") whenever a slot is
  free and free HBM allows, up to ``max_batch`` queries per batch;
* each slot owns a context (extraction workspace, FFT plans, scan workspace)
  and a HIP stream, so the extraction of one batch runs while the store scan
  of another is still on the device;
* per batch: host->device copy of the images, PRNU extraction of the whole
  batch (``prnu_preprocess_batch``, one launch per filter stage), then
  ``prnu_fpstore_query`` (FP8 MFMA scan, top-k, exact FP32 re-score), then
  device->host copy of the k results;
* a completion thread waits for each batch's event and resolves its requests;
* host-side pauses are kept off the request path: the pinned staging buffers
  are written once when a slot is made (a first write into fresh pinned pages
  faults every page in, 15-20 ms for a batch of 1-megapixel images), and
  while the server runs the objects alive at ``start()`` are frozen out of
  Python's cyclic collector (``gc.freeze``), so a full collection walks only
  what was made since instead of the whole interpreter (a full collection of
  a process holding torch took 34 ms).

A query's result does not depend on which other queries share its batch
(``prnu_fpstore_query`` is per-query by construction), so results are the
same alone and in any mix.
"""
from __future__ import annotations

import ctypes
import gc
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .csi import _Ctx, pattern_pointer_table

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


class FingerprintStore:
    """``prnu_fpstore``: FP8 scan copy + FP32 originals + FP64 stats."""

    def __init__(self, height: int, width: int, capacity: int):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.prnu_fpstore_create(int(height) * int(width), int(capacity), ctypes.byref(h)),
                   "prnu_fpstore_create")
        self.handle = h
        self.height, self.width, self.capacity = int(height), int(width), int(capacity)

    def put(self, patterns, first_slot: int = 0, sync: bool = True) -> None:
        ptrs = pattern_pointer_table(patterns)
        _lib.call("prnu_fpstore_put", self.handle, ptrs.data_ptr(), int(ptrs.numel()), int(first_slot),
                  torch.cuda.current_stream().cuda_stream)
        if sync:
            torch.cuda.current_stream().synchronize()

    def info(self) -> dict:
        n, cap, size, nbytes = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        _lib.call("prnu_fpstore_info", self.handle, ctypes.byref(n), ctypes.byref(cap), ctypes.byref(size),
                  ctypes.byref(nbytes))
        return {"n": n.value, "capacity": cap.value, "size": size.value, "bytes": nbytes.value}

    def query(self, ctx: _Ctx, patterns, k: int = 10, scan: bool = False, sync: bool = True):
        """Top-k of each query pattern: (idx int32 (nq, k), score float64 (nq, k)) as
        CUDA tensors, plus the FP8 scan scores (nq, size) when ``scan``."""
        ptrs = pattern_pointer_table(patterns)
        nq = int(ptrs.numel())
        idx = torch.empty((nq, k), dtype=torch.int32, device="cuda")
        score = torch.empty((nq, k), dtype=torch.float64, device="cuda")
        sc = torch.empty((nq, self.info()["size"]), dtype=torch.float64, device="cuda") if scan else None
        _lib.call("prnu_fpstore_query", ctx.handle, self.handle, ptrs.data_ptr(), nq, int(k), idx.data_ptr(),
                  score.data_ptr(), sc.data_ptr() if sc is not None else None,
                  torch.cuda.current_stream().cuda_stream)
        if sync:
            torch.cuda.current_stream().synchronize()
        return (idx, score, sc) if scan else (idx, score)

    def close(self) -> None:
        if self.handle:
            _lib.load().prnu_fpstore_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Request:
    image: np.ndarray               # uint8 (h, w)
    t_submit: float = 0.0
    t_done: float = 0.0
    idx: np.ndarray | None = None   # top-k slots
    score: np.ndarray | None = None  # exact NCC
    batch_size: int = 0
    error: BaseException | None = None   # set instead of idx/score when the batch failed
    done: threading.Event = field(default_factory=threading.Event)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit

    def wait(self, timeout: float | None = None) -> bool:
        return self.done.wait(timeout)

    def result(self, timeout: float | None = None) -> tuple[np.ndarray, np.ndarray]:
        """(idx, score) once answered; raises the batch's error, or
        TimeoutError when no answer came within `timeout` seconds."""
        if not self.done.wait(timeout):
            raise TimeoutError("request not answered in time")
        if self.error is not None:
            raise self.error
        return self.idx, self.score


class DeviceArena:
    """First-fit allocator over one device buffer of ``size`` bytes: the HBM
    that in-flight batches draw their per-query buffers from.  Offsets are
    multiples of ``align``; freed ranges are merged with their neighbours.
    ``base`` is the buffer's device address (0 for bookkeeping alone)."""

    def __init__(self, size: int, align: int = 256, base: int = 0):
        self.size, self.align, self.base = int(size), int(align), int(base)
        self.free_ranges: list[tuple[int, int]] = [(0, self.size)] if self.size > 0 else []
        self.used = 0
        self.high_water = 0

    def _round(self, nbytes: int) -> int:
        return (int(nbytes) + self.align - 1) // self.align * self.align

    def largest_free(self) -> int:
        return max((n for _, n in self.free_ranges), default=0)

    def alloc(self, nbytes: int) -> int | None:
        """Offset of a fresh range of ``nbytes`` (rounded up), or None."""
        need = self._round(nbytes)
        for i, (off, n) in enumerate(self.free_ranges):
            if n >= need:
                if n == need:
                    del self.free_ranges[i]
                else:
                    self.free_ranges[i] = (off + need, n - need)
                self.used += need
                self.high_water = max(self.high_water, self.used)
                return off
        return None

    def free(self, off: int, nbytes: int) -> None:
        need = self._round(nbytes)
        if off < 0 or off + need > self.size or off % self.align:
            raise ValueError(f"range [{off}, {off + need}) is not an allocation of this arena")
        rs = self.free_ranges
        i = 0
        while i < len(rs) and rs[i][0] < off:
            i += 1
        if (i > 0 and rs[i - 1][0] + rs[i - 1][1] > off) or (i < len(rs) and off + need > rs[i][0]):
            raise ValueError(f"range [{off}, {off + need}) is already free")
        rs.insert(i, (off, need))
        if i + 1 < len(rs) and off + need == rs[i + 1][0]:        # merge with the next range
            rs[i] = (off, need + rs[i + 1][1])
            del rs[i + 1]
        if i > 0 and rs[i - 1][0] + rs[i - 1][1] == off:          # and with the previous one
            rs[i - 1] = (rs[i - 1][0], rs[i - 1][1] + rs[i][1])
            del rs[i]
        self.used -= need


def batch_bytes(n: int, size: int, k: int, m: int, align: int = 256) -> int:
    """Arena bytes of one batch of ``m`` queries of ``n`` pixels against a
    store of ``size`` fingerprints: u8 images, FP32 patterns, FP64 scan
    scores (m x size), int32 indices and FP64 scores of the top k, each
    buffer rounded up to ``align``."""
    r = lambda b: (int(b) + align - 1) // align * align   # noqa: E731
    return r(m * n) + r(4 * m * n) + r(8 * m * size) + r(4 * m * k) + r(8 * m * k)


def _new_stream() -> "torch.cuda.ExternalStream":
    """A HIP stream created by the library, wrapped for torch.  It is never
    destroyed: torch's pinned-memory allocator remembers the streams that used
    a pinned block (the staging copies run on it) and records events on them
    when the block is freed, possibly after the server is gone."""
    h = ctypes.c_void_p()
    _lib.call("prnu_stream_create", ctypes.byref(h))
    return torch.cuda.ExternalStream(h.value)


def _spin_together(streams, cycles: int) -> float:
    """Seconds for a bounded spin kernel (torch's clock-cycle sleep, one
    thread) on each of `streams`, launched together."""
    torch.cuda.synchronize()
    t = time.perf_counter()
    for st in streams:
        with torch.cuda.stream(st):
            torch.cuda._sleep(cycles)
    torch.cuda.synchronize()
    return time.perf_counter() - t


class _Slot:
    """One in-flight batch: its own context, stream, pinned host staging and
    device pointer table; the batch's device buffers come from the arena."""

    def __init__(self, height: int, width: int, max_batch: int, k: int):
        n = height * width
        self.ctx = _Ctx(height, width, _lib.MODE_NCC)
        # the slot's own HIP stream (Server._spread_streams checks that the
        # slots' streams run side by side)
        self.stream = _new_stream()
        self.d_ptrs = torch.zeros(max_batch, dtype=torch.int64, device="cuda")
        self.h_ptrs = torch.zeros(max_batch, dtype=torch.int64, pin_memory=True)
        self.h_images = torch.empty((max_batch, n), dtype=torch.uint8, pin_memory=True)
        self.h_idx = torch.empty((max_batch, k), dtype=torch.int32, pin_memory=True)
        self.h_score = torch.empty((max_batch, k), dtype=torch.float64, pin_memory=True)
        for buf in (self.h_images, self.h_idx, self.h_score):
            buf.zero_()                 # fault the pinned pages in now, not under the first batch
        self.event = torch.cuda.Event()  # recorded after the batch's result copies
        self.busy = False
        self.requests: list[Request] = []
        self.lease: tuple[int, int] | None = None     # (arena offset, bytes) of the batch in flight


_gc_lock = threading.Lock()
_gc_servers = 0                 # running servers that asked for gc.freeze


def _gc_hold() -> None:
    global _gc_servers
    with _gc_lock:
        if _gc_servers == 0:
            gc.collect()
            gc.freeze()
        _gc_servers += 1


def _gc_release() -> None:
    global _gc_servers
    with _gc_lock:
        _gc_servers -= 1
        if _gc_servers == 0:
            gc.unfreeze()


class Server:
    """Continuous-batching query server over a :class:`FingerprintStore`.

    Admission by free HBM: the per-query device buffers of every batch in
    flight (images, patterns, scan scores, results) are carved out of one
    arena of ``hbm_budget`` bytes (default: the device's free memory at
    construction less ``hbm_reserve`` and the slots' fixed workspaces).  A
    batch takes as many pending requests as ``max_batch``, a free slot and
    the arena's largest free range allow; when the arena is full, requests
    wait until a batch in flight completes and gives its range back.

    ``freeze_gc``: from ``start()`` to ``stop()`` the objects alive at start
    are moved out of the cyclic collector's reach (``gc.freeze``; shared by
    all running servers of the process, undone when the last one stops)."""

    def __init__(self, store: FingerprintStore, k: int = 10, max_batch: int = 64, slots: int = 2,
                 hbm_reserve: int = 2 << 30, freeze_gc: bool = True, hbm_budget: int | None = None,
                 copy_threads: int = 4, timeline: bool = False):
        self.store = store
        self.k = int(k)
        self.h, self.w = store.height, store.width
        n = self.h * self.w
        self.size = store.info()["size"]
        if self.size < self.k:
            raise ValueError(f"the store holds {self.size} fingerprints, fewer than k = {self.k}")
        self.per_query_bytes = batch_bytes(n, self.size, self.k, 1)
        # fixed per slot: extraction workspace (about 25 n per query of a
        # batch), FP8 query copies, scan partial tiles, FFT plans
        fixed = 256 << 20
        per_slot_ws = 27 * n
        free, _ = self.mem_info()
        fit = (free - int(hbm_reserve) - slots * fixed) // max(1, slots * (per_slot_ws + self.per_query_bytes))
        self.max_batch = int(max(1, min(max_batch, fit)))
        if hbm_budget is None:
            hbm_budget = free - int(hbm_reserve) - slots * (fixed + per_slot_ws * self.max_batch)
        budget = int(min(int(hbm_budget), slots * batch_bytes(n, self.size, self.k, self.max_batch)))
        if budget < batch_bytes(n, self.size, self.k, 1):
            raise ValueError(f"an HBM budget of {budget} bytes holds no batch of one query "
                             f"({batch_bytes(n, self.size, self.k, 1)} bytes)")
        self.hbm_reserve = int(hbm_reserve)
        self._arena_buf = torch.empty(budget, dtype=torch.uint8, device="cuda")
        self.arena = DeviceArena(budget, base=self._arena_buf.data_ptr())
        self.slots = [_Slot(self.h, self.w, self.max_batch, self.k) for _ in range(slots)]
        self._spread_streams()
        self.pending: "queue.Queue[Request]" = queue.Queue()
        self.inflight: "queue.Queue[_Slot | None]" = queue.Queue()
        self._stop = threading.Event()
        self._free = threading.Condition()
        self._sched = threading.Thread(target=self._schedule, daemon=True)
        self._done = threading.Thread(target=self._complete, daemon=True)
        self.batches: list[int] = []
        self.freeze_gc = bool(freeze_gc)
        self._frozen = False
        # host threads that copy a batch's images into the slot's pinned
        # staging (one thread copies about 6 GB/s: 10 ms for 64 frames of 1 MB,
        # longer than the batch's device time, so the next batch would not
        # be ready while the previous one runs)
        self._copy_threads = max(1, int(copy_threads))
        self._copy_pool = ThreadPoolExecutor(self._copy_threads) if self._copy_threads > 1 else None
        # timeline: HIP events around each batch's work on its slot's stream
        # (device-side spans; a profiler's kernel trace serialises the slots'
        # queues, so overlap is measured this way)
        self._timeline: list | None = [] if timeline else None

    def _spread_streams(self, cycles: int = 2_000_000, attempts: int = 8) -> None:
        """Give every slot a stream on its own hardware queue.  HIP maps
        streams onto a few hardware queues (GPU_MAX_HW_QUEUES, 4 by default),
        each to the least used one, so two streams created one after the
        other can share a queue; the slots' batches then run one after the
        other on the device.  Checked by timing: spin kernels on streams of
        different queues run together, on one queue they take the sum.  A
        slot whose stream shares a queue gets a new stream (up to `attempts`
        times; the old one is kept, unused)."""
        self.stream_check: dict = {"slots": len(self.slots)}
        if len(self.slots) < 2:
            return
        _spin_together([self.slots[0].stream], cycles)           # the spin kernel's first launch loads it
        alone = _spin_together([self.slots[0].stream], cycles)
        self.stream_check.update(alone_ms=1e3 * alone, together_ms=[], new_streams=0)
        for i in range(1, len(self.slots)):
            for _ in range(attempts):
                t = _spin_together([s.stream for s in self.slots[:i + 1]], cycles)
                self.stream_check["together_ms"].append(1e3 * t)
                if t < 1.5 * alone:
                    break
                self.slots[i].stream = _new_stream()
                self.stream_check["new_streams"] += 1

    @staticmethod
    def mem_info() -> tuple[int, int]:
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.call("prnu_mem_info", ctypes.byref(f), ctypes.byref(t))
        return f.value, t.value

    # ---------------------------------------------------------------- client
    def submit(self, image: np.ndarray) -> Request:
        """Queue one 8-bit image of the store's frame size (h*w values)."""
        img = np.ascontiguousarray(image, dtype=np.uint8)
        if img.size != self.h * self.w:
            raise ValueError(f"query image of {img.size} values, the store holds {self.w}x{self.h} fingerprints")
        if self._stop.is_set():
            raise RuntimeError("the server is stopped")
        r = Request(image=img)
        r.t_submit = time.perf_counter()
        self.pending.put(r)
        return r

    def _views(self, off: int, m: int):
        """Device buffers of a batch of m queries at arena offset `off`."""
        n = self.h * self.w
        a = self.arena._round
        buf = self._arena_buf
        o = off
        images = buf[o:o + m * n].view(m, n)
        o += a(m * n)
        patterns = buf[o:o + 4 * m * n].view(torch.float32).view(m, n)
        o += a(4 * m * n)
        scan = buf[o:o + 8 * m * self.size].view(torch.float64).view(m, self.size)
        o += a(8 * m * self.size)
        idx = buf[o:o + 4 * m * self.k].view(torch.int32).view(m, self.k)
        o += a(4 * m * self.k)
        score = buf[o:o + 8 * m * self.k].view(torch.float64).view(m, self.k)
        return images, patterns, scan, idx, score

    def _launch(self, slot: _Slot, m: int, off: int) -> None:
        """Enqueue one batch of m queries (images already in slot.h_images) on
        the slot's stream."""
        lib = _lib.load()
        n = self.h * self.w
        images, patterns, scan, idx, score = self._views(off, m)
        base = patterns.data_ptr()
        hp = slot.h_ptrs.numpy()
        hp[:m] = base + 4 * n * np.arange(m, dtype=np.int64)
        with torch.cuda.stream(slot.stream):
            st = slot.stream.cuda_stream
            if self._timeline is not None:
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                ev[0].record(slot.stream)
            slot.d_ptrs[:m].copy_(slot.h_ptrs[:m], non_blocking=True)
            images.copy_(slot.h_images[:m], non_blocking=True)
            _lib.check(lib.prnu_preprocess_batch(slot.ctx.handle, _lib.INPUT_GRAY8, images.data_ptr(), m * n, n, m,
                                                 patterns.data_ptr(), 4 * n * m, 4 * n, st), "prnu_preprocess_batch")
            _lib.check(lib.prnu_fpstore_query(slot.ctx.handle, self.store.handle, slot.d_ptrs.data_ptr(), m, self.k,
                                              idx.data_ptr(), score.data_ptr(), scan.data_ptr(), st),
                       "prnu_fpstore_query")
            slot.h_idx[:m].copy_(idx, non_blocking=True)
            slot.h_score[:m].copy_(score, non_blocking=True)
            slot.event.record(slot.stream)
            if self._timeline is not None:
                ev[1].record(slot.stream)
                self._timeline.append((self.slots.index(slot), m, ev))

    def timeline(self) -> list[tuple[int, int, float, float]]:
        """(slot, batch size, start ms, end ms) of every batch launched since
        start() with ``timeline=True``, in launch order, times relative to the
        first batch's start (device clock; waits for the batches to finish)."""
        if self._timeline is None:
            raise ValueError("the server was created without timeline=True")
        if not self._timeline:
            return []
        torch.cuda.synchronize()
        t0 = self._timeline[0][2][0]
        return [(s, m, t0.elapsed_time(a), t0.elapsed_time(b)) for s, m, (a, b) in self._timeline]

    def start(self) -> "Server":
        # size every slot's workspaces once with a full batch (a growing
        # workspace synchronises the device, which would serialise the
        # in-flight batches later); the arena lends the full batch's buffers
        m = self.max_batch
        for slot in self.slots:
            off = self.arena.alloc(batch_bytes(self.h * self.w, self.size, self.k, m))
            if off is None:             # a budget below one full batch: borrow temporary buffers
                self._warm_outside_arena(slot, m)
                continue
            slot.h_images[:m].zero_()
            self._launch(slot, m, off)
            slot.stream.synchronize()
            self.arena.free(off, batch_bytes(self.h * self.w, self.size, self.k, m))
        self.arena.high_water = 0
        if self._timeline is not None:
            self._timeline.clear()
        if self.freeze_gc:
            _gc_hold()
            self._frozen = True
        self._sched.start()
        self._done.start()
        return self

    def _warm_outside_arena(self, slot: _Slot, m: int) -> None:
        saved_buf, saved_arena = self._arena_buf, self.arena
        try:
            self._arena_buf = torch.empty(batch_bytes(self.h * self.w, self.size, self.k, m), dtype=torch.uint8,
                                          device="cuda")
            slot.h_images[:m].zero_()
            self._launch(slot, m, 0)
            slot.stream.synchronize()
        finally:
            self._arena_buf, self.arena = saved_buf, saved_arena

    def stop(self) -> None:
        """Stop admitting; batches in flight complete, requests still pending
        are failed with RuntimeError."""
        self._stop.set()
        if self._sched.is_alive():
            self._sched.join()
        self.inflight.put(None)
        if self._done.is_alive():
            self._done.join()
        err = RuntimeError("the server stopped before this request was admitted")
        while True:
            try:
                r = self.pending.get_nowait()
            except queue.Empty:
                break
            self._fail([r], err)
        if self._copy_pool is not None:
            self._copy_pool.shutdown(wait=True)
        if self._frozen:
            self._frozen = False
            _gc_release()

    @staticmethod
    def _fail(requests: list[Request], err: BaseException) -> None:
        t = time.perf_counter()
        for r in requests:
            r.error, r.t_done = err, t
            r.done.set()

    # ------------------------------------------------------------- scheduler
    def _admit(self) -> tuple[list[Request], int]:
        """Up to max_batch pending requests that fit the arena, and the
        batch's arena offset (([], -1) when nothing is pending)."""
        try:
            first = self.pending.get(timeout=0.002)
        except queue.Empty:
            return [], -1
        n = self.h * self.w
        with self._free:
            while True:
                fit = 0
                room = self.arena.largest_free()
                while fit < self.max_batch and batch_bytes(n, self.size, self.k, fit + 1) <= room:
                    fit += 1
                if fit >= 1 or self._stop.is_set():
                    break
                self._free.wait(timeout=0.01)          # the arena is full: wait for a batch to finish
            if fit < 1:
                self.pending.put(first)
                return [], -1
            batch = [first]
            while len(batch) < fit:
                try:
                    batch.append(self.pending.get_nowait())
                except queue.Empty:
                    break
            off = self.arena.alloc(batch_bytes(n, self.size, self.k, len(batch)))
        return batch, off

    def _free_slot(self) -> _Slot | None:
        with self._free:
            while not self._stop.is_set():
                for s in self.slots:
                    if not s.busy:
                        return s
                self._free.wait(timeout=0.01)
        return None

    def _release(self, slot: _Slot) -> None:
        with self._free:
            if slot.lease is not None:
                self.arena.free(*slot.lease)
                slot.lease = None
            slot.busy = False
            slot.requests = []
            self._free.notify_all()

    def _schedule(self) -> None:
        n = self.h * self.w
        while not self._stop.is_set():
            slot = self._free_slot()
            if slot is None:
                return
            batch, off = self._admit()
            if not batch:
                continue
            m = len(batch)
            slot.busy = True
            slot.requests = batch
            slot.lease = (off, batch_bytes(n, self.size, self.k, m))
            try:
                self._stage(slot, batch)
                self._launch(slot, m, off)
            except Exception as e:          # a failed batch answers its requests with the error
                try:
                    slot.stream.synchronize()   # nothing of the batch may still use its lease
                except Exception:
                    pass
                self._fail(batch, e)
                self._release(slot)
                continue
            self.batches.append(m)
            self.inflight.put(slot)

    def _stage(self, slot: _Slot, batch: list[Request]) -> None:
        """Copy the batch's images into the slot's pinned staging, split over
        the copy threads (numpy copies release the GIL)."""
        dst = slot.h_images.numpy()

        def copy(lo: int, hi: int) -> None:
            for i in range(lo, hi):
                dst[i] = batch[i].image.reshape(-1)

        m = len(batch)
        if self._copy_pool is None or m < 2 * self._copy_threads:
            copy(0, m)
            return
        step = (m + self._copy_threads - 1) // self._copy_threads
        for f in [self._copy_pool.submit(copy, lo, min(m, lo + step)) for lo in range(0, m, step)]:
            f.result()

    def _complete(self) -> None:
        while True:
            slot = self.inflight.get()
            if slot is None:
                return
            try:
                # poll rather than block: a blocking wait in the HIP runtime
                # held up the scheduler thread's launches on the other slot
                # until it returned (measured: 13.6k against 18.0k queries/s
                # at saturation on a 4,000-fingerprint store)
                while not slot.event.query():
                    time.sleep(2e-4)
                t = time.perf_counter()
                m = len(slot.requests)
                idx = slot.h_idx[:m].numpy().copy()
                score = slot.h_score[:m].numpy().copy()
                for i, r in enumerate(slot.requests):
                    r.idx, r.score, r.t_done, r.batch_size = idx[i], score[i], t, m
                    r.done.set()
            except Exception as e:
                self._fail(slot.requests, e)
            self._release(slot)
