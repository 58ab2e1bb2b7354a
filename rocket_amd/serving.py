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
    done: threading.Event = field(default_factory=threading.Event)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit

    def wait(self, timeout: float | None = None) -> bool:
        return self.done.wait(timeout)


class _Slot:
    """One in-flight batch: its own context, stream and buffers."""

    def __init__(self, height: int, width: int, max_batch: int, k: int):
        n = height * width
        self.ctx = _Ctx(height, width, _lib.MODE_NCC)
        self.stream = torch.cuda.Stream()
        self.images = torch.empty((max_batch, n), dtype=torch.uint8, device="cuda")
        self.patterns = torch.empty((max_batch, n), dtype=torch.float32, device="cuda")
        self.ptrs = pattern_pointer_table(self.patterns)
        self.idx = torch.empty((max_batch, k), dtype=torch.int32, device="cuda")
        self.score = torch.empty((max_batch, k), dtype=torch.float64, device="cuda")
        self.h_images = torch.empty((max_batch, n), dtype=torch.uint8, pin_memory=True)
        self.h_idx = torch.empty((max_batch, k), dtype=torch.int32, pin_memory=True)
        self.h_score = torch.empty((max_batch, k), dtype=torch.float64, pin_memory=True)
        for buf in (self.h_images, self.h_idx, self.h_score):
            buf.zero_()                 # fault the pinned pages in now, not under the first batch
        self.event = torch.cuda.Event()
        self.busy = False
        self.requests: list[Request] = []


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

    ``freeze_gc``: from ``start()`` to ``stop()`` the objects alive at start
    are moved out of the cyclic collector's reach (``gc.freeze``; shared by
    all running servers of the process, undone when the last one stops)."""

    def __init__(self, store: FingerprintStore, k: int = 10, max_batch: int = 64, slots: int = 2,
                 hbm_reserve: int = 2 << 30, freeze_gc: bool = True):
        self.store = store
        self.k = int(k)
        self.h, self.w = store.height, store.width
        n = self.h * self.w
        # admission by free HBM: every slot holds max_batch images, patterns,
        # results and the scan workspace (FP8 queries, nq x size scores)
        free, _ = self.mem_info()
        size = max(1, store.info()["size"])
        # image, pattern, FP8 copy, scan scores, results, extraction workspace
        per_query = n + 4 * n + n + 8 * size + 64 * self.k + 25 * n
        fixed = 256 << 20                   # scan partial tiles, FFT plans
        fit = (free - hbm_reserve - slots * fixed) // max(1, slots * per_query)
        self.max_batch = int(max(1, min(max_batch, fit)))
        self.hbm_reserve = int(hbm_reserve)
        self.per_query_bytes = int(per_query)
        self.slots = [_Slot(self.h, self.w, self.max_batch, self.k) for _ in range(slots)]
        self.pending: "queue.Queue[Request]" = queue.Queue()
        self.inflight: "queue.Queue[_Slot | None]" = queue.Queue()
        self._stop = threading.Event()
        self._free = threading.Condition()
        self._sched = threading.Thread(target=self._schedule, daemon=True)
        self._done = threading.Thread(target=self._complete, daemon=True)
        self.batches: list[int] = []
        self.freeze_gc = bool(freeze_gc)
        self._frozen = False

    @staticmethod
    def mem_info() -> tuple[int, int]:
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.call("prnu_mem_info", ctypes.byref(f), ctypes.byref(t))
        return f.value, t.value

    # ---------------------------------------------------------------- client
    def submit(self, image: np.ndarray) -> Request:
        r = Request(image=np.ascontiguousarray(image, dtype=np.uint8))
        r.t_submit = time.perf_counter()
        self.pending.put(r)
        return r

    def start(self) -> "Server":
        # size every slot's workspaces once (a growing workspace synchronises
        # the device, which would serialise the in-flight batches later)
        lib = _lib.load()
        n = self.h * self.w
        for slot in self.slots:
            with torch.cuda.stream(slot.stream):
                st = slot.stream.cuda_stream
                slot.images.zero_()
                _lib.check(lib.prnu_preprocess_batch(slot.ctx.handle, _lib.INPUT_GRAY8, slot.images.data_ptr(),
                                                     self.max_batch * n, n, self.max_batch,
                                                     slot.patterns.data_ptr(), 4 * n * self.max_batch, 4 * n, st))
                slot.patterns.normal_()
                _lib.check(lib.prnu_fpstore_query(slot.ctx.handle, self.store.handle, slot.ptrs.data_ptr(),
                                                  self.max_batch, self.k, slot.idx.data_ptr(),
                                                  slot.score.data_ptr(), None, st))
            slot.stream.synchronize()
        if self.freeze_gc:
            _gc_hold()
            self._frozen = True
        self._sched.start()
        self._done.start()
        return self

    def stop(self) -> None:
        self._stop.set()
        self._sched.join()
        self.inflight.put(None)
        self._done.join()
        if self._frozen:
            self._frozen = False
            _gc_release()

    # ------------------------------------------------------------- scheduler
    def _admit(self) -> list[Request]:
        """Up to max_batch pending requests, fewer if free HBM has shrunk."""
        try:
            first = self.pending.get(timeout=0.002)
        except queue.Empty:
            return []
        batch = [first]
        free, _ = self.mem_info()
        room = max(1, (free - self.hbm_reserve) // self.per_query_bytes)
        limit = int(min(self.max_batch, room))
        while len(batch) < limit:
            try:
                batch.append(self.pending.get_nowait())
            except queue.Empty:
                break
        return batch

    def _free_slot(self) -> _Slot | None:
        with self._free:
            while not self._stop.is_set():
                for s in self.slots:
                    if not s.busy:
                        return s
                self._free.wait(timeout=0.01)
        return None

    def _schedule(self) -> None:
        lib = _lib.load()
        n = self.h * self.w
        while not self._stop.is_set():
            slot = self._free_slot()
            if slot is None:
                return
            batch = self._admit()
            if not batch:
                continue
            m = len(batch)
            for i, r in enumerate(batch):
                slot.h_images[i].numpy()[:] = r.image.reshape(-1)
            slot.busy = True
            slot.requests = batch
            self.batches.append(m)
            with torch.cuda.stream(slot.stream):
                st = slot.stream.cuda_stream
                slot.images[:m].copy_(slot.h_images[:m], non_blocking=True)
                _lib.check(lib.prnu_preprocess_batch(slot.ctx.handle, _lib.INPUT_GRAY8, slot.images.data_ptr(), m * n,
                                                     n, m, slot.patterns.data_ptr(), 4 * n * m, 4 * n, st),
                           "prnu_preprocess_batch")
                _lib.check(lib.prnu_fpstore_query(slot.ctx.handle, self.store.handle, slot.ptrs.data_ptr(), m,
                                                  self.k, slot.idx.data_ptr(), slot.score.data_ptr(), None, st),
                           "prnu_fpstore_query")
                slot.h_idx[:m].copy_(slot.idx[:m], non_blocking=True)
                slot.h_score[:m].copy_(slot.score[:m], non_blocking=True)
                slot.event.record(slot.stream)
            self.inflight.put(slot)

    def _complete(self) -> None:
        while True:
            slot = self.inflight.get()
            if slot is None:
                return
            slot.event.synchronize()
            t = time.perf_counter()
            m = len(slot.requests)
            idx = slot.h_idx[:m].numpy().copy()
            score = slot.h_score[:m].numpy().copy()
            for i, r in enumerate(slot.requests):
                r.idx, r.score, r.t_done, r.batch_size = idx[i], score[i], t, m
                r.done.set()
            with self._free:
                slot.busy = False
                slot.requests = []
                self._free.notify_all()
