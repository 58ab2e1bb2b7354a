"""Pattern spill tier: pattern sets larger than HBM.

The role of Rocket's host and file caches (``R/cache/HostCache.java:12-67``,
``R/cache/FileCache.java:18-130``, used by ``R/scheduler/HostWorker.java``):
what does not fit on the device waits in pinned host memory and, if asked, in
files.  This tier holds patterns as store *slot images*
(``prnu_store_save``): the K-blocked layout of a pattern store with the
per-slot FP64 statistics, so that a block of patterns comes back to the GPU
with one copy, without extraction or format conversion (an FP16 image of a
2048x1536 pattern is 6.3 MB, less than the 9.4 MB RGB frame it came from).

:meth:`HostPatternTier.ncc` computes the all-pairs NCC of the whole set on a
device store of a few blocks: R row blocks stay resident while every later
block streams through two column regions, the next one copied on a copy
stream while the current one is compared on the compute stream.  Each block
pair is one rectangle (or, on the diagonal, one triangle) of a grouped Gram
plan; with the same K slabs the result is bitwise the resident store's.
"""
from __future__ import annotations

import os

import numpy as np

from . import csi

# slot-image layout (prnu_store_save): a 64-byte header, then per K block
# `count` blocks of 128 bytes, then per slot FP64 sum and sum of squares and
# FP32 max|x| and scale
SLOT_HEADER_BYTES = 64
SLOT_BLOCK_BYTES = 128
SLOT_STATS_BYTES = 2 * 8 + 2 * 4
SLOT_BLOCK_VALUES = {"split_bf16": 32, "fp32": 32, "fp16": 64, "fp8": 128}


class HostPatternTier:
    """``count`` patterns of ``n`` values, in blocks of ``block`` slots, as
    slot images in pinned host memory (``format`` as PatternStore)."""

    def __init__(self, n: int, count: int, block: int = 1024, format: str = "fp16"):
        if count < 1 or block < 1:
            raise ValueError("count and block must be positive")
        self.n, self.count, self.block, self.format = int(n), int(count), int(block), format
        self.nblocks = (self.count + self.block - 1) // self.block
        self.images = [None] * self.nblocks
        self._stage = None

    def block_range(self, b: int) -> tuple[int, int]:
        return b * self.block, min(self.count, (b + 1) * self.block)

    def put_block(self, b: int, patterns) -> None:
        """FP32 CUDA patterns (rows of n values) of block b into its host image
        (statistics computed on the way, as PatternStore.put does)."""
        import torch
        r0, r1 = self.block_range(b)
        if patterns.shape[0] != r1 - r0:
            raise ValueError(f"block {b} holds {r1 - r0} patterns, got {patterns.shape[0]}")
        if self._stage is None:
            self._stage = csi.PatternStore(self.n, self.block, self.format)
        self._stage.put(patterns, 0)
        img = torch.empty(self._stage.image_bytes(r1 - r0), dtype=torch.uint8, pin_memory=True)
        self._stage.save(0, r1 - r0, img)
        self.images[b] = img

    def save_files(self, directory: str) -> list[str]:
        """One file per block (the FileCache role); the directory is created."""
        os.makedirs(directory, exist_ok=True)
        paths = []
        for b, img in enumerate(self._all_images()):
            path = os.path.join(directory, f"block-{b:05d}.slots")
            img.numpy().tofile(path)
            paths.append(path)
        return paths

    def load_files(self, directory: str) -> None:
        """Blocks written by save_files back into pinned host memory.  Each
        file must be a whole slot image of its block: the header's pattern
        length, format and count must match this tier and the block, and the
        file must hold exactly the bytes that header describes (a truncated
        or padded file raises ValueError before anything reaches a store)."""
        import torch
        paths = [os.path.join(directory, f"block-{b:05d}.slots") for b in range(self.nblocks)]
        for b, path in enumerate(paths):            # every file checked before any is read whole
            with open(path, "rb") as fh:
                head = np.frombuffer(fh.read(SLOT_HEADER_BYTES), dtype=np.uint8)
            r0, r1 = self.block_range(b)
            self._check_image(head, r1 - r0, path, size=os.path.getsize(path))
        for b, path in enumerate(paths):
            data = np.fromfile(path, dtype=np.uint8)
            r0, r1 = self.block_range(b)
            self._check_image(data, r1 - r0, path)
            img = torch.empty(data.size, dtype=torch.uint8, pin_memory=True)
            img.numpy()[:] = data
            self.images[b] = img

    def _check_image(self, data: np.ndarray, count: int, what: str, size: int | None = None) -> None:
        """The slot-image header (magic "PRNUSLOT", int64 n, int32 format,
        int32 count, int64 K blocks, 32 reserved bytes) against this tier;
        ``size`` is the image's length in bytes (default: ``data.size``)."""
        from . import _lib
        size = data.size if size is None else int(size)
        if data.size < SLOT_HEADER_BYTES or bytes(data[:8]) != b"PRNUSLOT":
            raise ValueError(f"{what}: not a slot image")
        n = int(data[8:16].view(np.int64)[0])
        fmt, cnt = (int(v) for v in data[16:24].view(np.int32))
        nkb = int(data[24:32].view(np.int64)[0])
        per_block = SLOT_BLOCK_VALUES[self.format]
        if n != self.n or fmt != _lib.PRECISIONS[self.format] or nkb != (self.n + per_block - 1) // per_block:
            raise ValueError(f"{what}: image of {n} values in format {fmt}, tier of {self.n} in {self.format}")
        if cnt != count:
            raise ValueError(f"{what}: image of {cnt} slots for a block of {count}")
        want = SLOT_HEADER_BYTES + count * (nkb * SLOT_BLOCK_BYTES + SLOT_STATS_BYTES)
        if size != want:
            raise ValueError(f"{what}: {size} bytes, a slot image of {count} slots has {want}")

    def host_bytes(self) -> int:
        return sum(int(img.numel()) for img in self._all_images())

    def _all_images(self):
        missing = [b for b, img in enumerate(self.images) if img is None]
        if missing:
            raise ValueError(f"blocks {missing} were never filled")
        return self.images

    def ncc(self, gram: "csi.GramNCC", slab_values: int = 0, device_blocks: int = 4,
            stats: dict | None = None) -> np.ndarray:
        """All-pairs NCC of the tier as a (count, count) float64 matrix
        (symmetric, the diagonal as the store Gram writes it).

        The device store holds `device_blocks` blocks: R = device_blocks - 2
        row blocks stay resident while every later block streams through two
        column regions, so the tier is copied in about P + P^2 / (2R) block
        loads for P blocks.  `stats`, when given, receives the bytes copied to
        the device, the block loads and the block pairs."""
        import torch
        if device_blocks < 3:
            raise ValueError("device_blocks must be at least 3 (a row block and two column regions)")
        images = self._all_images()
        B, P = self.block, self.nblocks
        R = min(device_blocks - 2, P)
        dev = csi.PatternStore(self.n, (R + 2) * B, self.format)
        region = [r * B for r in range(R + 2)]      # rows: 0..R-1; columns: R, R+1
        compute, copy = torch.cuda.Stream(), torch.cuda.Stream()
        loaded = [torch.cuda.Event() for _ in region]
        free = [torch.cuda.Event() for _ in region]
        nbuf = 3
        dbuf = [torch.empty((B, B), dtype=torch.float64, device="cuda") for _ in range(nbuf)]
        hbuf = [torch.empty(B * B, dtype=torch.float64, pin_memory=True) for _ in range(nbuf)]
        done = [torch.cuda.Event() for _ in range(nbuf)]
        pending = [None] * nbuf
        result = np.zeros((self.count, self.count), dtype=np.float64)
        plans = {}
        moved = 0
        loads = 0
        k = 0

        def drain(j):
            if pending[j] is not None:
                done[j].synchronize()
                r0, c0, rn, cn = pending[j]
                result[r0:r0 + rn, c0:c0 + cn] = hbuf[j][:rn * cn].view(rn, cn).numpy()
                pending[j] = None

        def load(b, reg):
            nonlocal moved, loads
            copy.wait_event(free[reg])
            with torch.cuda.stream(copy):
                dev.load(region[reg], images[b], sync=False)
            loaded[reg].record(copy)
            moved += int(images[b].numel())
            loads += 1

        def compare(rreg, rb, creg, cb):
            nonlocal k
            r0, r1 = self.block_range(rb)
            c0, c1 = self.block_range(cb)
            rn, cn = r1 - r0, c1 - c0
            key = (rreg, rn, creg, cn, k)
            drain(k)
            compute.wait_event(loaded[rreg])
            compute.wait_event(loaded[creg])
            with torch.cuda.stream(compute):
                if key not in plans:
                    plans[key] = csi.GramPlan(gram, blocks=[(dev, region[rreg], rn, region[creg], cn,
                                                             dbuf[k][:rn, :cn])], slab_values=slab_values)
                plans[key].run(use_graph=True, sync=False)
                hbuf[k][:rn * cn].view(rn, cn).copy_(dbuf[k][:rn, :cn], non_blocking=True)
            done[k].record(compute)
            pending[k] = (r0, c0, rn, cn)
            k = (k + 1) % nbuf

        col = 0
        for g0 in range(0, P, R):
            group = list(range(g0, min(P, g0 + R)))
            for t, i in enumerate(group):
                load(i, t)
            for t, i in enumerate(group):
                for u in range(t, len(group)):
                    compare(t, i, u, group[u])          # t == u: the block's triangle
            for j in range(group[-1] + 1, P):
                creg = R + col % 2
                col += 1
                load(j, creg)
                for t, i in enumerate(group):
                    compare(t, i, creg, j)
                free[creg].record(compute)
            for t in range(len(group)):
                free[t].record(compute)
        for j in range(nbuf):
            drain(j)
        torch.cuda.synchronize()
        for plan in plans.values():
            plan.close()
        dev.close()
        il = np.tril_indices(self.count, -1)
        result[il] = result.T[il]
        if stats is not None:
            stats.update({"bytes_to_device": moved, "block_loads": loads, "block_pairs": P * (P + 1) // 2,
                          "row_blocks_resident": R, "plans": len(plans)})
        return result

    def close(self) -> None:
        if self._stage is not None:
            self._stage.close()
            self._stage = None
        self.images = [None] * self.nblocks
