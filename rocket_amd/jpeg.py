"""JPEG frames to RGB8 on the GPU (``prnu_jpeg_*`` in include/prnu_api.h).

The counterpart of the reference's ``ReadJPEG.readJPEG``
(``src/main/java/nl/esciencecenter/common_source_identification/ReadJPEG.java:25-45``,
``readjpeg/src/main/cpp/native.cpp:59-134``): libjpeg's default decode of a
3-component JPEG into packed RGB8, bit for bit, for a batch of files at
once, written straight into device memory where ``PRNUFilter`` reads it.
"""
from __future__ import annotations

import ctypes

from . import _lib


def jpeg_size(data: bytes) -> tuple[int, int]:
    """(height, width) from the headers; raises for files the decoder rejects."""
    w, h = ctypes.c_int32(), ctypes.c_int32()
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    _lib.call("prnu_jpeg_info", buf, len(data), ctypes.byref(w), ctypes.byref(h))
    return h.value, w.value


class JpegDecoder:
    """Owns the pinned staging and device buffers of batched decodes."""

    def __init__(self):
        h = ctypes.c_void_p()
        _lib.call("prnu_jpeg_decoder_create", ctypes.byref(h))
        self.handle = h

    def decode(self, files, out=None, stream=None, strict: bool = True):
        """Decode a list of JPEG files (bytes).  Returns (frames, status):
        frames a (count, H, W, 3) uint8 CUDA tensor when every accepted file
        has one size (or the given `out`, file k at out[k]), status a list of
        per-file codes (0, JPEG_CORRUPT_DATA, or the prnu_status a file was
        rejected with).  With `strict`, raises when any file is rejected
        (after decoding the others)."""
        import torch
        files = [bytes(f) for f in files]
        n = len(files)
        if n == 0:
            raise ValueError("no files")
        if out is None:
            sizes = set()
            for f in files:
                try:
                    sizes.add(jpeg_size(f))
                except _lib.PrnuError:
                    pass
            if not sizes:
                raise ValueError("no decodable file")
            if len(sizes) != 1:
                raise ValueError("files of different sizes: pass `out` with room for the largest")
            (hh, ww), = sizes
            out = torch.empty((n, hh, ww, 3), dtype=torch.uint8, device="cuda")
        if not out.is_cuda or out.dtype != torch.uint8 or not out.is_contiguous():
            raise ValueError("out must be a contiguous uint8 CUDA tensor")
        # the files' own buffers (read only; `files` keeps them alive)
        ptrs = (ctypes.c_void_p * n)(*[ctypes.cast(ctypes.c_char_p(f), ctypes.c_void_p).value for f in files])
        lens = (ctypes.c_int64 * n)(*[len(f) for f in files])
        ws, hs, st = (ctypes.c_int32 * n)(), (ctypes.c_int32 * n)(), (ctypes.c_int32 * n)()
        stride = out.numel() // n
        rc = _lib.load().prnu_jpeg_decode_batch(self.handle, ptrs, lens, n, out.data_ptr(), stride, out.numel(),
                                                ws, hs, st, stream or torch.cuda.current_stream().cuda_stream)
        if strict or rc not in (_lib.OK, _lib.EINVAL, _lib.ENOTSUP):
            _lib.check(rc, "prnu_jpeg_decode_batch")
        return out, list(st)

    def close(self) -> None:
        if self.handle:
            _lib.load().prnu_jpeg_decoder_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class JpegDecoderPool:
    """Several decoders, each with its own stream on its own worker thread, so
    that one call's host work (parsing, unstuffing, the copy, the checks
    between linking rounds) overlaps the kernels of the others.  ``map``
    decodes a sequence of batches and yields (frames, status) in order; each
    batch is complete on the device when it is yielded (every decode call is
    synchronous on its stream)."""

    def __init__(self, workers: int = 3):
        import threading
        from concurrent.futures import ThreadPoolExecutor
        if workers < 1:
            raise ValueError("workers must be positive")
        self._local = threading.local()
        self._all = []
        self._lock = threading.Lock()
        self._pool = ThreadPoolExecutor(max_workers=workers)

    def _worker_state(self):
        import torch
        st = getattr(self._local, "state", None)
        if st is None:
            st = (JpegDecoder(), torch.cuda.Stream())
            self._local.state = st
            with self._lock:
                self._all.append(st[0])
        return st

    def _decode(self, files, shape, strict):
        import torch
        dec, stream = self._worker_state()
        with torch.cuda.stream(stream):
            out = None
            if shape is not None:
                out = torch.empty((len(files),) + tuple(shape), dtype=torch.uint8, device="cuda")
            return dec.decode(files, out=out, stream=stream.cuda_stream, strict=strict)

    def map(self, batches, shape=None, strict: bool = True, ahead: int = 3):
        """Decode each batch (a list of file bytes) of `batches`; `shape`, when
        given, is the (H, W, 3) of every frame (else taken from the files, as
        JpegDecoder.decode does).  At most `ahead` batches are in flight.
        Frames are allocated on the worker's stream: a consumer on another
        stream calls ``frames.record_stream(its_stream)`` before using them."""
        from collections import deque
        pending = deque()
        for files in batches:
            pending.append(self._pool.submit(self._decode, list(files), shape, strict))
            if len(pending) >= ahead:
                yield pending.popleft().result()
        while pending:
            yield pending.popleft().result()

    def close(self) -> None:
        self._pool.shutdown(wait=True)
        with self._lock:
            for dec in self._all:
                dec.close()
            self._all = []
