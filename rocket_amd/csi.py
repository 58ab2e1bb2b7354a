"""Host-side mirror of the reference's CSI plugin for this path, over the C ABI.

Names, argument meaning and error behaviour follow the reference's Java
classes (paths relative to
``src/main/java/nl/esciencecenter/common_source_identification/``):

* :class:`CommonSourceIdentificationContext` -- ``CommonSourceIdentificationContext.java:20-127``
  (``ApplicationContext<ImageIdentifier, Double>``, ``R/scheduler/ApplicationContext.java:11-99``)
* :class:`PRNUFilter` -- ``kernels/filter/PRNUFilter.java:41-194``
* :class:`NormalizedCrossCorrelation`, :class:`PeakToCorrelationEnergy` --
  ``kernels/compare/*.java`` behind :class:`PatternComparator`
  (``kernels/compare/PatternComparator.java:34-54``)
* :class:`GramNCC` -- the batched all-pairs extension (no reference counterpart:
  it replaces M(M-1)/2 ``correlateGPU`` calls).

Device buffers are torch CUDA tensors (PyTorch is only used for device memory
and streams here); every call runs on torch's current stream and, like the
reference's GPU methods, returns after that stream is synchronised unless
``sync=False``.  Undersized buffers raise :class:`ValueError`
(``IllegalArgumentException`` in the reference).
"""
from __future__ import annotations

import ctypes
import enum
import struct

from . import _lib

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


class ComparisonStrategy(enum.Enum):
    """``CommonSourceIdentificationContext.ComparisonStrategy`` (``:23-26``)."""
    PCE = _lib.MODE_PCE
    NCC = _lib.MODE_NCC


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _mem(t) -> tuple[int, int]:
    """(device address, size in bytes) of a CUDA tensor (the CudaMem analogue)."""
    if not (hasattr(t, "is_cuda") and t.is_cuda):
        raise ValueError("device buffer must be a CUDA tensor")
    if not t.is_contiguous():
        raise ValueError("device buffer must be contiguous")
    return t.data_ptr(), t.numel() * t.element_size()


def _sync(sync: bool) -> None:
    if sync:
        torch.cuda.current_stream().synchronize()


class _Ctx:
    """Owns one ``prnu_ctx``."""

    def __init__(self, height: int, width: int, mode: int):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.prnu_ctx_create(int(height), int(width), int(mode), ctypes.byref(h)), "prnu_ctx_create")
        self.handle = h
        self.height = int(height)
        self.width = int(width)
        self.mode = int(mode)

    def close(self) -> None:
        if self.handle:
            _lib.load().prnu_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ filters
class _FilterStage:
    def __init__(self, ctx: _Ctx):
        self._ctx = ctx

    def cleanup(self) -> None:
        pass


class GrayscaleFilter(_FilterStage):
    """``kernels/filter/GrayscaleFilter.java:72-85``."""

    def apply_gpu(self, image, output, input_format: int = _lib.INPUT_RGB8, sync: bool = True) -> None:
        ip, ib = _mem(image)
        op, ob = _mem(output)
        if ob < 4 * self._ctx.height * self._ctx.width:
            raise ValueError("output smaller than h*w floats")
        _lib.call("prnu_filter_grayscale", self._ctx.handle, input_format, ip, ib, op, _stream())
        _sync(sync)


class FastNoiseFilter(_FilterStage):
    """``kernels/filter/FastNoiseFilter.java:98-120`` (in place)."""

    def apply_gpu(self, image, sync: bool = True) -> None:
        p, b = _mem(image)
        if b < 4 * self._ctx.height * self._ctx.width:
            raise ValueError("image smaller than h*w floats")
        _lib.call("prnu_filter_fastnoise", self._ctx.handle, p, _stream())
        _sync(sync)


class ZeroMeanTotalFilter(_FilterStage):
    """``kernels/filter/ZeroMeanTotalFilter.java:77-99`` (in place)."""

    def apply_gpu(self, image, sync: bool = True) -> None:
        p, b = _mem(image)
        if b < 4 * self._ctx.height * self._ctx.width:
            raise ValueError("image smaller than h*w floats")
        _lib.call("prnu_filter_zeromean", self._ctx.handle, p, _stream())
        _sync(sync)


class WienerFilter(_FilterStage):
    """``kernels/filter/WienerFilter.java:187-240`` (in place)."""

    def apply_gpu(self, image, sync: bool = True) -> None:
        p, b = _mem(image)
        if b < 4 * self._ctx.height * self._ctx.width:
            raise ValueError("image smaller than h*w floats")
        _lib.call("prnu_filter_wiener", self._ctx.handle, p, _stream())
        _sync(sync)


class SpectralFilter(_FilterStage):
    """``kernels/filter/SpectralFilter.java:44-54``: R2C half spectrum."""

    def apply_gpu(self, image, output, sync: bool = True) -> None:
        ip, _ = _mem(image)
        op, ob = _mem(output)
        _lib.call("prnu_filter_spectral", self._ctx.handle, ip, op, ob, _stream())
        _sync(sync)


class PRNUFilter:
    """``kernels/filter/PRNUFilter.java:41-194``: gray -> FastNoise -> ZeroMean
    -> Wiener (-> R2C spectrum when ``apply_spectral_filter``)."""

    def __init__(self, height: int, width: int, apply_spectral_filter: bool = False, _ctx: _Ctx | None = None):
        self.h = int(height)
        self.w = int(width)
        self.apply_spectral_filter = bool(apply_spectral_filter)
        mode = _lib.MODE_PCE if apply_spectral_filter else _lib.MODE_NCC
        self._ctx = _ctx if _ctx is not None else _Ctx(height, width, mode)
        self._own = _ctx is None
        self.grayscale_filter = GrayscaleFilter(self._ctx)
        self.fast_noise_filter = FastNoiseFilter(self._ctx)
        self.zero_mean_total_filter = ZeroMeanTotalFilter(self._ctx)
        self.wiener_filter = WienerFilter(self._ctx)
        self.spectral_filter = SpectralFilter(self._ctx)

    def get_input_size(self) -> int:
        return self.h * self.w * 3

    def get_output_size(self) -> int:
        n = 2 * self.h * (self.w // 2 + 1) if self.apply_spectral_filter else self.h * self.w
        return 4 * n

    def apply_gpu(self, image, output, input_format: int = _lib.INPUT_RGB8, sync: bool = True) -> int:
        ip, ib = _mem(image)
        op, ob = _mem(output)
        out_bytes = ctypes.c_size_t()
        _lib.call("prnu_preprocess", self._ctx.handle, input_format, ip, ib, op, ob, ctypes.byref(out_bytes), _stream())
        _sync(sync)
        return out_bytes.value

    def apply_gpu_batch(self, images, outputs, input_format: int = _lib.INPUT_RGB8, sync: bool = True) -> None:
        """``apply_gpu`` of every row of ``images`` (B, ...) into the rows of
        ``outputs`` (B, ...) with one launch per filter stage
        (``prnu_preprocess_batch``); row strides are taken from the tensors."""
        count = int(images.shape[0])
        if int(outputs.shape[0]) < count:
            raise ValueError(f"outputs hold {outputs.shape[0]} patterns, need {count}")
        if count == 0:
            return
        for t in (images, outputs):
            if not (hasattr(t, "is_cuda") and t.is_cuda):
                raise ValueError("device buffer must be a CUDA tensor")
            if not t[0].is_contiguous():
                raise ValueError("every image / pattern row must be contiguous")
        ip, op = images.data_ptr(), outputs.data_ptr()
        in_stride = int(images.stride(0)) * images.element_size()
        out_stride = int(outputs.stride(0)) * outputs.element_size()
        ib = in_stride * (count - 1) + int(images[0].numel()) * images.element_size()
        ob = out_stride * (int(outputs.shape[0]) - 1) + int(outputs[0].numel()) * outputs.element_size()
        _lib.call("prnu_preprocess_batch", self._ctx.handle, input_format, ip, ib, in_stride, count, op, ob,
                  out_stride, _stream())
        _sync(sync)

    # getters, PRNUFilter.java:157-175
    def get_grayscale_filter(self):
        return self.grayscale_filter

    def get_fast_noise_filter(self):
        return self.fast_noise_filter

    def get_zero_mean_total_filter(self):
        return self.zero_mean_total_filter

    def get_wiener_filter(self):
        return self.wiener_filter

    def get_spectral_filter(self):
        return self.spectral_filter

    def cleanup(self) -> None:
        if self._own:
            self._ctx.close()


# -------------------------------------------------------------- comparators
class PatternComparator:
    """``kernels/compare/PatternComparator.java:34-54``."""

    def apply_gpu(self, left, right, result, sync: bool = True) -> None:
        raise NotImplementedError

    def cleanup(self) -> None:
        raise NotImplementedError


class NormalizedCrossCorrelation(PatternComparator):
    """``kernels/compare/NormalizedCrossCorrelation.java:18-145``."""

    def __init__(self, h: int, w: int, _ctx: _Ctx | None = None):
        self.h, self.w = int(h), int(w)
        self._ctx = _ctx if _ctx is not None else _Ctx(h, w, _lib.MODE_NCC)
        self._own = _ctx is None

    def get_input_size(self) -> int:
        return 4 * self.h * self.w

    def get_output_size(self) -> int:
        return 8

    def apply_gpu(self, left, right, result, sync: bool = True) -> None:
        lp, lb = _mem(left)
        rp, rb = _mem(right)
        op, ob = _mem(result)
        out = ctypes.c_size_t()
        _lib.call("prnu_correlate", self._ctx.handle, lp, lb, rp, rb, op, ob, ctypes.byref(out), _stream())
        _sync(sync)

    def cleanup(self) -> None:
        if self._own:
            self._ctx.close()


class PeakToCorrelationEnergy(PatternComparator):
    """``kernels/compare/PeakToCorrelationEnergy.java:36-263``.  ``variant``
    selects the reference semantics or the textbook PCE."""

    def __init__(self, h: int, w: int, variant: int = _lib.PCE_REFERENCE, _ctx: _Ctx | None = None):
        self.h, self.w = int(h), int(w)
        self.variant = int(variant)
        self._ctx = _ctx if _ctx is not None else _Ctx(h, w, _lib.MODE_PCE)
        self._own = _ctx is None

    def get_input_size(self) -> int:
        return 4 * 2 * self.h * (self.w // 2 + 1)

    def get_output_size(self) -> int:
        return 8

    def apply_gpu(self, left, right, result, sync: bool = True) -> None:
        lp, lb = _mem(left)
        rp, rb = _mem(right)
        op, ob = _mem(result)
        if lb < self.get_input_size() or rb < self.get_input_size() or ob < 8:
            raise ValueError("buffer smaller than the spectrum or the result")
        _lib.call("prnu_correlate_pce", self._ctx.handle, self.variant, lp, rp, op, _stream())
        _sync(sync)

    def compare(self, left, right, result, sync: bool = True) -> None:
        self.apply_gpu(left, right, result, sync)

    def pairs(self, spectra, pairs, out=None, sync: bool = True):
        """PCE of every (i, j) in ``pairs`` (int array (P, 2)) over a list of
        spectra (CUDA tensors) or a (M, ...) CUDA tensor, batched on the GPU."""
        ptrs = pattern_pointer_table(spectra)
        pr = torch.as_tensor(pairs, dtype=torch.int32).reshape(-1, 2).contiguous().cuda()
        n = int(pr.shape[0])
        if out is None:
            out = torch.empty(n, dtype=torch.float64, device="cuda")
        _lib.call("prnu_pce_pairs", self._ctx.handle, self.variant, ptrs.data_ptr(), pr.data_ptr(), n,
                  out.data_ptr(), _stream())
        _sync(sync)
        return out

    def cleanup(self) -> None:
        if self._own:
            self._ctx.close()


# ------------------------------------------------------------------ context
class CommonSourceIdentificationContext:
    """``CommonSourceIdentificationContext.java:20-127`` over the C ABI."""

    def __init__(self, height: int, width: int, comp: ComparisonStrategy = ComparisonStrategy.NCC):
        if not isinstance(comp, ComparisonStrategy):
            raise ValueError(f"unknown comparison strategy: {comp}")
        self.height, self.width = int(height), int(width)
        self.comp = comp
        self._ctx = _Ctx(height, width, comp.value)
        self.filter = PRNUFilter(height, width, comp is ComparisonStrategy.PCE, _ctx=self._ctx)
        if comp is ComparisonStrategy.PCE:
            self.comparator = PeakToCorrelationEnergy(height, width, _ctx=self._ctx)
        else:
            self.comparator = NormalizedCrossCorrelation(height, width, _ctx=self._ctx)

    def get_max_file_size(self) -> int:
        return int(_lib.load().prnu_max_file_size(self._ctx.handle))

    def get_max_input_size(self) -> int:
        return int(_lib.load().prnu_max_input_size(self._ctx.handle))

    def get_max_output_size(self) -> int:
        return int(_lib.load().prnu_max_output_size(self._ctx.handle))

    def preprocess_input_gpu(self, key, input, output, input_format: int = _lib.INPUT_RGB8) -> int:
        """``:105-108``; returns the bytes of the pattern written."""
        self.filter.apply_gpu(input, output, input_format)
        return self.filter.get_output_size()

    def correlate_gpu(self, left_key, left, right_key, right, output) -> int:
        """``:111-115``; returns the bytes to copy back (8)."""
        lp, lb = _mem(left)
        rp, rb = _mem(right)
        op, ob = _mem(output)
        out = ctypes.c_size_t()
        _lib.call("prnu_correlate", self._ctx.handle, lp, lb, rp, rb, op, ob, ctypes.byref(out), _stream())
        _sync(True)
        return out.value

    def postprocess_output(self, left_key, right_key, buffer: bytes) -> float:
        """``:118-120``: the first FP64 of the output (native byte order)."""
        return struct.unpack_from("=d", buffer, 0)[0]

    def destroy(self) -> None:
        self._ctx.close()


# ------------------------------------------------------------------ batched
def pattern_pointer_table(patterns) -> "torch.Tensor":
    """Device int64 array of the row addresses of a (M, ...) CUDA tensor or a
    list of CUDA tensors."""
    if isinstance(patterns, (list, tuple)):
        addrs = [_mem(p)[0] for p in patterns]
    else:
        base, _ = _mem(patterns)
        row = patterns[0].numel() * patterns.element_size()
        addrs = [base + i * row for i in range(patterns.shape[0])]
    return torch.tensor(addrs, dtype=torch.int64, device="cuda")


class GramNCC:
    """All-pairs NCC of M patterns of one size as a triangular Gram on MFMA.

    ``precision``: "fp32" (FP32 MFMA), "split_bf16" (hi+lo BF16, three
    products) or "fp8" (e4m3, per-pattern scale).  The statistics (sum x,
    sum x^2 in FP64) are computed once per pattern set."""

    def __init__(self, height: int, width: int):
        self.height, self.width = int(height), int(width)
        self._ctx = _Ctx(height, width, _lib.MODE_NCC)

    def stats(self, ptrs, count: int, sync: bool = True):
        n = self.height * self.width
        s = torch.empty(count, dtype=torch.float64, device="cuda")
        ss = torch.empty(count, dtype=torch.float64, device="cuda")
        am = torch.empty(count, dtype=torch.float32, device="cuda")
        _lib.call("prnu_pattern_stats", ptrs.data_ptr(), count, n, s.data_ptr(), ss.data_ptr(), am.data_ptr(), _stream())
        _sync(sync)
        return s, ss, am

    def ncc(self, patterns=None, precision: str = "split_bf16", stats=None, out=None, sync: bool = True,
            ptrs=None):
        """NCC matrix of ``patterns`` (a (M, h*w) CUDA tensor or a list of
        CUDA tensors), or of a prebuilt device pointer table ``ptrs``."""
        if ptrs is None:
            ptrs = pattern_pointer_table(patterns)
        count = int(ptrs.numel())
        if stats is None:
            stats = self.stats(ptrs, count, sync=False)
        s, ss, am = stats
        if out is None:
            out = torch.empty((count, count), dtype=torch.float64, device="cuda")
        _lib.call("prnu_gram_ncc", self._ctx.handle, ptrs.data_ptr(), count, _lib.PRECISIONS[precision],
                  s.data_ptr(), ss.data_ptr(), am.data_ptr(), out.data_ptr(), _stream())
        _sync(sync)
        return out

    def close(self) -> None:
        self._ctx.close()


class PatternStore:
    """Device-resident, K-blocked pattern set for the batched NCC path (see
    ``prnu_store_*`` in include/prnu_api.h): ``format`` "split_bf16", "fp32",
    "fp8" (e4m3, scale max|x|/448) or "fp16" (power-of-two scale, one product).
    The role of Rocket's device cache slots (``R/cache/DeviceCache.java``),
    laid out so that one Gram K stage reads contiguous memory."""

    def __init__(self, n: int, capacity: int, format: str = "split_bf16"):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.prnu_store_create(int(n), int(capacity), _lib.PRECISIONS[format], ctypes.byref(h)),
                   "prnu_store_create")
        self.handle = h
        self.n, self.capacity, self.format = int(n), int(capacity), format

    def put(self, patterns, first_slot: int = 0, sync: bool = True) -> None:
        ptrs = pattern_pointer_table(patterns)
        _lib.call("prnu_store_put", self.handle, ptrs.data_ptr(), int(ptrs.numel()), int(first_slot), _stream())
        _sync(sync)

    def put_ptrs(self, ptrs, count: int, first_slot: int = 0, sync: bool = True) -> None:
        _lib.call("prnu_store_put", self.handle, ptrs.data_ptr(), int(count), int(first_slot), _stream())
        _sync(sync)

    def stats(self, count: int, first_slot: int = 0):
        """(sum, sumsq, absmax) of slots [first, first+count) as CPU numpy arrays."""
        import numpy as np
        s, ss, am = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _lib.call("prnu_store_stats", self.handle, ctypes.byref(s), ctypes.byref(ss), ctypes.byref(am))
        out = []
        for ptr, dt, size in [(s, np.float64, 8), (ss, np.float64, 8), (am, np.float32, 4)]:
            host = np.empty(count, dtype=dt)
            _lib.call("prnu_memcpy_d2h", host.ctypes.data, ptr.value + size * first_slot, size * count, None)
            out.append(host)
        return tuple(out)

    def gram(self, gram_ctx: "GramNCC", count: int, first_slot: int = 0, out=None, sync: bool = True):
        if out is None:
            out = torch.empty((count, count), dtype=torch.float64, device="cuda")
        _lib.call("prnu_store_gram_ncc", gram_ctx._ctx.handle, self.handle, int(first_slot), int(count),
                  out.data_ptr(), _stream())
        _sync(sync)
        return out

    def image_bytes(self, count: int) -> int:
        """Bytes of a slot image of `count` slots (``prnu_store_image_bytes``)."""
        return int(_lib.load().prnu_store_image_bytes(self.handle, int(count)))

    def save(self, first_slot: int, count: int, image, sync: bool = True) -> None:
        """Slots [first_slot, first_slot + count) into `image`, a CPU uint8
        tensor of image_bytes(count) bytes (pinned for an asynchronous copy)."""
        if image.is_cuda or image.dtype != torch.uint8 or image.numel() < self.image_bytes(count):
            raise ValueError("image must be a CPU uint8 tensor of image_bytes(count) bytes")
        _lib.call("prnu_store_save", self.handle, int(first_slot), int(count), image.data_ptr(), _stream())
        _sync(sync)

    def load(self, first_slot: int, image, sync: bool = True) -> None:
        """A slot image (written by save, same length and format) into slots
        from first_slot on."""
        if image.is_cuda or image.dtype != torch.uint8 or not image.is_contiguous():
            raise ValueError("image must be a contiguous CPU uint8 tensor")
        _lib.call("prnu_store_load", self.handle, int(first_slot), image.data_ptr(), int(image.numel()), _stream())
        _sync(sync)

    def close(self) -> None:
        if self.handle:
            _lib.load().prnu_store_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _GramGroup(ctypes.Structure):
    _fields_ = [("store", ctypes.c_void_p), ("first_slot", ctypes.c_int32), ("count", ctypes.c_int32),
                ("d_ncc", ctypes.c_void_p)]


class _GramBlock(ctypes.Structure):
    _fields_ = [("store", ctypes.c_void_p), ("row_first", ctypes.c_int32), ("row_count", ctypes.c_int32),
                ("col_first", ctypes.c_int32), ("col_count", ctypes.c_int32), ("d_out", ctypes.c_void_p),
                ("ld_out", ctypes.c_int64)]


class GramPlan:
    """All-pairs NCC of several pattern groups (e.g. one per image size) in one
    launch (``prnu_gram_plan_*``).  ``groups`` is a list of
    ``(PatternStore, first_slot, count)``; outputs are (count, count) FP64
    CUDA tensors, one per group, in :attr:`outputs`.

    With ``blocks`` instead, a list of ``(PatternStore, row_first, row_count,
    col_first, col_count, out)`` where ``out`` is a 2-D FP64 CUDA tensor view
    (row stride = leading dimension): each block's NCC goes into its view
    (``prnu_gram_plan_create_blocks``)."""

    def __init__(self, gram_ctx: "GramNCC", groups=None, slab_values: int = 0, blocks=None):
        lib = _lib.load()
        h = ctypes.c_void_p()
        if blocks is None:
            self._keep = [g[0] for g in groups]          # stores must outlive the plan
            self.outputs = [torch.empty((int(c), int(c)), dtype=torch.float64, device="cuda") for _, _, c in groups]
            arr = (_GramGroup * len(groups))()
            for i, ((store, first, count), out) in enumerate(zip(groups, self.outputs)):
                arr[i] = _GramGroup(store.handle.value, int(first), int(count), out.data_ptr())
            _lib.check(lib.prnu_gram_plan_create(gram_ctx._ctx.handle, arr, len(groups), int(slab_values),
                                                 ctypes.byref(h)), "prnu_gram_plan_create")
        else:
            self._keep = [b[0] for b in blocks]
            self.outputs = [b[5] for b in blocks]
            arr = (_GramBlock * len(blocks))()
            for i, (store, r0, rn, c0, cn, out) in enumerate(blocks):
                if out.dtype != torch.float64 or not out.is_cuda or out.stride(1) != 1:
                    raise ValueError("block outputs must be FP64 CUDA views with unit column stride")
                if out.shape[0] < rn or out.shape[1] < cn:
                    raise ValueError("block output view smaller than the block")
                arr[i] = _GramBlock(store.handle.value, int(r0), int(rn), int(c0), int(cn), out.data_ptr(),
                                    int(out.stride(0)))
            _lib.check(lib.prnu_gram_plan_create_blocks(gram_ctx._ctx.handle, arr, len(blocks), int(slab_values),
                                                        ctypes.byref(h)), "prnu_gram_plan_create_blocks")
        self.handle = h
        self._ctx = gram_ctx

    def run(self, use_graph: bool = True, sync: bool = True):
        _lib.call("prnu_gram_plan_run", self.handle, int(bool(use_graph)), _stream())
        _sync(sync)
        return self.outputs

    def autotune(self) -> tuple[int, float]:
        v, ms = ctypes.c_int64(), ctypes.c_float()
        _lib.call("prnu_gram_plan_autotune", self.handle, _stream(), ctypes.byref(v), ctypes.byref(ms))
        _sync(True)
        return v.value, ms.value

    @property
    def persistent(self) -> bool:
        """True: one workgroup per CU takes work items from per-XCD queues;
        False: one workgroup per work item.  Results are bitwise the same."""
        v = ctypes.c_int()
        _lib.call("prnu_gram_plan_get_launch", self.handle, ctypes.byref(v))
        return v.value == _lib.PLAN_PERSISTENT

    @persistent.setter
    def persistent(self, on: bool) -> None:
        _lib.call("prnu_gram_plan_set_launch", self.handle, _lib.PLAN_PERSISTENT if on else _lib.PLAN_PER_ITEM)

    def info(self) -> dict:
        items, wg, slab, part, flop = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int64(), ctypes.c_size_t(), \
            ctypes.c_double()
        _lib.call("prnu_gram_plan_info", self.handle, ctypes.byref(items), ctypes.byref(wg), ctypes.byref(slab),
                  ctypes.byref(part), ctypes.byref(flop))
        return {"work_items": items.value, "workgroups": wg.value, "slab_values": slab.value,
                "partial_bytes": part.value, "flop": flop.value}

    def close(self) -> None:
        if self.handle:
            _lib.load().prnu_gram_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
