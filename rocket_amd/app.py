"""The common-source-identification application end to end on one GPU.

Mirrors the flow of the reference's ``Main`` (``CSI/Main.java:40-139``, under
``src/main/java/nl/esciencecenter/common_source_identification/``): every image
is preprocessed into a PRNU pattern, every pair is compared (NCC or PCE), the
scores fill a symmetric matrix with a zero diagonal, and the linkage,
clustering, edge list and matrix files are written.  The reference spreads
the pairs over Rocket's distributed scheduler; here extraction runs in
batches, NCC as one triangular Gram over a device-resident pattern store and
PCE as batched pair lists.  ``common_source_identification_dir`` takes a
folder of JPEG files as ``Main`` does; they are decoded on the GPU (libjpeg's
pixels, bit for bit) into the buffer the extraction reads.
"""
from __future__ import annotations

import os
import re

import numpy as np

from . import _lib, csi, csi_io


def _compare(pats, n: int, h: int, w: int, filt, pce: bool, precision: str, pce_variant: int):
    """All-pairs scores of n device-resident patterns (or spectra)."""
    if pce:
        cmp = csi.PeakToCorrelationEnergy(h, w, pce_variant, _ctx=filt._ctx)
        pairs = np.array([(i, j) for i in range(n) for j in range(i + 1, n)], dtype=np.int32)
        scores = cmp.pairs(pats, pairs).cpu().numpy() if len(pairs) else np.zeros(0)
        return csi_io.correlation_matrix(n, pairs, scores)
    gram = csi.GramNCC(h, w)
    store = csi.PatternStore(h * w, n, precision)
    store.put(pats)
    corrs = store.gram(gram, n).cpu().numpy()
    np.fill_diagonal(corrs, 0.0)               # Main.java:121-126 never sets the diagonal
    store.close()
    gram.close()
    return corrs


def correlation_matrix_gpu(images, strategy: str = "NCC", precision: str = "split_bf16",
                           pce_variant: int = _lib.PCE_REFERENCE, batch: int = 64):
    """Scores of all pairs of `images` (numpy or CUDA uint8, (N, h, w) gray or
    (N, h, w, 3) RGB) as an (N, N) float64 numpy matrix, zero diagonal."""
    import torch
    imgs = images if isinstance(images, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(images))
    n, h, w = int(imgs.shape[0]), int(imgs.shape[1]), int(imgs.shape[2])
    fmt = _lib.INPUT_RGB8 if imgs.dim() == 4 else _lib.INPUT_GRAY8
    pce = strategy.upper() == "PCE"
    filt = csi.PRNUFilter(h, w, apply_spectral_filter=pce)
    pats = torch.empty((n, filt.get_output_size() // 4), device="cuda")
    for b in range(0, n, batch):
        chunk = imgs[b:b + batch].reshape(min(batch, n - b), -1).cuda()
        filt.apply_gpu_batch(chunk, pats[b:b + batch], fmt)
    corrs = _compare(pats, n, h, w, filt, pce, precision, pce_variant)
    filt.cleanup()
    return corrs


def common_source_identification(images, filenames, testcase: str, outdir: str = ".", strategy: str = "NCC",
                                 precision: str = "split_bf16", pce_variant: int = _lib.PCE_REFERENCE):
    """Scores, clustering and result files (``Main.java:121-139``); returns
    (corrs, linkage, paths)."""
    corrs = correlation_matrix_gpu(images, strategy, precision, pce_variant)
    paths = csi_io.write_all(testcase, corrs, list(filenames), outdir)
    linkage = csi_io.hierarchical_clustering(corrs, list(filenames))
    return corrs, linkage, paths


def scan_jpeg_directory(folder: str) -> list[str]:
    """Names of the entries of `folder` ending in ".jpg" (any case), sorted by
    name: ``Util.scanDirectory`` with the filter of ``Main.java:88``
    (``rocket/util/Util.java:91-106``)."""
    if not os.path.isdir(folder):
        raise FileNotFoundError(f"folderpath does not exist or is not a directory: {folder}")
    return sorted(f for f in os.listdir(folder) if f.lower().endswith(".jpg"))


def correlation_matrix_jpeg(paths, strategy: str = "NCC", precision: str = "split_bf16",
                            pce_variant: int = _lib.PCE_REFERENCE, batch: int = 32):
    """All-pairs scores of the JPEG files `paths`, decoded on the GPU
    (``prnu_jpeg_decode_batch``, libjpeg's result bit for bit) straight into
    the device buffer the extraction reads.  The frame size is the first
    file's (``Main.java:97-104``); a file of another size raises ValueError
    ("image has incorrect dimensions", ``CommonSourceIdentificationContext.java:96-99``),
    one that does not decode raises RuntimeError ("error while reading",
    ``:91-95``)."""
    import torch
    from . import jpeg
    n = len(paths)
    if n == 0:
        raise ValueError("no JPEG files")
    data = [open(p, "rb").read() for p in paths]
    try:
        h, w = jpeg.jpeg_size(data[0])
    except _lib.PrnuError as e:
        raise RuntimeError(f"error while reading: {paths[0]}") from e
    for p, d in zip(paths, data):
        try:
            size = jpeg.jpeg_size(d)
        except _lib.PrnuError as e:
            raise RuntimeError(f"error while reading: {p}") from e
        if size != (h, w):
            raise ValueError(f"image has incorrect dimensions: {p}")
    pce = strategy.upper() == "PCE"
    filt = csi.PRNUFilter(h, w, apply_spectral_filter=pce)
    pats = torch.empty((n, filt.get_output_size() // 4), device="cuda")
    pool = jpeg.JpegDecoderPool(workers=3)      # batches decode while earlier ones are extracted
    try:
        starts = list(range(0, n, batch))
        decoded = pool.map((data[b:b + batch] for b in starts), shape=(h, w, 3), strict=False)
        for b, (out, status) in zip(starts, decoded):
            out.record_stream(torch.cuda.current_stream())   # decoded on a worker stream, read on this one
            m = len(status)
            for k, st in enumerate(status):
                if st not in (_lib.OK, _lib.JPEG_CORRUPT_DATA):   # libjpeg only warns on corrupt data
                    raise RuntimeError(f"error while reading: {paths[b + k]}")
            filt.apply_gpu_batch(out.view(m, -1), pats[b:b + m], _lib.INPUT_RGB8)
        corrs = _compare(pats, n, h, w, filt, pce, precision, pce_variant)
    finally:
        pool.close()
        filt.cleanup()
    return corrs


def common_source_identification_dir(testcase: str, folder: str, outdir: str = ".", strategy: str = "NCC",
                                      precision: str = "split_bf16", pce_variant: int = _lib.PCE_REFERENCE):
    """``Main``'s flow from a folder of JPEG files (``Main.java:69-139``): the
    test-case name check, the sorted ``.jpg`` listing, GPU decode and
    extraction, all pairs, clustering and the five result files named after
    the test case.  Returns (corrs, linkage, paths, filenames)."""
    if not re.fullmatch(r"[a-zA-Z0-9_.-]+", testcase):
        raise ValueError("testcase will be used for file names, please do not use special characters")
    names = scan_jpeg_directory(folder)
    corrs = correlation_matrix_jpeg([os.path.join(folder, f) for f in names], strategy, precision, pce_variant)
    paths = csi_io.write_all(testcase, corrs, names, outdir)
    linkage = csi_io.hierarchical_clustering(corrs, names)
    return corrs, linkage, paths, names
