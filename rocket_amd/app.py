"""The common-source-identification application end to end on one GPU.

Mirrors the flow of the reference's ``Main`` (``CSI/Main.java:40-139``, under
``src/main/java/nl/esciencecenter/common_source_identification/``): every image
is preprocessed into a PRNU pattern, every pair is compared (NCC or PCE), the
scores fill a symmetric matrix with a zero diagonal, and the linkage,
clustering, edge list and matrix files are written.  The reference spreads
the pairs over Rocket's distributed scheduler; here extraction runs in
batches, NCC as one triangular Gram over a device-resident pattern store and
PCE as batched pair lists.
"""
from __future__ import annotations

import numpy as np

from . import _lib, csi, csi_io


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
    if pce:
        cmp = csi.PeakToCorrelationEnergy(h, w, pce_variant, _ctx=filt._ctx)
        pairs = np.array([(i, j) for i in range(n) for j in range(i + 1, n)], dtype=np.int32)
        scores = cmp.pairs(pats, pairs).cpu().numpy() if len(pairs) else np.zeros(0)
        corrs = csi_io.correlation_matrix(n, pairs, scores)
    else:
        gram = csi.GramNCC(h, w)
        store = csi.PatternStore(h * w, n, precision)
        store.put(pats)
        corrs = store.gram(gram, n).cpu().numpy()
        np.fill_diagonal(corrs, 0.0)               # Main.java:121-126 never sets the diagonal
        store.close()
        gram.close()
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
