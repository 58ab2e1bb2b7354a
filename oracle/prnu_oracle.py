"""FP64 CPU oracle for the PRNU hot path (TEST INFRASTRUCTURE ONLY).

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may import
it.  The HIP library (``rocket_amd``) must never call into it.

It restates, in float64 numpy, the semantics of the reference's GPU path
("reference mode", quirks Q1-Q7 reproduced) and a textbook PCE.  Reference
paths below are relative to
``src/main/java/nl/esciencecenter/common_source_identification/``.

Pinning (how each function is checked independently of its own formula; the
tests live in ``tests/test_oracle.py``):

* ``grayscale``          closed form per pixel; channel order from libjpeg's RGB.
* ``fastnoise``          equals the divergence built from ``numpy.gradient``
                         (same one-sided borders); constant -> 0; ramp -> 0;
                         transpose equivariance.
* ``zeromean``           invariants (every row/column even/odd mean is 0),
                         idempotence, and the closed form x - c - r + g.
* ``wiener``             numpy FFT checked against an O(N^2) DFT matrix on
                         radix-2/3/5 sizes; box means against
                         ``scipy.ndimage.uniform_filter(mode='constant')``;
                         constant image -> closed form c*81/(n-1).
* ``spectrum``           ``numpy.fft.rfft2`` against the DFT matrix.
* ``ncc_reference``      NCC(x,x) = Sxx/(Sxx-Sx^2); NCC(x,-x) = -1;
                         NCC(x,a x) = sign(a); symmetry; Gram vs pairs.
* ``pce_reference``      delta patterns: closed-form correlation for shifts in
                         row 0 (where the row-mirroring quirk is exact) and an
                         O(N^2) brute-force correlation on 8x8.
* ``pce_textbook``       delta patterns with any shift: closed form; brute force.

No function here is left unpinned.
"""
from __future__ import annotations

import numpy as np

# FP32 constants exactly as the reference kernels use them (grayscalefilter.cu:63)
W_R = float(np.float32(0.299))
W_G = float(np.float32(0.587))
W_B = float(np.float32(0.114))
FILTER_SIZES = (3, 5, 7, 9)          # wienerfilter.cu:52, WienerFilter.java:253
PCE_RADIUS = 5                       # peaktocorrelationenergy.cu:218
PCE_SQUARE = 11                      # peaktocorrelationenergy.cu:217


# --------------------------------------------------------------------------- a2
def grayscale(rgb: np.ndarray) -> np.ndarray:
    """u8 (h, w, 3) -> float64 (h, w).

    ``grayscalefilter.cu:58-63``: with ``uchar3 c`` = bytes (0, 1, 2),
    ``out = 0.299*c.z + 0.587*c.y + 0.114*c.x``.  libjpeg emits R, G, B, so the
    effective weights are 0.299*B + 0.587*G + 0.114*R (quirk Q1).
    """
    c = rgb.astype(np.float64)
    return W_R * c[..., 2] + W_G * c[..., 1] + W_B * c[..., 0]


def grayscale_from_gray8(gray: np.ndarray) -> np.ndarray:
    """u8 (h, w) fed as R = G = B (the 8-bit gray workloads)."""
    g = gray.astype(np.float64)
    return (W_R + W_G + W_B) * g


# --------------------------------------------------------------------------- a3
def _hgrad(a: np.ndarray) -> np.ndarray:
    """fastnoisefilter.cu:150-161: one-sided at j=0 and j=w-1, central inside."""
    d = np.empty_like(a)
    d[:, 1:-1] = 0.5 * (a[:, 2:] - a[:, :-2])
    d[:, 0] = a[:, 1] - a[:, 0]
    d[:, -1] = a[:, -1] - a[:, -2]
    return d


def _vgrad(a: np.ndarray) -> np.ndarray:
    """fastnoisefilter.cu:163-175."""
    return _hgrad(a.T).T


def fastnoise(x: np.ndarray) -> np.ndarray:
    """FastNoise (first step total variation), ``fastnoisefilter.cu:177-204``.

    dx, dy = gradients of x; s = 1 / (1 + sqrt(dx^2 + dy^2)) (EPS = 1,
    ``FastNoiseFilter.java:30``); the output *replaces* the image with
    hgrad(dx*s) + vgrad(dy*s) (quirk Q2).  Requires h, w >= 2.
    """
    x = np.asarray(x, dtype=np.float64)
    if x.shape[0] < 2 or x.shape[1] < 2:
        raise ValueError("fastnoise needs h >= 2 and w >= 2")
    dx = _hgrad(x)
    dy = _vgrad(x)
    s = 1.0 / (1.0 + np.sqrt(dx * dx + dy * dy))
    return _hgrad(dx * s) + _vgrad(dy * s)


# --------------------------------------------------------------------------- a4
def _zeromean_cols(x: np.ndarray) -> np.ndarray:
    """``zeromeantotalfilter.cu:65-113``: per column, subtract the mean of the
    even rows from even rows and of the odd rows from odd rows.  Counts are
    (h+1)>>1 and h>>1."""
    out = x.copy()
    out[0::2, :] -= x[0::2, :].mean(axis=0, keepdims=True)
    if x.shape[0] > 1:
        out[1::2, :] -= x[1::2, :].mean(axis=0, keepdims=True)
    return out


def zeromean(x: np.ndarray) -> np.ndarray:
    """ZeroMeanTotal, GPU order: columns, then rows
    (``ZeroMeanTotalFilter.java:94-98``; ``zeromeantotalfilter.cu:176-224``).
    The CPU applies rows first (``ZeroMeanTotalFilter.java:108-115``); both
    orders agree in exact arithmetic (quirk Q3)."""
    x = np.asarray(x, dtype=np.float64)
    y = _zeromean_cols(x)
    return _zeromean_cols(y.T).T


# --------------------------------------------------------------------------- a5
def box_mean_zero_padded(p: np.ndarray, size: int) -> np.ndarray:
    """Centered size x size mean with zero padding and no wrap
    (``wienerfilter.cu:164-247``).  Computed with 2-D prefix sums."""
    h, w = p.shape
    r = size // 2
    pad = np.zeros((h + 2 * r + 1, w + 2 * r + 1))
    pad[r + 1:r + 1 + h, r + 1:r + 1 + w] = p
    c = pad.cumsum(0).cumsum(1)
    s = c[size:size + h, size:size + w] - c[0:h, size:size + w] - c[size:size + h, 0:w] + c[0:h, 0:w]
    return s / float(size * size)


def wiener_variance_estimate(p: np.ndarray) -> np.ndarray:
    """min over f in {3,5,7,9} of the f x f zero-padded box mean of P."""
    est = np.full(p.shape, np.inf)
    for f in FILTER_SIZES:
        est = np.minimum(est, box_mean_zero_padded(p, f))
    return est


def wiener(x: np.ndarray) -> np.ndarray:
    """Frequency-domain Wiener filter, ``WienerFilter.java:187-240`` and
    ``wienerfilter.cu``:

    F = FFT2(x) (unnormalised, row-major h x w); P = |F|^2;
    varest = min_f box_f(P) (zero padded, unshifted spectrum);
    var = n/(n-1) * sum(x^2)  (``wienerfilter.cu:304-340``);
    F *= var / max(var, varest); out = Re(IFFT2(F)) / n (quirk Q4).
    """
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    f = np.fft.fft2(x)
    p = f.real * f.real + f.imag * f.imag
    varest = wiener_variance_estimate(p)
    var = float(np.sum(x * x)) * n / (n - 1)
    f = f * (var / np.maximum(var, varest))
    return np.fft.ifft2(f).real


# --------------------------------------------------------------------------- a1
def prnu_pattern(image: np.ndarray) -> np.ndarray:
    """PRNUFilter chain (``PRNUFilter.java:124-138``): gray -> FastNoise ->
    ZeroMean -> Wiener.  ``image`` is u8 (h, w) gray or (h, w, 3) RGB."""
    if image.ndim == 3:
        g = grayscale(image)
    else:
        g = grayscale_from_gray8(image)
    return wiener(zeromean(fastnoise(g)))


# --------------------------------------------------------------------------- a6
def spectrum(pattern: np.ndarray) -> np.ndarray:
    """R2C FFT2 half spectrum h x (w/2+1), ``SpectralFilter.java:44-54``."""
    return np.fft.rfft2(np.asarray(pattern, dtype=np.float64))


# --------------------------------------------------------------------------- a8
def pattern_sums(x: np.ndarray, fp32_products: bool = True) -> tuple[float, float]:
    """(sum x, sum x^2); with ``fp32_products`` each square is rounded to FP32
    before the FP64 sum, as ``normalizedcrosscorrelation.cu:153-161`` does."""
    x32 = np.asarray(x, dtype=np.float32).ravel()
    sx = float(np.sum(x32.astype(np.float64)))
    if fp32_products:
        sxx = float(np.sum((x32 * x32).astype(np.float64)))
    else:
        x64 = x32.astype(np.float64)
        sxx = float(np.dot(x64, x64))
    return sx, sxx


def ncc_reference(x: np.ndarray, y: np.ndarray, fp32_products: bool = True) -> float:
    """NCC exactly as ``normalizedcrosscorrelation.cu:85-178`` defines it
    (quirk Q5): sums of FP32 patterns, products rounded to FP32 and
    accumulated in FP64, and

        NCC = Sxy / (sqrt(Sxx - Sx^2) * sqrt(Syy - Sy^2))

    with no 1/n on Sx^2 and no mean subtraction in Sxy."""
    x32 = np.asarray(x, dtype=np.float32).ravel()
    y32 = np.asarray(y, dtype=np.float32).ravel()
    sx, sxx = pattern_sums(x32, fp32_products)
    sy, syy = pattern_sums(y32, fp32_products)
    if fp32_products:
        sxy = float(np.sum((x32 * y32).astype(np.float64)))
    else:
        sxy = float(np.dot(x32.astype(np.float64), y32.astype(np.float64)))
    return sxy / (np.sqrt(sxx - sx * sx) * np.sqrt(syy - sy * sy))


def ncc_matrix(patterns: np.ndarray) -> np.ndarray:
    """M x M NCC in FP64 with exact products (the FP64 oracle every Gram
    precision is compared with).  ``patterns`` is (M, n) FP32."""
    p = np.asarray(patterns, dtype=np.float32).reshape(len(patterns), -1).astype(np.float64)
    g = p @ p.T
    s = p.sum(axis=1)
    ss = np.einsum("ij,ij->i", p, p)
    d = np.sqrt(ss - s * s)
    return g / np.outer(d, d)


# --------------------------------------------------------------------------- a9
def cross_spectrum_reference(xs: np.ndarray, ys: np.ndarray, w: int) -> np.ndarray:
    """Full h x w cross spectrum built from two R2C halves as
    ``peaktocorrelationenergy.cu:102-134`` does: for j <= w/2,
    conj(X[i][j]) * Y[i][j]; for j > w/2, X[i][w-j] * conj(Y[i][w-j]) on the
    *same* row i, not (h-i) mod h (quirk Q6)."""
    h = xs.shape[0]
    c = np.empty((h, w), dtype=np.complex128)
    half = w // 2 + 1
    c[:, :half] = np.conj(xs[:, :half]) * ys[:, :half]
    j = np.arange(half, w)
    c[:, half:] = xs[:, w - j] * np.conj(ys[:, w - j])
    return c


def pce_reference_from_corr(c: np.ndarray) -> tuple[float, int, float]:
    """Peak/energy stage of ``peaktocorrelationenergy.cu:155-368`` on the real
    correlation plane ``c`` (any scale).  Returns (pce, peak_index, energy).

    * peak location = argmax |c| (first in row-major order on ties);
    * the exclusion window keeps rows/cols strictly inside (p-5, p+5), i.e.
      9 x 9, clipped at the borders, not wrapped (``:240-242``);
    * energy = sum of c^2 outside / (N - 121) in FP64 (``:245-265``);
    * the peak *value* is the last element c[h-1][w-1] (``:199``), squared in
      FP32 (``:366``).
    """
    h, w = c.shape
    n = h * w
    idx = int(np.argmax(np.abs(c)))
    py, px = divmod(idx, w)
    rows = np.arange(h)[:, None]
    cols = np.arange(w)[None, :]
    inside = (rows > py - PCE_RADIUS) & (rows < py + PCE_RADIUS) & \
             (cols > px - PCE_RADIUS) & (cols < px + PCE_RADIUS)
    energy = float(np.sum(np.where(inside, 0.0, c * c))) / float(n - PCE_SQUARE * PCE_SQUARE)
    peak = np.float32(c[h - 1, w - 1])
    pce = float(np.float32(peak * peak)) / energy
    return pce, idx, energy


def pce_reference(x: np.ndarray, y: np.ndarray) -> float:
    """Reference-mode PCE of two spatial FP32 patterns (both go through the
    R2C spectrum first, as the PCE strategy's preprocess does)."""
    h, w = x.shape
    c = np.fft.ifft2(cross_spectrum_reference(spectrum(x), spectrum(y), w)).real * (h * w)
    return pce_reference_from_corr(c)[0]


def pce_textbook_from_corr(c: np.ndarray) -> tuple[float, int, float]:
    """Textbook PCE on a circular correlation plane: peak at argmax |c|,
    11 x 11 window centred on it with cyclic wrap, energy = mean of c^2
    outside the window (FP64), PCE = sign(c_p) * c_p^2 / energy."""
    h, w = c.shape
    n = h * w
    idx = int(np.argmax(np.abs(c)))
    py, px = divmod(idx, w)
    mask = np.ones((h, w), dtype=bool)
    ry = (py + np.arange(-PCE_RADIUS, PCE_RADIUS + 1)) % h
    rx = (px + np.arange(-PCE_RADIUS, PCE_RADIUS + 1)) % w
    mask[np.ix_(ry, rx)] = False
    excluded = n - int(mask.sum())
    energy = float(np.sum(c[mask] ** 2)) / float(n - excluded)
    cp = float(c[py, px])
    return float(np.sign(cp) * cp * cp / energy), idx, energy


def circular_xcorr(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """c[t] = sum_s x[s] y[s + t] (cyclic), = IFFT2(conj(X) Y) * n / n."""
    return np.fft.ifft2(np.conj(np.fft.fft2(x)) * np.fft.fft2(y)).real


def pce_textbook(x: np.ndarray, y: np.ndarray) -> float:
    return pce_textbook_from_corr(circular_xcorr(np.asarray(x, np.float64),
                                                 np.asarray(y, np.float64)))[0]


# --------------------------------------------------------------------------- a11
def triangle_size(n: int) -> int:
    """``R/util/Util.java:75-77``."""
    return n * (n + 1) // 2


def triangle_size_rect(num_rows: int, num_cols: int) -> int:
    """``R/util/Util.java:66-73``."""
    if num_cols > num_rows:
        return triangle_size(num_rows) + num_rows * (num_cols - num_rows)
    return triangle_size(num_cols)


def triangle_size_range(row_begin: int, row_end: int, col_begin: int, col_end: int,
                        include_diagonal: bool) -> int:
    """Number of pairs (i, j) with i in [rb, re), j in [cb, ce), i < j (or
    i <= j), ``R/util/Util.java:51-64``."""
    if not include_diagonal:
        return triangle_size_range(row_begin + 1, row_end + 1, col_begin, col_end, True)
    if row_begin >= row_end and col_begin >= col_end:
        return 0
    return (triangle_size_rect(row_begin, col_begin) + triangle_size_rect(row_end, col_end)
            - triangle_size_rect(row_begin, col_end) - triangle_size_rect(row_end, col_begin))
