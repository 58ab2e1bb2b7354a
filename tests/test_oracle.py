"""Pins of the FP64 oracle against mathematics and independent library routines.

The reference holds no golden vectors for this path (its filter tests compare
CPU and GPU on unseeded random pixels at 5%), so each oracle function is pinned
here by closed forms, invariants, numpy/scipy routines that reduce to it, or
brute force on tiny inputs.  Each check is chosen so that a plausible mistake
(wrong border stencil, transposed axes, dropped term, wrong sign, wrong
mirroring) fails.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import prnu_oracle as O


def dft_matrix(n, sign=-1):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


def dft2_brute(x):
    h, w = x.shape
    return dft_matrix(h) @ x @ dft_matrix(w)


# ----------------------------------------------------------------- grayscale
def test_grayscale_closed_form(rng):
    rgb = rng.integers(0, 256, size=(5, 7, 3), dtype=np.uint8)
    g = O.grayscale(rgb)
    b0, b1, b2 = (rgb[..., i].astype(np.float64) for i in range(3))
    # byte 2 carries the 0.299 weight (quirk Q1): byte order R,G,B from libjpeg
    assert np.allclose(g, 0.299 * b2 + 0.587 * b1 + 0.114 * b0, rtol=1e-7)
    assert not np.allclose(g, 0.299 * b0 + 0.587 * b1 + 0.114 * b2, rtol=1e-3)


def test_grayscale_gray8_is_identity_to_1e7():
    g = np.arange(256, dtype=np.uint8).reshape(16, 16)
    assert np.allclose(O.grayscale_from_gray8(g), g.astype(np.float64), rtol=1e-8, atol=0)
    rgb = np.repeat(g[..., None], 3, axis=2)
    assert np.allclose(O.grayscale(rgb), O.grayscale_from_gray8(g), rtol=1e-15)


# ----------------------------------------------------------------- fastnoise
def fastnoise_via_numpy_gradient(x):
    dy, dx = np.gradient(x)            # edge_order=1: one-sided at borders
    s = 1.0 / (1.0 + np.hypot(dx, dy))
    return np.gradient(dx * s, axis=1) + np.gradient(dy * s, axis=0)


@pytest.mark.parametrize("shape", [(8, 8), (13, 7), (2, 2), (3, 17)])
def test_fastnoise_matches_numpy_gradient_divergence(rng, shape):
    x = rng.normal(size=shape) * 10
    assert np.allclose(O.fastnoise(x), fastnoise_via_numpy_gradient(x), rtol=1e-13, atol=1e-13)


def test_fastnoise_constant_and_ramp_are_zero():
    assert np.abs(O.fastnoise(np.full((9, 6), 3.5))).max() == 0.0
    ramp = np.add.outer(2.0 * np.arange(9), -3.0 * np.arange(6))
    assert np.abs(O.fastnoise(ramp)).max() < 1e-14


def test_fastnoise_quadratic_closed_form():
    # x = j^2: dx = 1 at j=0, 2j inside, 2w-3 at w-1; dy = 0 -> s = 1/(1+|dx|)
    w = 10
    x = np.tile(np.arange(w, dtype=np.float64) ** 2, (4, 1))
    dx = np.array([1.0] + [2.0 * j for j in range(1, w - 1)] + [2.0 * w - 3.0])
    u = dx / (1.0 + np.abs(dx))
    expect = np.empty(w)
    expect[0] = u[1] - u[0]
    expect[-1] = u[-1] - u[-2]
    expect[1:-1] = 0.5 * (u[2:] - u[:-2])
    assert np.allclose(O.fastnoise(x), np.tile(expect, (4, 1)), rtol=1e-14)


def test_fastnoise_transpose_equivariant(rng):
    x = rng.normal(size=(11, 6))
    assert np.allclose(O.fastnoise(x.T), O.fastnoise(x).T, rtol=1e-14, atol=1e-14)


def test_fastnoise_rejects_degenerate():
    with pytest.raises(ValueError):
        O.fastnoise(np.zeros((1, 5)))


# ----------------------------------------------------------------- zeromean
@pytest.mark.parametrize("shape", [(8, 6), (7, 9), (2, 3), (5, 2)])
def test_zeromean_invariants(rng, shape):
    x = rng.normal(size=shape) + 5.0
    z = O.zeromean(x)
    for par in (0, 1):
        assert np.abs(z[par::2, :].mean(axis=0)).max() < 1e-12   # columns, per row parity
        assert np.abs(z[:, par::2].mean(axis=1)).max() < 1e-12   # rows, per column parity
    assert np.allclose(O.zeromean(z), z, atol=1e-12)             # idempotent


def test_zeromean_closed_form_and_order(rng):
    x = rng.normal(size=(9, 8)) * 3.0
    h, w = x.shape
    pi = (np.arange(h) % 2)[:, None]
    pj = (np.arange(w) % 2)[None, :]
    c = np.zeros((h, w))
    r = np.zeros((h, w))
    g = np.zeros((h, w))
    for a in (0, 1):
        c += np.where(pi == a, x[a::2, :].mean(axis=0, keepdims=True), 0.0)
        r += np.where(pj == a, x[:, a::2].mean(axis=1, keepdims=True), 0.0)
        for b in (0, 1):
            g += np.where((pi == a) & (pj == b), x[a::2, b::2].mean(), 0.0)
    assert np.allclose(O.zeromean(x), x - c - r + g, atol=1e-12)
    # rows-first (CPU order) equals columns-first (GPU order)
    rows_first = O._zeromean_cols(O._zeromean_cols(x.T).T)
    assert np.allclose(rows_first, O.zeromean(x), atol=1e-12)


# ----------------------------------------------------------------- wiener
@pytest.mark.parametrize("shape", [(16, 12), (30, 20), (15, 10)])
def test_numpy_fft_is_the_dft(rng, shape):
    x = rng.normal(size=shape)
    assert np.allclose(np.fft.fft2(x), dft2_brute(x), rtol=1e-10, atol=1e-9)
    half = shape[1] // 2 + 1
    assert np.allclose(O.spectrum(x), dft2_brute(x)[:, :half], rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("size", O.FILTER_SIZES)
def test_box_mean_matches_scipy(rng, size):
    p = rng.random(size=(13, 11))
    ref = ndimage.uniform_filter(p, size=size, mode="constant", cval=0.0)
    assert np.allclose(O.box_mean_zero_padded(p, size), ref, rtol=1e-12, atol=1e-14)


def test_wiener_constant_closed_form():
    # x = c: F = n c at DC only; the 9x9 box at DC averages n^2 c^2 / 81 > var,
    # so scale = var / varest = 81 / (n - 1) and the output is c * 81 / (n - 1).
    h, w, c = 32, 24, 1.7
    n = h * w
    out = O.wiener(np.full((h, w), c))
    assert np.allclose(out, c * 81.0 / (n - 1), rtol=1e-12)


def test_wiener_matches_brute_force(rng):
    x = rng.normal(size=(10, 6))
    h, w = x.shape
    n = h * w
    f = dft2_brute(x)
    p = np.abs(f) ** 2
    est = np.full(p.shape, np.inf)
    for s in O.FILTER_SIZES:
        r = s // 2
        box = np.zeros_like(p)
        for i in range(h):
            for j in range(w):
                acc = 0.0
                for di in range(-r, r + 1):
                    for dj in range(-r, r + 1):
                        if 0 <= i + di < h and 0 <= j + dj < w:
                            acc += p[i + di, j + dj]
                box[i, j] = acc / (s * s)
        est = np.minimum(est, box)
    var = np.sum(x * x) * n / (n - 1)
    g = f * var / np.maximum(var, est)
    inv = (dft_matrix(h, +1) @ g @ dft_matrix(w, +1)).real / n
    assert np.allclose(O.wiener(x), inv, rtol=1e-10, atol=1e-12)


# ----------------------------------------------------------------- NCC
def test_ncc_identities(rng):
    x = rng.normal(size=4096).astype(np.float32)
    sx, sxx = O.pattern_sums(x)
    assert O.ncc_reference(x, x) == pytest.approx(sxx / (sxx - sx * sx), rel=1e-12)
    assert O.ncc_reference(x, -x) == pytest.approx(-sxx / (sxx - sx * sx), rel=1e-12)
    xz = (x - x.mean()).astype(np.float32)
    assert O.ncc_reference(xz, xz) == pytest.approx(1.0, abs=1e-9)
    assert O.ncc_reference(xz, -xz) == pytest.approx(-1.0, abs=1e-9)
    assert O.ncc_reference(xz, np.float32(4.0) * xz) == pytest.approx(1.0, abs=1e-9)
    y = rng.normal(size=4096).astype(np.float32)
    assert O.ncc_reference(x, y) == pytest.approx(O.ncc_reference(y, x), rel=1e-14)


def test_ncc_sums_fp32_products():
    # 1 + 2^-12 squared = 1 + 2^-11 + 2^-24: the FP32 square drops 2^-24 (ties to even)
    x = np.full(8, 1.0 + 2.0 ** -12, dtype=np.float32)
    _, sxx32 = O.pattern_sums(x, fp32_products=True)
    _, sxx64 = O.pattern_sums(x, fp32_products=False)
    assert sxx32 == 8 * (1.0 + 2.0 ** -11)
    assert sxx64 == 8 * (1.0 + 2.0 ** -11 + 2.0 ** -24)


def test_ncc_matrix_vs_pairwise(rng):
    # zero-mean rows: the reference formula has no 1/n on Sx^2 (quirk Q5) and
    # goes NaN when Sx^2 > Sxx, which extracted patterns never approach
    p = rng.normal(size=(7, 1000))
    p = (p - p.mean(axis=1, keepdims=True)).astype(np.float32)
    m = O.ncc_matrix(p)
    for i in range(7):
        for j in range(7):
            assert m[i, j] == pytest.approx(O.ncc_reference(p[i], p[j], fp32_products=False), rel=1e-12, abs=1e-14)
    assert np.allclose(m, m.T)


# ----------------------------------------------------------------- PCE
def delta(h, w, dy, dx, v=1.0):
    d = np.zeros((h, w))
    d[dy % h, dx % w] = v
    return d


def test_reference_cross_spectrum_equals_true_on_row0_and_mirror_elsewhere(rng):
    h, w = 8, 10
    x = rng.normal(size=(h, w))
    y = rng.normal(size=(h, w))
    c = O.cross_spectrum_reference(O.spectrum(x), O.spectrum(y), w)
    true = np.conj(np.fft.fft2(x)) * np.fft.fft2(y)
    assert np.allclose(c[:, : w // 2 + 1], true[:, : w // 2 + 1])
    assert np.allclose(c[0], true[0]) and np.allclose(c[h // 2], true[h // 2])
    # right half, row i holds the true value of row (h - i) mod h
    rows = (h - np.arange(h)) % h
    assert np.allclose(c[:, w // 2 + 1:], true[rows][:, w // 2 + 1:])
    assert not np.allclose(c, true)


def test_pce_reference_delta_in_row0_closed_form():
    # x = delta(0,0), y = delta(0,dx): correlation is n * delta at (0, dx);
    # peak value is c[h-1][w-1] = 0 unless the shift lands there.
    h, w = 16, 12
    x = delta(h, w, 0, 0)
    y = delta(h, w, 0, 3)
    c = np.fft.ifft2(O.cross_spectrum_reference(O.spectrum(x), O.spectrum(y), w)).real * (h * w)
    expect = delta(h, w, 0, 3, h * w)
    assert np.allclose(c, expect, atol=1e-9)
    _, idx, _ = O.pce_reference_from_corr(expect + 1.0)
    assert idx == 3
    # shift to (-1, -1): peak value at the last element, window clipped at the borders
    y2 = delta(h, w, -1, -1)
    c2 = np.fft.ifft2(np.conj(np.fft.fft2(x)) * np.fft.fft2(y2)).real * (h * w)
    c2 = c2 + 0.5                                  # add a floor so the energy is non-zero
    pce2, idx2, e2 = O.pce_reference_from_corr(c2)
    assert idx2 == h * w - 1
    # window rows/cols (p-5, p+5) exclusive, clipped: 5 x 5 cells at the corner
    outside = h * w - 25
    assert e2 == pytest.approx(outside * 0.25 / (h * w - 121))
    peak = np.float32(h * w + 0.5)
    assert pce2 == pytest.approx(float(peak * peak) / e2, rel=1e-7)


def test_pce_reference_brute_force_8x8(rng):
    h, w = 8, 8
    x = rng.normal(size=(h, w))
    y = rng.normal(size=(h, w))
    xs = dft2_brute(x)[:, : w // 2 + 1]
    ys = dft2_brute(y)[:, : w // 2 + 1]
    full = np.empty((h, w), complex)
    for i in range(h):
        for j in range(w):
            if j < w // 2 + 1:
                full[i, j] = np.conj(xs[i, j]) * ys[i, j]
            else:
                full[i, j] = xs[i, w - j] * np.conj(ys[i, w - j])
    c = (dft_matrix(h, +1) @ full @ dft_matrix(w, +1)).real
    p = int(np.argmax(np.abs(c)))
    py, px = divmod(p, w)
    e = 0.0
    for i in range(h):
        for j in range(w):
            if py - 5 < i < py + 5 and px - 5 < j < px + 5:
                continue
            e += c[i, j] ** 2
    e /= h * w - 121
    want = float(np.float32(c[-1, -1]) ** 2) / e if e else np.inf
    got = O.pce_reference(x.astype(np.float32), y.astype(np.float32))
    assert got == pytest.approx(want, rel=1e-6)


def test_circular_xcorr_brute_force(rng):
    h, w = 6, 5
    x = rng.normal(size=(h, w))
    y = rng.normal(size=(h, w))
    c = np.zeros((h, w))
    for ty in range(h):
        for tx in range(w):
            c[ty, tx] = sum(x[a, b] * y[(a + ty) % h, (b + tx) % w] for a in range(h) for b in range(w))
    assert np.allclose(O.circular_xcorr(x, y), c)


def test_pce_textbook_delta_any_shift():
    h, w = 32, 24
    x = delta(h, w, 0, 0) * 3.0
    for dy, dx in [(0, 0), (5, 7), (-3, 2), (16, -1)]:
        y = delta(h, w, dy, dx) * 2.0 + 0.01
        c = O.circular_xcorr(x, y)
        pce, idx, energy = O.pce_textbook_from_corr(c)
        assert divmod(idx, w) == (dy % h, dx % w)
        # c = 6 at the shift, 0.03 elsewhere; window 11 x 11 wraps
        assert energy == pytest.approx(0.03 ** 2 * (h * w - 121) / (h * w - 121))
        assert pce == pytest.approx(6.03 ** 2 / 0.03 ** 2, rel=1e-9)


def test_pce_textbook_matches_reference_mode_on_row0_shift(rng):
    # Both modes see the same correlation plane when the shift is in row 0 and
    # the pattern is constant along rows (spectrum only in row 0).
    h, w = 16, 16
    base = rng.normal(size=(1, w))
    x = np.repeat(base, h, axis=0)
    y = np.roll(x, 3, axis=1)
    c_ref = np.fft.ifft2(O.cross_spectrum_reference(O.spectrum(x), O.spectrum(y), w)).real
    assert np.allclose(c_ref, O.circular_xcorr(x, y), atol=1e-9)


# ----------------------------------------------------------------- pair counting
@pytest.mark.parametrize("rb,re,cb,ce", [(0, 10, 0, 10), (0, 4, 6, 10), (3, 9, 2, 7), (5, 5, 0, 3), (2, 7, 0, 1)])
@pytest.mark.parametrize("diag", [False, True])
def test_triangle_size_brute_force(rb, re, cb, ce, diag):
    brute = sum(1 for i in range(rb, re) for j in range(cb, ce) if i < j or (diag and i == j))
    assert O.triangle_size_range(rb, re, cb, ce, diag) == brute
