"""CPU tests of the C restatement that times the reference's CPU path
(oracle/cpu_port.c, built by oracle/Makefile): its FFT against numpy's, its
extraction chain (the applyCPU methods of kernels/filter/) and spectrum
against the FP64 oracle, its PCE (both variants) against the oracle's, and
its NCC against the FP64 definition."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import prnu_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(ROOT, "oracle", "_build", "libprnu_cpu_port.so")


@pytest.fixture(scope="module")
def port():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    lib = ctypes.CDLL(LIB)
    lib.port_fft.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.port_prnu_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.port_pce_pairs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    return lib


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 12, 15, 16, 30, 45, 60, 64, 100, 120, 243, 256, 360, 512, 625,
                               1000, 1024, 1536, 2048])
def test_fft_matches_numpy(port, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    for inverse, ref in ((0, np.fft.fft(x.astype(np.complex128))), (1, np.fft.ifft(x.astype(np.complex128)) * n)):
        buf = x.copy()
        assert port.port_fft(buf.ctypes.data, n, inverse) == 0
        scale = np.abs(ref).max()
        assert np.abs(buf - ref).max() <= 2e-6 * scale * max(1.0, np.log2(n)), (n, inverse)


def test_fft_rejects_other_factors(port):
    buf = np.zeros(7, dtype=np.complex64)
    assert port.port_fft(buf.ctypes.data, 7, 0) == -1


def _chain(port, imgs, fmt, spectra=False):
    count, h, w = imgs.shape[:3]
    pats = np.zeros((count, h, w), dtype=np.float32)
    spec = np.zeros((count, h, w // 2 + 1), dtype=np.complex64) if spectra else None
    rc = port.port_prnu_batch(np.ascontiguousarray(imgs).ctypes.data, count, h, w, fmt, pats.ctypes.data,
                              spec.ctypes.data if spectra else None, 2)
    assert rc == 0
    return pats, spec


@pytest.mark.parametrize("h,w,rgb", [(48, 64, True), (45, 60, False), (64, 64, True), (30, 50, True), (75, 40, False)])
def test_extraction_chain_matches_oracle(port, h, w, rgb):
    rng = np.random.default_rng(h * w)
    shape = (3, h, w, 3) if rgb else (3, h, w)
    imgs = rng.integers(0, 256, size=shape, dtype=np.uint8)
    pats, spec = _chain(port, imgs, 0 if rgb else 1, spectra=True)
    for k in range(3):
        want = O.prnu_pattern(imgs[k])
        assert np.abs(pats[k] - want).max() <= 1e-4 * np.abs(want).max(), k
        ws = O.spectrum(pats[k].astype(np.float64))
        assert np.abs(spec[k] - ws).max() <= 1e-5 * np.abs(ws).max() * np.log2(h * w), k


@pytest.mark.parametrize("variant,fn", [(0, O.pce_reference), (1, O.pce_textbook)])
@pytest.mark.parametrize("h,w", [(48, 64), (45, 60), (64, 50)])
def test_pce_matches_oracle(port, variant, fn, h, w):
    rng = np.random.default_rng(7 * h + w)
    k = rng.standard_normal((h, w))
    pats = np.stack([0.3 * k + rng.standard_normal((h, w)), 0.3 * k + rng.standard_normal((h, w)),
                     rng.standard_normal((h, w))]).astype(np.float32)
    spec = np.stack([np.fft.rfft2(p.astype(np.float64)) for p in pats]).astype(np.complex64)
    pairs = np.array([(0, 1), (0, 2), (1, 1), (2, 0)], dtype=np.int32)
    out = np.zeros(len(pairs))
    assert port.port_pce_pairs(spec.ctypes.data, h, w, pairs.ctypes.data, len(pairs), variant, out.ctypes.data, 2) == 0
    for q, (i, j) in enumerate(pairs):
        assert out[q] == pytest.approx(fn(pats[i].astype(np.float64), pats[j].astype(np.float64)), rel=1e-4), (i, j)


def test_ncc_matches_fp64_definition(port):
    """NormalizedCrossCorrelation.applyCPU: sum of FP32 products in FP64 over
    sqrt(sum x^2 * sum y^2) (no mean removal, as the CPU code does)."""
    port.port_sumsq.restype = ctypes.c_double
    port.port_sumsq.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    port.port_ncc_cpu.restype = ctypes.c_double
    port.port_ncc_cpu.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_double, ctypes.c_int64]
    rng = np.random.default_rng(3)
    x, y = (rng.standard_normal(10007).astype(np.float32) for _ in range(2))
    sx, sy = port.port_sumsq(x.ctypes.data, x.size), port.port_sumsq(y.ctypes.data, y.size)
    assert sx == pytest.approx(float(np.sum((x * x).astype(np.float64))), rel=1e-12)
    got = port.port_ncc_cpu(x.ctypes.data, sx, y.ctypes.data, sy, x.size)
    want = float(np.sum((x * y).astype(np.float64))) / np.sqrt(sx * sy)
    assert got == pytest.approx(want, rel=1e-12)
    assert port.port_ncc_cpu(x.ctypes.data, sx, x.ctypes.data, sx, x.size) == pytest.approx(1.0, rel=1e-12)
