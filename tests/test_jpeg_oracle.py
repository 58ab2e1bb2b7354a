"""CPU tests pinning the JPEG oracle (oracle/jpeg_oracle.py).

* the 12 ISLOW constants against the 9-digit values their names carry;
* the integer IDCT: a DC-only block in closed form, random blocks within one
  level of the exact (float64) inverse DCT;
* whole decodes against the committed fixtures (Pillow / libjpeg-turbo RGB)
  and, where Pillow is importable, against Pillow bit for bit over the
  reference test's frame sizes (ReadJPEGTest.java:74-95), 4:4:4 / 4:2:2 /
  4:2:0, qualities, optimized tables and restart intervals;
* what the reference rejects (readjpeg/src/main/cpp/native.cpp:99-103):
  gray (1 component); and what this decoder does not handle: progressive.
"""
import glob
import io
import math
import os

import numpy as np
import pytest

from oracle import jpeg_oracle as J

HERE = os.path.dirname(os.path.abspath(__file__))


def test_islow_constants_match_their_names():
    for name, v in J.FIX.items():
        assert v == round(float(name.replace("_", ".", 1)) * 8192), name


def test_idct_dc_only_closed_form():
    q = np.ones(64, dtype=np.int64)
    for d in (-1024, -500, -8, -1, 0, 1, 7, 100, 1016):
        c = np.zeros((1, 64), dtype=np.int64)
        c[0, 0] = d
        want = min(255, max(0, ((d + 4) >> 3) + 128))
        assert np.all(J.idct_islow(c, q) == want), d


def test_idct_within_one_of_exact():
    rng = np.random.default_rng(5)
    u = np.arange(8)
    cu = np.where(u == 0, 1 / math.sqrt(2), 1.0)
    basis = cu[None, :] * np.cos((2 * u[:, None] + 1) * u[None, :] * np.pi / 16)   # [x, u]
    for _ in range(200):
        f = np.zeros((8, 8))
        k = rng.integers(1, 20)
        idx = rng.choice(64, k, replace=False)
        f.flat[idx] = rng.integers(-300, 300, k)
        exact = 0.25 * basis @ f @ basis.T + 128          # rows x, cols y
        got = J.idct_islow(f.reshape(1, 64).astype(np.int64), np.ones(64, dtype=np.int64))[0].astype(float)
        assert np.abs(got - np.clip(np.round(exact), 0, 255)).max() <= 1


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "jpeg", "*.jpg"))))
def test_decode_matches_fixture(path):
    want = np.load(path[:-4] + ".rgb.npy")
    got = J.decode(open(path, "rb").read())
    assert got.shape == want.shape and np.array_equal(got, want)


def _frame(w, h, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([(x * 7 + y * 3) % 256, (x * x + y) % 256, 128 + 60 * np.sin(x / 5.0) * np.cos(y / 7.0)], -1)
    return (img + rng.normal(0, 12, img.shape)).clip(0, 255).astype(np.uint8)


@pytest.mark.parametrize("size", [(10, 10), (5, 3), (14, 8), (100, 100), (150, 123), (1, 1), (2, 5), (3, 9), (33, 2)])
@pytest.mark.parametrize("opts", [dict(quality=75, subsampling=0), dict(quality=75, subsampling=1),
                                  dict(quality=75, subsampling=2), dict(quality=92, subsampling=2, optimize=True),
                                  dict(quality=80, subsampling=2, restart_marker_blocks=2)])
def test_decode_matches_pillow(size, opts):
    Image = pytest.importorskip("PIL.Image")
    w, h = size
    buf = io.BytesIO()
    Image.fromarray(_frame(w, h, w * 1000 + h)).save(buf, "JPEG", **opts)
    data = buf.getvalue()
    want = np.asarray(Image.open(io.BytesIO(data)).convert("RGB"))
    assert np.array_equal(J.decode(data), want)


def test_rejects_gray_and_progressive():
    Image = pytest.importorskip("PIL.Image")
    buf = io.BytesIO()
    Image.fromarray(_frame(16, 16, 1)[:, :, 0]).save(buf, "JPEG")
    with pytest.raises(J.JpegError, match="RGB"):
        J.decode(buf.getvalue())
    buf = io.BytesIO()
    Image.fromarray(_frame(16, 16, 1)).save(buf, "JPEG", progressive=True)
    with pytest.raises(J.JpegError, match="progressive"):
        J.decode(buf.getvalue())
    with pytest.raises(J.JpegError):
        J.decode(b"not a jpeg at all")
