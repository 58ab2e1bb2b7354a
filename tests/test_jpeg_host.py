"""Host-side JPEG header parsing of the library (prnu_jpeg_info), no GPU:
sizes of the committed fixtures, and the files the decoder rejects as the
reference's readJPEG does (readjpeg/src/main/cpp/native.cpp:99-103: not
3-component RGB; :26-29: input ends before libjpeg is done) or does not
implement (progressive)."""
import ctypes
import glob
import io
import os

import numpy as np
import pytest

from rocket_amd import _lib, jpeg

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "jpeg", "*.jpg"))))
def test_info_reads_fixture_sizes(path):
    want = np.load(path[:-4] + ".rgb.npy").shape[:2]
    assert jpeg.jpeg_size(open(path, "rb").read()) == want


def _encode(img, **kw):
    Image = pytest.importorskip("PIL.Image")
    buf = io.BytesIO()
    Image.fromarray(img).save(buf, "JPEG", **kw)
    return buf.getvalue()


def test_info_rejects_what_readjpeg_rejects():
    rng = np.random.default_rng(0)
    rgb = (rng.random((16, 24, 3)) * 255).astype(np.uint8)
    with pytest.raises(_lib.PrnuError) as e:
        jpeg.jpeg_size(_encode(rgb[:, :, 0]))            # gray: 1 component
    assert e.value.status == _lib.ENOTSUP
    with pytest.raises(_lib.PrnuError) as e:
        jpeg.jpeg_size(_encode(rgb, progressive=True))
    assert e.value.status == _lib.ENOTSUP
    good = _encode(rgb)
    with pytest.raises(_lib.PrnuError) as e:
        jpeg.jpeg_size(good[: len(good) * 2 // 3])       # no EOI
    assert e.value.status == _lib.EINVAL
    with pytest.raises(_lib.PrnuError):
        jpeg.jpeg_size(b"\xff\xd8\xff\xdb\x00")          # truncated segment
    with pytest.raises(_lib.PrnuError):
        jpeg.jpeg_size(b"GIF89a not a jpeg")
    assert jpeg.jpeg_size(good) == (16, 24)


def _read_jpeg(data, capacity):
    lib = _lib.load()
    w, h = ctypes.c_int32(-1), ctypes.c_int32(-1)
    out = ctypes.create_string_buffer(max(1, capacity))
    rc = lib.readJPEG(data, len(data), ctypes.byref(w), ctypes.byref(h), out, capacity)
    return rc, w.value, h.value, out


def test_readjpeg_drop_in_rejects_before_the_device():
    """readJPEG's own failures (native.cpp:99-108): 1, sizes untouched, for a
    gray file, a non-JPEG and an output buffer smaller than 3 * w * h; all
    decided on the host."""
    gray = _encode(np.zeros((12, 20), dtype=np.uint8))
    rgb = _encode(np.zeros((12, 20, 3), dtype=np.uint8))
    for data, cap in ((gray, 10000), (b"\xff\xd8 not a jpeg", 10000), (rgb, 12 * 20 * 3 - 1)):
        rc, w, h, _ = _read_jpeg(data, cap)
        assert (rc, w, h) == (1, -1, -1)
