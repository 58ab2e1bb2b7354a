"""GPU JPEG decoding (prnu_jpeg_decode_batch) against libjpeg's default
decode, bit for bit: the committed fixtures (libjpeg-turbo RGB from Pillow),
the JPEG oracle, and Pillow itself on the frame sizes of the reference's
ReadJPEGTest (ReadJPEGTest.java:74-95) and the 2048x1536 bench frame, in
4:4:4, 4:2:2 and 4:2:0, with optimized tables and restart markers, batched
with mixed sizes; rejected files in a batch leave the others decoded."""
import glob
import io
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import jpeg_oracle as J  # noqa: E402
from rocket_amd import _lib, jpeg  # noqa: E402

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def frame(w, h, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([(x * 7 + y * 3) % 256, (x * x + y) % 256, 128 + 60 * np.sin(x / 5.0) * np.cos(y / 7.0)], -1)
    return (img + rng.normal(0, 12, img.shape)).clip(0, 255).astype(np.uint8)


def encode(img, **kw):
    Image = pytest.importorskip("PIL.Image")
    buf = io.BytesIO()
    Image.fromarray(img).save(buf, "JPEG", **kw)
    return buf.getvalue()


def pil_rgb(data):
    Image = pytest.importorskip("PIL.Image")
    return np.asarray(Image.open(io.BytesIO(data)).convert("RGB"))


def decode_mixed(dec, files, sizes):
    """Batch of files of different sizes into one buffer sized for the largest."""
    hmax = max(h for h, _ in sizes)
    wmax = max(w for _, w in sizes)
    out = torch.zeros((len(files), hmax * wmax * 3), dtype=torch.uint8, device="cuda")
    out, st = dec.decode(files, out=out)
    host = out.cpu().numpy()
    return [host[k, : h * w * 3].reshape(h, w, 3) for k, (h, w) in enumerate(sizes)], st


def test_fixtures_bitwise():
    paths = sorted(glob.glob(os.path.join(HERE, "golden", "jpeg", "*.jpg")))
    files = [open(p, "rb").read() for p in paths]
    wants = [np.load(p[:-4] + ".rgb.npy") for p in paths]
    dec = jpeg.JpegDecoder()
    got, st = decode_mixed(dec, files, [w.shape[:2] for w in wants])
    assert st == [0] * len(files)
    for p, g, w in zip(paths, got, wants):
        assert np.array_equal(g, w), os.path.basename(p)
    dec.close()


@pytest.mark.parametrize("sub", [0, 1, 2])
def test_reference_test_sizes_bitwise(sub):
    sizes = [(10, 10), (3, 5), (8, 14), (100, 100), (123, 150), (210, 170), (1000, 1000), (1523, 2147)]
    files = [encode(frame(w, h, 7 * w + h), quality=90, subsampling=sub) for h, w in sizes]
    dec = jpeg.JpegDecoder()
    got, st = decode_mixed(dec, files, sizes)
    assert st == [0] * len(files)
    for f, g, s in zip(files, got, sizes):
        assert np.array_equal(g, pil_rgb(f)), (s, sub)
    dec.close()


def test_restart_markers_optimized_tables_and_oracle():
    opts = [dict(quality=85, subsampling=2, restart_marker_blocks=4),
            dict(quality=70, subsampling=1, restart_marker_rows=1, optimize=True),
            dict(quality=95, subsampling=0, optimize=True)]
    files = [encode(frame(61, 47, k), **o) for k, o in enumerate(opts)]
    dec = jpeg.JpegDecoder()
    out, st = dec.decode(files)
    assert st == [0, 0, 0]
    for k, f in enumerate(files):
        g = out[k].cpu().numpy()
        assert np.array_equal(g, pil_rgb(f))
        assert np.array_equal(g, J.decode(f))
    dec.close()


def test_bench_frame_and_same_size_batch():
    files = [encode(frame(2048, 1536, s), quality=92, subsampling=2) for s in range(3)]
    dec = jpeg.JpegDecoder()
    out, st = dec.decode(files)
    assert out.shape == (3, 1536, 2048, 3) and st == [0, 0, 0]
    for k in (0, 2):
        assert np.array_equal(out[k].cpu().numpy(), pil_rgb(files[k]))
    dec.close()


@pytest.mark.parametrize("sub", [0, 2])
def test_noise_frames_bitwise(sub):
    """Pixel noise gives long codes and the slowest resynchronisation of
    chunk decoders guessed mid-stream; the result must not depend on it."""
    rng = np.random.default_rng(11 + sub)
    sizes = [(480, 640), (97, 333)]
    files = [encode(rng.integers(0, 256, (h, w, 3), dtype=np.uint8), quality=q, subsampling=sub)
             for (h, w), q in zip(sizes, (97, 60))]
    dec = jpeg.JpegDecoder()
    got, st = decode_mixed(dec, files, sizes)
    assert st == [0, 0]
    for f, g in zip(files, got):
        assert np.array_equal(g, pil_rgb(f))
    dec.close()


def test_corrupt_entropy_data_keeps_the_batch():
    """Bytes changed inside the entropy-coded data: the file still decodes
    (flagged when a code does not exist), the other files stay exact."""
    good = encode(frame(200, 150, 3), quality=90, subsampling=2)
    bad = bytearray(good)
    sos = bad.index(b"\xff\xda")
    start = sos + 2 + int.from_bytes(bad[sos + 2:sos + 4], "big")
    rng = np.random.default_rng(2)
    for i in rng.integers(start + 50, len(bad) - 50, 40):
        if bad[i] != 0xFF and bad[i - 1] != 0xFF and (bad[i] ^ 0x5A) != 0xFF:
            bad[i] ^= 0x5A
    files = [good, bytes(bad), good]
    dec = jpeg.JpegDecoder()
    out, st = dec.decode(files, strict=False)
    assert st[0] == 0 and st[2] == 0 and st[1] in (0, _lib.JPEG_CORRUPT_DATA)
    want = pil_rgb(good)
    assert np.array_equal(out[0].cpu().numpy(), want) and np.array_equal(out[2].cpu().numpy(), want)
    dec.close()


def test_rejected_files_leave_the_rest_decoded():
    good = [encode(frame(40, 30, 1)), encode(frame(40, 30, 2), subsampling=1)]
    prog = encode(frame(40, 30, 3), progressive=True)
    gray = encode(frame(40, 30, 4)[:, :, 0])
    trunc = good[0][: len(good[0]) // 2]
    files = [good[0], prog, b"not a jpeg", gray, good[1], trunc]
    dec = jpeg.JpegDecoder()
    out = torch.zeros((len(files), 30, 40, 3), dtype=torch.uint8, device="cuda")
    out, st = dec.decode(files, out=out, strict=False)
    assert st == [0, _lib.ENOTSUP, _lib.EINVAL, _lib.ENOTSUP, 0, _lib.EINVAL]
    assert np.array_equal(out[0].cpu().numpy(), pil_rgb(good[0]))
    assert np.array_equal(out[4].cpu().numpy(), pil_rgb(good[1]))
    with pytest.raises(_lib.PrnuError):
        dec.decode(files, out=out)
    small = torch.zeros((1, 10), dtype=torch.uint8, device="cuda")   # stride too small for 40x30x3
    _, st = dec.decode(good[:1], out=small, strict=False)
    assert st == [_lib.EINVAL]
    dec.close()


def test_readjpeg_drop_in_matches_libjpeg():
    """The exported readJPEG (native.cpp:59-134's signature): one file into a
    host buffer, sizes out, 0 on success; then a rejected file gives 1."""
    import ctypes
    lib = _lib.load()
    for sub in (0, 2):
        data = encode(frame(77, 45, sub), quality=88, subsampling=sub)
        cap = 77 * 45 * 3 + 5
        w, h = ctypes.c_int32(), ctypes.c_int32()
        out = ctypes.create_string_buffer(cap)
        assert lib.readJPEG(data, len(data), ctypes.byref(w), ctypes.byref(h), out, cap) == 0
        assert (w.value, h.value) == (77, 45)
        got = np.frombuffer(out.raw[: 77 * 45 * 3], dtype=np.uint8).reshape(45, 77, 3)
        assert np.array_equal(got, pil_rgb(data))
    prog = encode(frame(30, 20, 1), progressive=True)
    assert lib.readJPEG(prog, len(prog), ctypes.byref(w), ctypes.byref(h), out, cap) == 1


def test_decoder_pool_matches_single_decoder():
    """JpegDecoderPool: batches come back in order, each equal to one
    decoder's result, with several batches in flight on worker streams."""
    files = [encode(frame(96, 64, s), quality=85, subsampling=2) for s in range(12)]
    single = jpeg.JpegDecoder()
    want = [single.decode(files[b:b + 3])[0].cpu().numpy() for b in range(0, 12, 3)]
    single.close()
    pool = jpeg.JpegDecoderPool(workers=3)
    got = []
    for out, st in pool.map((files[b:b + 3] for b in range(0, 12, 3)), shape=(64, 96, 3)):
        assert st == [0, 0, 0]
        out.record_stream(torch.cuda.current_stream())
        got.append(out.cpu().numpy())
    pool.close()
    assert len(got) == 4 and all(np.array_equal(g, w) for g, w in zip(got, want))
