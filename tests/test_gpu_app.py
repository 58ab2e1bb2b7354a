"""GPU test of the application flow: synthetic images of 4 cameras ->
batched extraction -> all-pairs NCC (split-BF16 store Gram) and PCE -> result
files.  Checks the matrix against the FP64 oracle chain, that the
hierarchical clustering merges images of one camera before any cross-camera
link, and that the files round-trip."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import prnu_oracle as O  # noqa: E402
from rocket_amd import _lib, app, csi_io, synthetic  # noqa: E402

pytestmark = pytest.mark.gpu


def test_ncc_app_end_to_end(tmp_path):
    h, w, n, cams = 64, 80, 16, 4
    imgs = synthetic.make_images(n, h, w, cams, workers=1)
    names = [f"img{i:02d}.jpg" for i in range(n)]
    corrs, linkage, paths = app.common_source_identification(imgs, names, "synthetic16", str(tmp_path))
    pats = np.stack([O.prnu_pattern(im) for im in imgs])
    want = O.ncc_matrix(pats)
    np.fill_diagonal(want, 0.0)
    assert np.abs(corrs - want).max() <= 1e-5
    # all same-camera merges (12 of them) come before the first cross-camera merge
    cam = lambda members: {m % cams for m in members}
    members = {i: [i] for i in range(n)}
    nid = n
    for l in linkage[: n - cams]:
        a, b = members.pop(l.n1), members.pop(l.n2)
        assert len(cam(a + b)) == 1
        members[nid] = a + b
        nid += 1
    back = np.fromfile(paths["matrix_bin"], dtype=">f8").reshape(n, n)
    assert np.array_equal(back, corrs)
    assert open(paths["clustering"]).read().startswith("flat clustering:\n0=[0]\n")   # NCC < 60: singletons


def test_pce_app_scores():
    h, w, n, cams = 64, 64, 8, 2
    imgs = synthetic.make_images(n, h, w, cams, workers=1)
    c = app.correlation_matrix_gpu(imgs, "PCE", pce_variant=_lib.PCE_TEXTBOOK)
    pats = [O.prnu_pattern(im) for im in imgs]
    for i, j in [(0, 1), (0, 2), (3, 5)]:
        assert c[i, j] == pytest.approx(O.pce_textbook(pats[i], pats[j]), rel=1e-4)
    assert c[0, 2] > c[0, 1]                       # images 0 and 2 share a camera


def _write_jpegs(folder, imgs, names, **kw):
    Image = pytest.importorskip("PIL.Image")
    for im, nm in zip(imgs, names):
        Image.fromarray(im).save(folder / nm, "JPEG", **kw)


def test_jpeg_folder_app_matches_decoded_images(tmp_path):
    """Main's flow from a folder of JPEG files: the GPU decode gives libjpeg's
    pixels, so the scores equal those of the same pipeline fed with Pillow's
    decode of the files, bit for bit."""
    Image = pytest.importorskip("PIL.Image")
    h, w, n, cams = 96, 128, 8, 2
    gray = synthetic.make_images(n, h, w, cams, workers=1)
    rgb = np.stack([np.stack([g, np.roll(g, 3, axis=1), 255 - g], -1) for g in gray])
    names = [f"cam{i % cams}_{i:02d}.{'JPG' if i % 3 == 0 else 'jpg'}" for i in range(n)]
    _write_jpegs(tmp_path, rgb, names, quality=95, subsampling=2)
    (tmp_path / "notes.txt").write_text("not an image")
    out = tmp_path / "out"
    out.mkdir()
    corrs, linkage, paths, listed = app.common_source_identification_dir("jpeg8", str(tmp_path), str(out))
    assert listed == sorted(names)
    decoded = np.stack([np.asarray(Image.open(tmp_path / f).convert("RGB")) for f in listed])
    assert np.array_equal(corrs, app.correlation_matrix_gpu(decoded))
    assert np.array_equal(np.fromfile(paths["matrix_bin"], dtype=">f8").reshape(n, n), corrs)
    for kind in ("PCE",):
        c = app.correlation_matrix_jpeg([str(tmp_path / f) for f in listed], kind, pce_variant=_lib.PCE_TEXTBOOK)
        assert np.array_equal(c, app.correlation_matrix_gpu(decoded, kind, pce_variant=_lib.PCE_TEXTBOOK))


def test_jpeg_folder_app_rejects_like_main(tmp_path):
    rgb = np.repeat(synthetic.make_images(3, 32, 48, 1, workers=1)[..., None], 3, axis=-1)
    _write_jpegs(tmp_path, list(rgb[:2]), ["a.jpg", "b.jpg"])
    _write_jpegs(tmp_path, [np.ascontiguousarray(rgb[2][:16])], ["c.jpg"])   # another size
    with pytest.raises(ValueError, match="incorrect dimensions"):
        app.correlation_matrix_jpeg([str(tmp_path / f) for f in ("a.jpg", "b.jpg", "c.jpg")])
    (tmp_path / "d.jpg").write_bytes(b"not a jpeg")
    with pytest.raises(RuntimeError, match="error while reading"):
        app.correlation_matrix_jpeg([str(tmp_path / f) for f in ("a.jpg", "d.jpg")])
