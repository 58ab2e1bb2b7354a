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
