"""GPU parity of extraction (gray, FastNoise, zero-mean, Wiener, spectrum),
the full PRNUFilter chain and PCE against the FP64 oracle.

Tolerances (FP32 kernels vs FP64 oracle), each relative to the largest
magnitude of the expected output:
* grayscale, FastNoise, zero-mean: 2e-6 (a few FP32 roundings per pixel);
* Wiener / spectrum / full chain: 2e-5 (two 2-D FP32 FFTs);
* PCE: relative PCE error <= 1e-4 (the target stated for the path);
* end-to-end NCC on GPU-extracted patterns vs the FP64 oracle chain: 1e-5.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import prnu_oracle as O  # noqa: E402
from rocket_amd import _lib, csi, synthetic  # noqa: E402

pytestmark = pytest.mark.gpu


def rel_err(got, want):
    return float(np.abs(np.asarray(got, np.float64) - want).max() / max(np.abs(want).max(), 1e-30))


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


SIZES = [(48, 64), (80, 96), (45, 60), (128, 100), (512, 512)]


@pytest.mark.parametrize("h,w", [(5, 7), (512, 512)])
def test_grayscale(h, w, rng):
    rgb = rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8)
    f = csi.PRNUFilter(h, w)
    out = torch.empty(h * w, device="cuda")
    f.get_grayscale_filter().apply_gpu(cuda(rgb), out, _lib.INPUT_RGB8)
    assert rel_err(out.cpu().numpy().reshape(h, w), O.grayscale(rgb)) <= 2e-6
    gray = rgb[..., 0].copy()
    f.get_grayscale_filter().apply_gpu(cuda(gray), out, _lib.INPUT_GRAY8)
    assert rel_err(out.cpu().numpy().reshape(h, w), O.grayscale_from_gray8(gray)) <= 2e-6
    with pytest.raises(ValueError):
        f.get_grayscale_filter().apply_gpu(cuda(gray), out, _lib.INPUT_RGB8)   # undersized input
    f.cleanup()


@pytest.mark.parametrize("h,w", SIZES + [(2, 2), (3, 130), (67, 2)])
def test_fastnoise(h, w, rng):
    x = rng.uniform(0, 255, size=(h, w)).astype(np.float32)
    f = csi.PRNUFilter(h, w)
    t = cuda(x.ravel())
    f.get_fast_noise_filter().apply_gpu(t)
    # the output is a sum of differences of normalised gradients, each in
    # (-1, 1): FP32 rounding is bounded absolutely (outputs may cancel to ~0)
    assert np.abs(t.cpu().numpy().reshape(h, w) - O.fastnoise(x)).max() <= 2e-6
    f.cleanup()


@pytest.mark.parametrize("h,w", SIZES + [(7, 9), (2, 3)])
def test_zeromean(h, w, rng):
    x = (rng.normal(size=(h, w)) + 3.0).astype(np.float32)
    f = csi.PRNUFilter(h, w)
    t = cuda(x.ravel())
    f.get_zero_mean_total_filter().apply_gpu(t)
    got = t.cpu().numpy().reshape(h, w)
    assert np.abs(got - O.zeromean(x)).max() <= 2e-6 * np.abs(x).max()
    f.cleanup()


@pytest.mark.parametrize("h,w", SIZES + [(1536, 2048), (1024, 768)])
def test_wiener(h, w, rng):
    x = O.zeromean(rng.normal(size=(h, w))).astype(np.float32)
    f = csi.PRNUFilter(h, w)
    t = cuda(x.ravel())
    f.get_wiener_filter().apply_gpu(t)
    assert rel_err(t.cpu().numpy().reshape(h, w), O.wiener(x)) <= 2e-5
    f.cleanup()


def test_wiener_constant_closed_form():
    h, w, c = 32, 24, 1.7
    f = csi.PRNUFilter(h, w)
    t = torch.full((h * w,), c, device="cuda")
    f.get_wiener_filter().apply_gpu(t)
    assert np.allclose(t.cpu().numpy(), c * 81.0 / (h * w - 1), rtol=2e-5)
    f.cleanup()


@pytest.mark.parametrize("h,w", [(48, 64), (45, 60), (80, 96), (1024, 1024), (1536, 2048), (1024, 2048)])
def test_spectral(h, w, rng):
    x = rng.normal(size=(h, w)).astype(np.float32)
    f = csi.PRNUFilter(h, w, apply_spectral_filter=True)
    out = torch.empty(2 * h * (w // 2 + 1), device="cuda")
    f.get_spectral_filter().apply_gpu(cuda(x.ravel()), out)
    got = out.cpu().numpy().view(np.complex64).reshape(h, w // 2 + 1)
    want = O.spectrum(x)
    assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()
    f.cleanup()


@pytest.mark.parametrize("h,w,rgb", [(80, 96, True), (96, 80, False), (512, 512, False), (45, 60, True),
                                   (1024, 1024, False)])
def test_preprocess_chain(h, w, rgb, rng):
    if rgb:
        img = rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8)
        fmt = _lib.INPUT_RGB8
    else:
        img = synthetic.make_image(3, h, w, 2)
        fmt = _lib.INPUT_GRAY8
    want = O.prnu_pattern(img)
    ctx = csi.CommonSourceIdentificationContext(h, w, csi.ComparisonStrategy.NCC)
    out = torch.empty(h * w, device="cuda")
    assert ctx.preprocess_input_gpu(0, cuda(img), out, fmt) == 4 * h * w
    assert rel_err(out.cpu().numpy().reshape(h, w), want) <= 2e-5
    ctx.destroy()
    # PCE strategy: the same pattern's R2C spectrum
    ctx = csi.CommonSourceIdentificationContext(h, w, csi.ComparisonStrategy.PCE)
    assert ctx.get_max_input_size() == 8 * h * (w // 2 + 1)
    spec = torch.empty(2 * h * (w // 2 + 1), device="cuda")
    ctx.preprocess_input_gpu(0, cuda(img), spec, fmt)
    got = spec.cpu().numpy().view(np.complex64).reshape(h, w // 2 + 1)
    s_want = O.spectrum(want)
    assert np.abs(got - s_want).max() <= 2e-5 * np.abs(s_want).max()
    ctx.destroy()


def test_preprocess_rejects_undersized():
    h, w = 16, 16
    ctx = csi.CommonSourceIdentificationContext(h, w, csi.ComparisonStrategy.NCC)
    img = torch.zeros(h * w * 3, dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError):
        ctx.preprocess_input_gpu(0, img, torch.empty(h * w - 1, device="cuda"))
    with pytest.raises(ValueError):
        ctx.preprocess_input_gpu(0, img[: h * w], torch.empty(h * w, device="cuda"))
    ctx.destroy()


def _spectra(h, w, pats):
    f = csi.PRNUFilter(h, w, apply_spectral_filter=True)
    out = []
    for p in pats:
        s = torch.empty(2 * h * (w // 2 + 1), device="cuda")
        f.get_spectral_filter().apply_gpu(cuda(p.astype(np.float32).ravel()), s)
        out.append(s)
    f.cleanup()
    return out


@pytest.mark.parametrize("h,w", [(48, 64), (45, 60), (64, 64), (256, 200)])
@pytest.mark.parametrize("variant", [_lib.PCE_REFERENCE, _lib.PCE_TEXTBOOK])
def test_pce_matches_oracle(h, w, variant, rng):
    k = rng.normal(size=(h, w))
    pats = [0.3 * k + rng.normal(size=(h, w)), 0.3 * k + rng.normal(size=(h, w)), rng.normal(size=(h, w))]
    pats = [p.astype(np.float32) for p in pats]
    spec = _spectra(h, w, pats)
    cmp = csi.PeakToCorrelationEnergy(h, w, variant)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    fn = O.pce_reference if variant == _lib.PCE_REFERENCE else O.pce_textbook
    for i, j in [(0, 1), (0, 2), (1, 1)]:
        cmp.apply_gpu(spec[i], spec[j], out)
        want = fn(pats[i].astype(np.float64), pats[j].astype(np.float64))
        assert out.item() == pytest.approx(want, rel=1e-4), (i, j)
    cmp.cleanup()


def test_pce_textbook_delta_closed_form():
    h, w = 32, 24
    x = np.zeros((h, w), np.float32)
    x[0, 0] = 3.0
    y = np.full((h, w), 0.01, np.float32)
    y[5, 7] += 2.0
    spec = _spectra(h, w, [x, y])
    cmp = csi.PeakToCorrelationEnergy(h, w, _lib.PCE_TEXTBOOK)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    cmp.apply_gpu(spec[0], spec[1], out)
    assert out.item() == pytest.approx(6.03 ** 2 / 0.03 ** 2, rel=1e-4)
    cmp.cleanup()


def test_correlate_pce_mode_is_reference_variant(rng):
    h, w = 64, 48
    a, b = (rng.normal(size=(h, w)).astype(np.float32) for _ in range(2))
    spec = _spectra(h, w, [a, b])
    ctx = csi.CommonSourceIdentificationContext(h, w, csi.ComparisonStrategy.PCE)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    assert ctx.correlate_gpu(0, spec[0], 1, spec[1], out) == 8
    assert out.item() == pytest.approx(O.pce_reference(a.astype(np.float64), b.astype(np.float64)), rel=1e-4)
    ctx.destroy()


def test_end_to_end_ncc_prnu64_subset():
    """Synthetic 512x512 images from 4 cameras: GPU extraction + split-BF16
    Gram vs the FP64 oracle chain; same-camera pairs score above others."""
    h, w, count, cams = 512, 512, 16, 4
    imgs = synthetic.make_images(count, h, w, cams, workers=1)
    ctx = csi.CommonSourceIdentificationContext(h, w, csi.ComparisonStrategy.NCC)
    pats = torch.empty((count, h * w), device="cuda")
    for i in range(count):
        ctx.preprocess_input_gpu(i, cuda(imgs[i]), pats[i], _lib.INPUT_GRAY8)
    gram = csi.GramNCC(h, w)
    got = gram.ncc(pats, "split_bf16").cpu().numpy()
    want = O.ncc_matrix(np.stack([O.prnu_pattern(im) for im in imgs]).astype(np.float32))
    assert np.abs(got - want).max() <= 1e-5
    cam = np.arange(count) % cams
    same = got[(cam[:, None] == cam[None, :]) & ~np.eye(count, dtype=bool)]
    diff = got[cam[:, None] != cam[None, :]]
    assert same.min() > diff.max()
    gram.close()
    ctx.destroy()


@pytest.mark.parametrize("variant", [_lib.PCE_REFERENCE, _lib.PCE_TEXTBOOK])
@pytest.mark.parametrize("h,w", [(48, 64), (45, 60)])
def test_pce_pairs_batched_equals_single(variant, h, w, rng):
    k = rng.normal(size=(h, w))
    pats = [(0.3 * k * (i % 2) + rng.normal(size=(h, w))).astype(np.float32) for i in range(6)]
    spec = _spectra(h, w, pats)
    cmp = csi.PeakToCorrelationEnergy(h, w, variant)
    pairs = np.array([(i, j) for i in range(6) for j in range(i + 1, 6)], dtype=np.int32)
    got = cmp.pairs(spec, pairs).cpu().numpy()
    one = torch.zeros(1, dtype=torch.float64, device="cuda")
    fn = O.pce_reference if variant == _lib.PCE_REFERENCE else O.pce_textbook
    for q, (i, j) in enumerate(pairs):
        cmp.apply_gpu(spec[i], spec[j], one)
        # pairs share complex transforms two by two: equal to the single-pair
        # result up to FP32 rounding of the shared transform.  That rounding
        # moves the peak by delta ~1e-6 of the plane's rms, so the PCE
        # (peak^2 / energy) moves by ~2 delta sqrt(PCE): a relative bound for
        # large PCE, a sqrt bound for noise-level ones
        assert got[q] == pytest.approx(one.item(), rel=1e-5, abs=2e-5 * np.sqrt(abs(one.item())))
        assert got[q] == pytest.approx(fn(pats[i].astype(np.float64), pats[j].astype(np.float64)), rel=1e-4)
    cmp.cleanup()


@pytest.mark.parametrize("h,w,fmt,spectral", [(45, 60, _lib.INPUT_RGB8, False), (80, 96, _lib.INPUT_GRAY8, False),
                                              (48, 64, _lib.INPUT_GRAY8, True), (96, 80, _lib.INPUT_RGB8, True)])
def test_preprocess_batch_equals_single(h, w, fmt, spectral, rng):
    """prnu_preprocess_batch over strided rows (padding between images and
    between patterns) is bitwise prnu_preprocess of each image alone, and
    matches the oracle chain."""
    count, bpp = 7, (3 if fmt == _lib.INPUT_RGB8 else 1)
    imgs = rng.integers(0, 256, size=(count, h * w * bpp + 37), dtype=np.uint8)   # 37 pad bytes per row
    f = csi.PRNUFilter(h, w, apply_spectral_filter=spectral)
    plen = f.get_output_size() // 4
    out = torch.full((count, plen + 16), -7.0, device="cuda")                      # 16 pad floats per row
    d_imgs = cuda(imgs)
    f.apply_gpu_batch(d_imgs[:, : h * w * bpp], out[:, :plen], fmt)
    one = torch.empty(plen, device="cuda")
    for i in range(count):
        f.apply_gpu(d_imgs[i, : h * w * bpp].contiguous(), one, fmt)
        assert torch.equal(out[i, :plen], one), f"image {i}"
        assert torch.all(out[i, plen:] == -7.0)                                   # padding untouched
    img0 = imgs[0, : h * w * bpp].reshape((h, w, 3) if bpp == 3 else (h, w))
    want = O.prnu_pattern(img0)
    got = out[0, :plen].cpu().numpy()
    if spectral:
        s_want = O.spectrum(want)
        got = got.view(np.complex64).reshape(h, w // 2 + 1)
        assert np.abs(got - s_want).max() <= 2e-5 * np.abs(s_want).max()
    else:
        assert rel_err(got.reshape(h, w), want) <= 2e-5
    f.cleanup()


def test_preprocess_batch_rejects_bad_strides():
    h, w = 16, 16
    f = csi.PRNUFilter(h, w)
    imgs = torch.zeros((3, h * w * 3), dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError):                    # pattern rows shorter than a pattern
        f.apply_gpu_batch(imgs, torch.empty((3, h * w - 1), device="cuda"))
    with pytest.raises(ValueError):                    # image rows shorter than an RGB image
        f.apply_gpu_batch(imgs[:, : h * w], torch.empty((3, h * w), device="cuda"))
    f.apply_gpu_batch(imgs[:0], torch.empty((0, h * w), device="cuda"))   # empty batch: no-op
    f.cleanup()
