"""GPU parity at the full workload sizes (each test runs in seconds).

* PCE of 1024x1024 spectra (256 pairs of 24 extracted patterns, which crosses
  the batched path's pair batches and takes its 8-column strip column pass):
  both variants against the FP64 restatement on sampled pairs, relative
  error <= 1e-4, and the batched scores against the single-pair drop-in.
* Uneven groups at the configured frame sizes and counts (2000 @ 1024x1024,
  1500 @ 2048x1536, 596 @ 4000x3000) in one grouped split-BF16 launch:
  symmetry, unit diagonal, agreement with one store Gram per group, camera
  separation over every pair, and an FP64 oracle sample of 24 patterns per
  group (|dNCC| <= 1e-5).
* Extraction at 1536x2048 and 3000x4000 (rows of 2048 and 4000 values, the
  radix-5 length 4000, columns longer than 2048): the whole chain against the
  FP64 oracle (2e-5 of the largest magnitude), NCC patterns and spectra.
* A serving store of 20,000 fingerprints of 1024x1024: queries from stored
  fingerprints find them first, exact re-scores match the FP64 oracle NCC
  (relative 1e-10), and each query's answer is bitwise the same alone and in
  a batch.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import prnu_oracle as O  # noqa: E402
from rocket_amd import _lib, csi, serving, synthetic  # noqa: E402

pytestmark = pytest.mark.gpu

NCC_TOL = 1e-5


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rel_err(got, want):
    return float(np.abs(np.asarray(got, np.float64) - want).max() / max(np.abs(want).max(), 1e-30))


@pytest.fixture(autouse=True)
def _release_cached_memory():
    yield
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


# ---------------------------------------------------------------- extraction
@pytest.mark.parametrize("h,w,rgb", [(1536, 2048, True), (3000, 4000, False), (4000, 3000, True)])
def test_chain_at_camera_frame_sizes(h, w, rgb, rng):
    print(f"chain {w}x{h} rgb={rgb}", flush=True)
    if rgb:
        gray = synthetic.make_image(5, h, w, 3)
        img = np.stack([gray, np.roll(gray, 3, axis=1), 255 - gray], axis=-1)
        fmt = _lib.INPUT_RGB8
    else:
        img = synthetic.make_image(5, h, w, 3)
        fmt = _lib.INPUT_GRAY8
    want = O.prnu_pattern(img)
    f = csi.PRNUFilter(h, w)
    out = torch.empty(h * w, device="cuda")
    assert f.apply_gpu(cuda(img), out, fmt) == 4 * h * w
    assert rel_err(out.cpu().numpy().reshape(h, w), want) <= 2e-5
    f.cleanup()
    fs = csi.PRNUFilter(h, w, apply_spectral_filter=True)
    spec = torch.empty(fs.get_output_size() // 4, device="cuda")
    fs.apply_gpu(cuda(img), spec, fmt)
    got = spec.cpu().numpy().view(np.complex64).reshape(h, w // 2 + 1)
    s_want = O.spectrum(want)
    assert np.abs(got - s_want).max() <= 2e-5 * np.abs(s_want).max()
    fs.cleanup()


@pytest.mark.parametrize("h,w", [(8, 6000), (6000, 8), (5, 8192), (12, 4000)])
def test_wiener_long_rows_and_columns(h, w, rng):
    """Rows above 4096 values take the largest row-FFT tier; 6000 = 2^4 3 5^3."""
    print(f"wiener {w}x{h}", flush=True)
    x = O.zeromean(rng.normal(size=(h, w))).astype(np.float32)
    f = csi.PRNUFilter(h, w)
    t = cuda(x.ravel())
    f.get_wiener_filter().apply_gpu(t)
    assert rel_err(t.cpu().numpy().reshape(h, w), O.wiener(x)) <= 2e-5
    f.cleanup()


# ----------------------------------------------------------------------- PCE
def test_pce_pairs_1024x1024_against_oracle():
    """24 extracted patterns of 1024x1024 from 3 cameras; 256 of their pairs
    through prnu_pce_pairs (several pair batches); 8 pairs scored by the FP64
    restatement (4 same-camera), both variants."""
    h = w = 1024
    count, cams = 24, 3
    print("pce 1024x1024", flush=True)
    imgs = synthetic.make_images(count, h, w, cams, workers=1)
    fn = csi.PRNUFilter(h, w)
    pats = torch.empty((count, h * w), device="cuda")
    fn.apply_gpu_batch(cuda(imgs.reshape(count, -1)), pats, _lib.INPUT_GRAY8)
    fn.cleanup()
    fs = csi.PRNUFilter(h, w, apply_spectral_filter=True)
    spec = torch.empty((count, fs.get_output_size() // 4), device="cuda")
    for i in range(count):
        fs.get_spectral_filter().apply_gpu(pats[i], spec[i])
    host = pats.cpu().numpy().reshape(count, h, w).astype(np.float64)
    allp = np.array([(i, j) for i in range(count) for j in range(i + 1, count)], dtype=np.int32)
    pairs = allp[np.random.default_rng(4).choice(len(allp), 256, replace=False)]
    check = [(0, 3), (1, 4), (2, 5), (6, 9), (0, 1), (1, 2), (7, 8), (22, 23)]   # camera of image i: i % 3
    pairs[:len(check)] = check
    one = torch.zeros(1, dtype=torch.float64, device="cuda")
    for variant, oracle in ((_lib.PCE_REFERENCE, O.pce_reference), (_lib.PCE_TEXTBOOK, O.pce_textbook)):
        cmp = csi.PeakToCorrelationEnergy(h, w, variant, _ctx=fs._ctx)
        got = cmp.pairs(spec, pairs).cpu().numpy()
        assert np.all(np.isfinite(got))
        for q, (i, j) in enumerate(check):
            want = oracle(host[i], host[j])
            assert got[q] == pytest.approx(want, rel=1e-4), f"variant {variant} pair {(i, j)}"
        for q in (0, 4, 100, 255):
            i, j = pairs[q]
            cmp.apply_gpu(spec[i], spec[j], one)
            # the batched path shares complex transforms two pairs at a time:
            # equal to the single pair up to that FP32 rounding (see
            # test_gpu_extract.test_pce_pairs_batched_equals_single)
            assert got[q] == pytest.approx(one.item(), rel=1e-5, abs=2e-5 * np.sqrt(abs(one.item())))
        if variant == _lib.PCE_TEXTBOOK:
            same = np.array([i % cams == j % cams for i, j in pairs])
            assert got[same].min() > 100 * max(got[~same].max(), 1.0)
    fs.cleanup()


# -------------------------------------------------------------------- groups
GROUPS = [(2000, 1024, 1024, 16), (1500, 1536, 2048, 16), (596, 3000, 4000, 8)]


def test_uneven_groups_at_config_sizes():
    gain = 0.2            # same-camera NCC ~0.04, far above the 1/sqrt(n) noise of 2M pairs (5.3e-3 at n = 1M)
    g = torch.Generator(device="cuda").manual_seed(2718)
    stores, samples = [], []
    for count, h, w, cams in GROUPS:
        n = h * w
        print(f"group {count} x {w}x{h}", flush=True)
        k = torch.randn((cams, n), generator=g, device="cuda")
        st = csi.PatternStore(n, count, "split_bf16")
        chunk = max(24, (1 << 30) // (4 * n))          # the first chunk holds the 24-pattern sample
        for b in range(0, count, chunk):
            e = min(count, b + chunk)
            blk = torch.randn((e - b, n), generator=g, device="cuda")
            blk += gain * k[torch.arange(b, e, device="cuda") % cams]
            blk -= blk.mean(dim=1, keepdim=True)
            st.put(blk, first_slot=b)
            if b == 0:
                samples.append(blk[:24].cpu().numpy())
            del blk
        del k
        stores.append(st)
    gram = csi.GramNCC(8, 8)
    plan = csi.GramPlan(gram, [(st, 0, c) for st, (c, _, _, _) in zip(stores, GROUPS)])
    outs = [o.cpu().numpy() for o in plan.run(use_graph=True)]
    for (count, h, w, cams), st, got, host in zip(GROUPS, stores, outs, samples):
        assert np.array_equal(got, got.T)
        assert np.abs(np.diag(got) - 1.0).max() <= NCC_TOL
        err = np.abs(got[:24, :24] - O.ncc_matrix(host)).max()
        assert err <= NCC_TOL, f"{count} x {w}x{h}: max |dNCC| vs FP64 = {err:.3e}"
        single = st.gram(gram, count).cpu().numpy()
        assert np.abs(single - got).max() <= NCC_TOL            # only the slab summation order differs
        cam = np.arange(count) % cams
        same = cam[:, None] == cam[None, :]
        off = ~np.eye(count, dtype=bool)
        assert got[same & off].min() > got[~same].max(), f"{count} x {w}x{h}: cameras not separated"
    plan.close()
    for st in stores:
        st.close()
    gram.close()


# ------------------------------------------------------------------- serving
SERVE_N, SERVE_SIZE, SERVE_CHUNK = 1024 * 1024, 20000, 500


def _fingerprint_chunk(c):
    g = torch.Generator(device="cuda").manual_seed(10_000 + c)
    f = torch.randn((SERVE_CHUNK, SERVE_N), generator=g, device="cuda")
    return f - f.mean(dim=1, keepdim=True)


def test_serving_store_of_20000_1024x1024():
    print("fingerprint store 20000 x 1024x1024", flush=True)
    store = serving.FingerprintStore(1024, 1024, SERVE_SIZE)
    for c in range(SERVE_SIZE // SERVE_CHUNK):
        store.put(_fingerprint_chunk(c), first_slot=c * SERVE_CHUNK)
    assert store.info()["size"] == SERVE_SIZE
    owners = [5, 19999, 12345, -1, 777, -1, 10001, 4999]
    g = torch.Generator(device="cuda").manual_seed(77)
    q = torch.randn((len(owners), SERVE_N), generator=g, device="cuda")
    rows = {}
    for i, o in enumerate(owners):
        if o >= 0:
            rows[o] = _fingerprint_chunk(o // SERVE_CHUNK)[o % SERVE_CHUNK].clone()
            q[i] += 0.3 * rows[o]
    q = (q - q.mean(dim=1, keepdim=True)).contiguous()
    ctx = csi._Ctx(1024, 1024, _lib.MODE_NCC)
    idx, score = (t.cpu().numpy() for t in store.query(ctx, q, k=10))
    qh = q.cpu().numpy()
    for r, o in enumerate(owners):
        if o >= 0:
            assert idx[r, 0] == o, f"query {r}: owner {o} not first"
            want = O.ncc_reference(qh[r], rows[o].cpu().numpy())
            assert score[r, 0] == pytest.approx(want, rel=1e-10)
        assert np.all(np.diff(score[r]) <= 0)
    # exact re-scores of other candidates: fingerprints regenerated from their chunk
    for r, c in ((3, 0), (3, 9), (5, 4), (0, 1)):
        slot = int(idx[r, c])
        fp = _fingerprint_chunk(slot // SERVE_CHUNK)[slot % SERVE_CHUNK].cpu().numpy()
        assert score[r, c] == pytest.approx(O.ncc_reference(qh[r], fp), rel=1e-10, abs=1e-13)
    # alone and in the mix
    for r in (0, 3, 7):
        i1, s1 = (t.cpu().numpy() for t in store.query(ctx, q[r:r + 1], k=10))
        assert np.array_equal(i1[0], idx[r]) and np.array_equal(s1[0], score[r])
    i2, s2 = (t.cpu().numpy() for t in store.query(ctx, q[2:6], k=10))
    assert np.array_equal(i2, idx[2:6]) and np.array_equal(s2, score[2:6])
    ctx.close()
    store.close()
