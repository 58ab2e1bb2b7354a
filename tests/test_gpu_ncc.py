"""GPU parity of the NCC path (HIP via the C ABI) against the FP64 oracle.

Tolerances:
* per-pattern sums and the per-pair reference comparator: the same FP32
  products summed in FP64 in another order -> relative 1e-10;
* Gram NCC in FP32 MFMA and split-BF16 MFMA: max |dNCC| <= 1e-5 against the
  FP64 oracle (the target stated for the path).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import prnu_oracle as O  # noqa: E402
from rocket_amd import _lib, csi  # noqa: E402

pytestmark = pytest.mark.gpu

NCC_TOL = 1e-5


def patterns(count, h, w, cams, seed, gain=0.05):
    """Zero-mean FP32 patterns: a camera fingerprint times `gain` plus unit noise."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    n = h * w
    k = torch.randn((cams, n), generator=g, device="cuda")
    p = torch.randn((count, n), generator=g, device="cuda")
    p += gain * k[torch.arange(count, device="cuda") % cams]
    p -= p.mean(dim=1, keepdim=True)
    return p.float().contiguous()


def test_library_loads():
    lib = _lib.load()
    assert lib.prnu_api_version() == 1
    n = _lib.ctypes.c_int()
    _lib.check(lib.prnu_device_count(_lib.ctypes.byref(n)))
    assert n.value >= 1


@pytest.mark.parametrize("h,w", [(97, 131), (64, 64)])
def test_pattern_stats_match_oracle(h, w):
    p = patterns(5, h, w, 2, seed=1) + 0.25
    gram = csi.GramNCC(h, w)
    ptrs = csi.pattern_pointer_table(p)
    s, ss, am = gram.stats(ptrs, 5)
    host = p.cpu().numpy()
    for i in range(5):
        sx, sxx = O.pattern_sums(host[i])
        assert s[i].item() == pytest.approx(sx, rel=1e-10, abs=1e-9)
        assert ss[i].item() == pytest.approx(sxx, rel=1e-10)
        assert am[i].item() == np.abs(host[i]).max()
    gram.close()


def test_pattern_stats_unaligned_pointer():
    h, w = 33, 31
    buf = patterns(1, h * w + 1, 1, 1, seed=2).view(-1)
    view = buf[1:]                                    # 4-byte aligned, not 16
    gram = csi.GramNCC(h, w)
    s, ss, _ = gram.stats(csi.pattern_pointer_table([view]), 1)
    sx, sxx = O.pattern_sums(view.cpu().numpy())
    assert ss[0].item() == pytest.approx(sxx, rel=1e-10)
    assert s[0].item() == pytest.approx(sx, rel=1e-9, abs=1e-9)
    gram.close()


@pytest.mark.parametrize("h,w", [(512, 512), (61, 77)])
def test_pair_correlate_reference_semantics(h, w):
    p = patterns(3, h, w, 2, seed=3)
    ctx = csi.CommonSourceIdentificationContext(h, w, csi.ComparisonStrategy.NCC)
    assert ctx.get_max_input_size() == 4 * h * w
    assert ctx.get_max_output_size() == 8
    assert ctx.get_max_file_size() == 3 * h * w
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    host = p.cpu().numpy()
    for i, j in [(0, 1), (0, 2), (1, 1)]:
        assert ctx.correlate_gpu(i, p[i], j, p[j], out) == 8
        want = O.ncc_reference(host[i], host[j])
        assert out.item() == pytest.approx(want, rel=1e-10, abs=1e-12)
    assert ctx.postprocess_output(0, 1, out.cpu().numpy().tobytes()) == out.item()
    ctx.destroy()


def test_pair_correlate_rejects_undersized():
    h, w = 32, 32
    ctx = csi.CommonSourceIdentificationContext(h, w, csi.ComparisonStrategy.NCC)
    a = torch.zeros(h * w, device="cuda")
    small = torch.zeros(h * w - 1, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        ctx.correlate_gpu(0, a, 1, small, out)
    with pytest.raises(ValueError):
        ctx.correlate_gpu(0, a, 1, a, torch.zeros(1, dtype=torch.float32, device="cuda"))
    ctx.destroy()


@pytest.mark.parametrize("precision", ["fp32", "split_bf16"])
@pytest.mark.parametrize("count,h,w", [(64, 512, 512), (300, 200, 321), (7, 16, 9)])
def test_gram_matches_fp64_oracle(precision, count, h, w):
    p = patterns(count, h, w, 4, seed=10 + count, gain=0.2)
    gram = csi.GramNCC(h, w)
    got = gram.ncc(p, precision).cpu().numpy()
    want = O.ncc_matrix(p.cpu().numpy())
    err = np.abs(got - want).max()
    assert err <= NCC_TOL, f"{precision}: max |dNCC| = {err:.3e}"
    assert np.array_equal(got, got.T)
    # same-camera pairs (NCC ~ 0.04) separate from the rest (noise ~ 1/sqrt(n))
    cam = np.arange(count) % 4
    same = got[(cam[:, None] == cam[None, :]) & ~np.eye(count, dtype=bool)]
    diff = got[cam[:, None] != cam[None, :]]
    if count >= 64:
        assert same.min() > diff.max()
    gram.close()


def test_gram_patterns_as_pointer_list():
    h, w, count = 48, 40, 9
    p = patterns(count, h, w, 3, seed=21)
    rows = [p[i].clone() for i in range(count)]       # separate allocations
    gram = csi.GramNCC(h, w)
    got = gram.ncc(rows, "split_bf16").cpu().numpy()
    want = O.ncc_matrix(p.cpu().numpy())
    assert np.abs(got - want).max() <= NCC_TOL
    gram.close()


def test_gram_full_size_properties():
    """1024 patterns of 2048x1536: split-BF16 vs FP32 agree, symmetric, unit
    diagonal, and a 48-pattern sample matches the FP64 oracle."""
    h, w, count = 1536, 2048, 1024
    p = patterns(count, h, w, 16, seed=99, gain=0.02)
    gram = csi.GramNCC(h, w)
    stats = gram.stats(csi.pattern_pointer_table(p), count)
    bf = gram.ncc(p, "split_bf16", stats=stats).cpu().numpy()
    fp = gram.ncc(p, "fp32", stats=stats).cpu().numpy()
    assert np.array_equal(bf, bf.T)
    assert np.abs(bf - fp).max() <= NCC_TOL
    assert np.abs(np.diag(bf) - 1.0).max() <= 1e-5
    idx = np.random.default_rng(0).choice(count, 48, replace=False)
    want = O.ncc_matrix(p[torch.as_tensor(idx, device="cuda")].cpu().numpy())
    err = np.abs(bf[np.ix_(idx, idx)] - want).max()
    assert err <= NCC_TOL, f"max |dNCC| = {err:.3e}"
    gram.close()


@pytest.mark.parametrize("count,h,w", [(64, 512, 512), (300, 200, 321), (7, 16, 9)])
def test_gram_fp8_within_e4m3_bound(count, h, w):
    """FP8-e4m3 Gram with per-pattern scale.  e4m3 rounds to 3 mantissa bits
    (relative error <= 2^-4, rms ~0.036 for uniform mantissas), so an
    off-diagonal NCC of n-sample patterns errs by ~0.05/sqrt(n) rms; the bound
    is 6 sigma.  The diagonal carries the E[eps^2] bias (~1.3e-3) plus a
    fluctuation of 0.036*sqrt(3/n) rms (sum eps x^2 / sum x^2 for Gaussian x)."""
    p = patterns(count, h, w, 4, seed=30 + count, gain=0.2)
    gram = csi.GramNCC(h, w)
    got = gram.ncc(p, "fp8").cpu().numpy()
    want = O.ncc_matrix(p.cpu().numpy())
    off = ~np.eye(count, dtype=bool)
    n = h * w
    assert np.abs(got - want)[off].max() <= 6 * 0.05 / np.sqrt(n)
    assert np.abs(np.diag(got) - np.diag(want)).max() <= 1.5e-3 + 6 * 0.036 * np.sqrt(3.0 / n)
    assert np.array_equal(got, got.T)
    gram.close()


def test_gram_config_and_autotune():
    h, w, count = 256, 256, 300
    p = patterns(count, h, w, 4, seed=41)
    gram = csi.GramNCC(h, w)
    ptrs = csi.pattern_pointer_table(p)
    lib = _lib.load()
    cfg = (_lib.ctypes.c_int32 * 3)()
    _lib.check(lib.prnu_gram_get_config(gram._ctx.handle, _lib.PREC_SPLIT_BF16, count, cfg))
    assert cfg[0] >= 1 and cfg[1] == 1
    ref = gram.ncc(p, "split_bf16").cpu().numpy()
    for slabs, xcd, var in [(8, 0, 0), (16, 1, 0), (3, 0, 1), (16, 1, 1), (1, 1, 1)]:
        c = (_lib.ctypes.c_int32 * 3)(slabs, xcd, var)
        _lib.check(lib.prnu_gram_set_config(gram._ctx.handle, _lib.PREC_SPLIT_BF16, c))
        got = gram.ncc(p, "split_bf16").cpu().numpy()
        assert np.abs(got - ref).max() <= NCC_TOL        # only the slab summation order changes
    best = (_lib.ctypes.c_int32 * 3)()
    ms = _lib.ctypes.c_float()
    _lib.check(lib.prnu_gram_autotune(gram._ctx.handle, ptrs.data_ptr(), count, _lib.PREC_SPLIT_BF16, 2, best,
                                      _lib.ctypes.byref(ms), torch.cuda.current_stream().cuda_stream))
    assert best[0] >= 1 and ms.value > 0
    got = gram.ncc(p, "split_bf16").cpu().numpy()
    assert np.abs(got - ref).max() <= NCC_TOL
    gram.close()


@pytest.mark.parametrize("fmt", ["split_bf16", "fp32"])
@pytest.mark.parametrize("count,h,w,first,cap", [(64, 512, 512, 0, 64), (300, 200, 321, 5, 310), (7, 16, 9, 3, 12),
                                                 (840, 40, 50, 2, 845), (1100, 32, 33, 0, 1100)])
def test_store_gram_matches_fp64_oracle(fmt, count, h, w, first, cap):
    """Ragged counts with an odd number of 128-pattern quarters (300, 840,
    1100: the edge and one-quarter items of the packing) and an even one."""
    p = patterns(count, h, w, 4, seed=50 + count, gain=0.2)
    store = csi.PatternStore(h * w, cap, fmt)
    store.put(p, first_slot=first)
    s, ss, am = store.stats(count, first)
    host = p.cpu().numpy()
    for i in range(0, count, max(1, count // 5)):
        sx, sxx = O.pattern_sums(host[i])
        assert ss[i] == pytest.approx(sxx, rel=1e-10) and s[i] == pytest.approx(sx, rel=1e-9, abs=1e-9)
        assert am[i] == np.abs(host[i]).max()
    gram = csi.GramNCC(h, w)
    got = store.gram(gram, count, first).cpu().numpy()
    want = O.ncc_matrix(host)
    err = np.abs(got - want).max()
    assert err <= NCC_TOL, f"{fmt}: max |dNCC| = {err:.3e}"
    assert np.array_equal(got, got.T)
    gram.close()
    store.close()


def test_store_full_size_properties():
    """1024 x 2048x1536 through the blocked store: split-BF16 vs FP32 store,
    vs the pointer-table Gram, and a 48-pattern FP64 sample."""
    h, w, count = 1536, 2048, 1024
    p = patterns(count, h, w, 16, seed=77, gain=0.02)
    gram = csi.GramNCC(h, w)
    sb = csi.PatternStore(h * w, count, "split_bf16")
    sb.put(p)
    bf = sb.gram(gram, count).cpu().numpy()
    sb.close()
    sf = csi.PatternStore(h * w, count, "fp32")
    sf.put(p)
    fp = sf.gram(gram, count).cpu().numpy()
    sf.close()
    assert np.array_equal(bf, bf.T) and np.abs(bf - fp).max() <= NCC_TOL
    ptr = gram.ncc(p, "split_bf16").cpu().numpy()
    assert np.abs(bf - ptr).max() <= NCC_TOL
    idx = np.random.default_rng(1).choice(count, 48, replace=False)
    want = O.ncc_matrix(p[torch.as_tensor(idx, device="cuda")].cpu().numpy())
    assert np.abs(bf[np.ix_(idx, idx)] - want).max() <= NCC_TOL
    gram.close()
