"""prnu_pattern_stats_noalloc (statistics with the partial sums in a fixed
device array, for callers that cannot allocate) through the C ABI: bitwise
equal to prnu_pattern_stats on two calls in a row, the second with fewer and
shorter patterns (the shared array keeps the first call's partials, which the
second must not read), within the FP64 oracle's tolerance, and a call beyond
the array's capacity refused with a clear error."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import prnu_oracle as O  # noqa: E402
from rocket_amd import _lib, csi  # noqa: E402

pytestmark = pytest.mark.gpu


def _stats(fn, pats):
    lib = _lib.load()
    count = pats.shape[0]
    ptrs = csi.pattern_pointer_table(pats)
    s = torch.full((count,), float("nan"), dtype=torch.float64, device="cuda")
    ss = torch.full((count,), float("nan"), dtype=torch.float64, device="cuda")
    am = torch.full((count,), float("nan"), dtype=torch.float32, device="cuda")
    _lib.check(getattr(lib, fn)(ptrs.data_ptr(), count, pats.shape[1], s.data_ptr(), ss.data_ptr(), am.data_ptr(),
                                torch.cuda.current_stream().cuda_stream), fn)
    torch.cuda.current_stream().synchronize()
    return s.cpu().numpy(), ss.cpu().numpy(), am.cpu().numpy()


def test_noalloc_stats_equal_workspace_stats():
    g = torch.Generator(device="cuda").manual_seed(7)
    for count, n in ((40, 2048 * 1536), (7, 61 * 77)):        # the second call is smaller
        print(f"noalloc stats: {count} patterns of {n} values", flush=True)
        p = torch.randn((count, n), generator=g, device="cuda") + 0.125
        got = _stats("prnu_pattern_stats_noalloc", p)
        want = _stats("prnu_pattern_stats", p)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
        host = p.cpu().numpy()
        for i in (0, count - 1):
            sx, sxx = O.pattern_sums(host[i])
            assert got[0][i] == pytest.approx(sx, rel=1e-10, abs=1e-9)
            assert got[1][i] == pytest.approx(sxx, rel=1e-10)
            assert got[2][i] == np.abs(host[i]).max()


def test_noalloc_stats_refuses_more_than_the_scratch():
    lib = _lib.load()
    p = torch.zeros((1, 64), device="cuda")
    ptrs = csi.pattern_pointer_table(p)
    out = torch.zeros(4096, dtype=torch.float64, device="cuda")
    # 2731 patterns of 2048x1536 need 2731 * 48 > 131072 partials: refused on
    # the host before any launch (the pointer table is never read)
    st = lib.prnu_pattern_stats_noalloc(ptrs.data_ptr(), 2731, 2048 * 1536, out.data_ptr(), out.data_ptr(), None,
                                        torch.cuda.current_stream().cuda_stream)
    assert st == _lib.EINVAL
    assert b"at most 2730 patterns" in lib.prnu_last_error()
