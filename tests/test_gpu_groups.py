"""GPU tests of the grouped all-pairs plan (several image sizes in one launch).

* every group's output matches the FP64 oracle NCC matrix of that group
  (|dNCC| <= 1e-5, the bar of the batched path) and the single-group store
  Gram of the same slots;
* groups with ragged counts (not multiples of the 256 tile), groups sharing a
  store through slot ranges, FP32 and split-BF16 stores;
* HIP-graph replays are bitwise equal to plain launches, replay after replay,
  with unrelated eager work and synchronisations between the replays;
* the autotuner keeps a slab length whose results stay within the bar;
* invalid plans are rejected with ValueError.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import prnu_oracle as O  # noqa: E402
from rocket_amd import csi  # noqa: E402

pytestmark = pytest.mark.gpu

NCC_TOL = 1e-5


def patterns(count, n, cams, seed, gain=0.2):
    g = torch.Generator(device="cuda").manual_seed(seed)
    k = torch.randn((cams, n), generator=g, device="cuda")
    p = torch.randn((count, n), generator=g, device="cuda")
    p += gain * k[torch.arange(count, device="cuda") % cams]
    p -= p.mean(dim=1, keepdim=True)
    return p.float().contiguous()


# (n, count, store capacity, first slot): ragged counts, different K
GROUPS = [(96 * 128, 300, 310, 4), (64 * 80, 37, 37, 0), (150 * 200, 513, 520, 7)]


@pytest.fixture(scope="module", params=["split_bf16", "fp32"])
def groups(request):
    fmt = request.param
    out = []
    for i, (n, count, cap, first) in enumerate(GROUPS):
        p = patterns(count, n, 4, seed=300 + i)
        st = csi.PatternStore(n, cap, fmt)
        st.put(p, first_slot=first)
        out.append((st, first, count, p))
    gram = csi.GramNCC(8, 8)   # workspace/timing context; sizes come from the stores
    yield fmt, out, gram
    for st, *_ in out:
        st.close()
    gram.close()


def test_plan_matches_oracle_and_single_group(groups):
    fmt, gs, gram = groups
    # a fourth group: a slot range inside the first store
    st0, first0, _, p0 = gs[0]
    spec = [(st, first, count) for st, first, count, _ in gs] + [(st0, first0 + 100, 190)]
    plan = csi.GramPlan(gram, spec)
    info = plan.info()
    assert info["work_items"] > 0 and info["slab_values"] % 32 == 0
    outs = [o.cpu().numpy() for o in plan.run(use_graph=False)]
    hosts = [p.cpu().numpy() for *_, p in gs] + [p0[100:290].cpu().numpy()]
    for g, (got, host) in enumerate(zip(outs, hosts)):
        want = O.ncc_matrix(host)
        err = np.abs(got - want).max()
        assert err <= NCC_TOL, f"{fmt} group {g}: max |dNCC| = {err:.3e}"
        assert np.array_equal(got, got.T)
    for g, (st, first, count, _) in enumerate(gs):
        single = st.gram(gram, count, first).cpu().numpy()
        assert np.abs(single - outs[g]).max() <= NCC_TOL       # only the slab summation order differs
    plan.close()


def test_plan_graph_replays_bitwise(groups):
    fmt, gs, gram = groups
    plan = csi.GramPlan(gram, [(st, first, count) for st, first, count, _ in gs])
    eager = [o.clone() for o in plan.run(use_graph=False)]
    scratch = torch.zeros(1 << 20, device="cuda")
    for rep in range(4):
        for o in plan.outputs:
            o.fill_(float("nan"))
        got = plan.run(use_graph=True)
        for g, (a, b) in enumerate(zip(got, eager)):
            assert torch.equal(a, b), f"replay {rep} group {g}"
        scratch.add_(1.0)                                   # unrelated eager work between replays
        torch.cuda.synchronize()
    plan.close()


def test_plan_autotune(groups):
    fmt, gs, gram = groups
    plan = csi.GramPlan(gram, [(st, first, count) for st, first, count, _ in gs])
    slab, ms = plan.autotune()
    assert slab > 0 and slab % 32 == 0 and ms > 0
    assert plan.info()["slab_values"] == slab
    outs = [o.cpu().numpy() for o in plan.run(use_graph=True)]
    for g, (st, first, count, p) in enumerate(gs):
        assert np.abs(outs[g] - O.ncc_matrix(p.cpu().numpy())).max() <= NCC_TOL
    plan.close()


def test_plan_rejects_invalid(groups):
    fmt, gs, gram = groups
    st, first, count, _ = gs[0]
    with pytest.raises(ValueError):                          # beyond the store's capacity
        csi.GramPlan(gram, [(st, first, st.capacity)])
    with pytest.raises(ValueError):                          # slab not a multiple of 32
        csi.GramPlan(gram, [(st, first, count)], slab_values=100)
    other = csi.PatternStore(64, 8, "fp32" if fmt == "split_bf16" else "split_bf16")
    with pytest.raises(ValueError):                          # mixed store formats
        csi.GramPlan(gram, [(st, first, count), (other, 0, 8)])
    other.close()
    with pytest.raises(ValueError):
        csi.GramPlan(gram, [])


@pytest.mark.parametrize("variant", [1, 2, 3])
@pytest.mark.parametrize("fmt", ["split_bf16", "fp32"])
def test_store_gram_load_variants(fmt, variant):
    """Store Gram variants (1: register-staged loads, 2: direct loads with
    raised priority, 3: four waves of 128 x 128 for split-BF16) against the
    oracle and bitwise against the default (eight waves, direct global-to-LDS
    loads): every output element sees the same MFMA sequence; ragged count,
    offset slot range."""
    import ctypes
    from rocket_amd import _lib
    n, count, cap, first = 120 * 97, 300, 310, 5
    p = patterns(count, n, 4, seed=911)
    st = csi.PatternStore(n, cap, fmt)
    st.put(p, first_slot=first)
    gram = csi.GramNCC(8, 8)
    base = st.gram(gram, count, first).cpu().numpy()
    prec = _lib.PRECISIONS[fmt]
    cfg = (ctypes.c_int32 * 3)(0, 1, variant)
    _lib.check(_lib.load().prnu_gram_set_config(gram._ctx.handle, prec, cfg))
    got = st.gram(gram, count, first).cpu().numpy()
    plan = csi.GramPlan(gram, [(st, first, count)])
    via_plan = plan.run(use_graph=True)[0].cpu().numpy()
    plan.close()
    _lib.check(_lib.load().prnu_gram_set_config(gram._ctx.handle, prec, None))
    want = O.ncc_matrix(p.cpu().numpy())
    assert np.abs(got - want).max() <= NCC_TOL
    assert np.abs(via_plan - want).max() <= NCC_TOL
    assert np.array_equal(got, base)          # same slabs and order: bitwise equal to the default
    st.close()
    gram.close()


def fp8_bounds(n):
    """e4m3 error bounds of test_gpu_ncc.test_gram_fp8_within_e4m3_bound:
    off-diagonal 6 sigma of 0.05/sqrt(n); diagonal bias + 6 sigma."""
    return 6 * 0.05 / np.sqrt(n), 1.5e-3 + 6 * 0.036 * np.sqrt(3.0 / n)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_fp8_store_gram_within_e4m3_bound(variant):
    """FP8 store (blocks of 128 values, per-pattern scale): n not a multiple
    of 128, ragged count, offset slots; both load variants; close to the
    pointer-table FP8 Gram, which quantises the same way."""
    import ctypes
    from rocket_amd import _lib
    n, count, cap, first = 200 * 321, 300, 310, 3
    p = patterns(count, n, 4, seed=777)
    st = csi.PatternStore(n, cap, "fp8")
    st.put(p, first_slot=first)
    gram = csi.GramNCC(200, 321)
    cfg = (ctypes.c_int32 * 3)(0, 1, variant)
    _lib.check(_lib.load().prnu_gram_set_config(gram._ctx.handle, _lib.PREC_FP8, cfg))
    got = st.gram(gram, count, first).cpu().numpy()
    _lib.check(_lib.load().prnu_gram_set_config(gram._ctx.handle, _lib.PREC_FP8, None))
    want = O.ncc_matrix(p.cpu().numpy())
    off = ~np.eye(count, dtype=bool)
    b_off, b_diag = fp8_bounds(n)
    assert np.abs(got - want)[off].max() <= b_off
    assert np.abs(np.diag(got) - np.diag(want)).max() <= b_diag
    assert np.array_equal(got, got.T)
    ptr = gram.ncc(p, "fp8").cpu().numpy()
    assert np.abs(got - ptr).max() <= NCC_TOL        # same quantisation, different K slabs
    s, ss, am = st.stats(count, first)
    assert np.array_equal(am, np.abs(p.cpu().numpy()).max(axis=1))
    st.close()
    gram.close()


def test_fp8_plan_two_sizes():
    """Grouped plan over two FP8 stores of different sizes equals each store's
    own Gram within the summation-order tolerance."""
    specs = [(64 * 64, 130), (100 * 90, 257)]
    stores, ps = [], []
    for i, (n, count) in enumerate(specs):
        p = patterns(count, n, 3, seed=880 + i)
        st = csi.PatternStore(n, count, "fp8")
        st.put(p)
        stores.append(st)
        ps.append(p)
    gram = csi.GramNCC(8, 8)
    plan = csi.GramPlan(gram, [(st, 0, c) for st, (_, c) in zip(stores, specs)])
    assert plan.info()["slab_values"] % 128 == 0
    outs = [o.cpu().numpy() for o in plan.run(use_graph=True)]
    with pytest.raises(ValueError):                  # FP8 slabs are whole 128-value blocks
        csi.GramPlan(gram, [(stores[0], 0, 130)], slab_values=96)
    plan.close()
    for st, (n, count), p, got in zip(stores, specs, ps, outs):
        single = st.gram(gram, count).cpu().numpy()
        assert np.abs(single - got).max() <= NCC_TOL
        b_off, _ = fp8_bounds(n)
        want = O.ncc_matrix(p.cpu().numpy())
        off = ~np.eye(count, dtype=bool)
        assert np.abs(got - want)[off].max() <= b_off
        st.close()
    gram.close()


def test_block_plan_rectangles_and_index_space_schedule():
    """Rectangle blocks (overlapping slot ranges included) against the FP64
    oracle, and the all-pairs NCC scheduled over an index space of 256-key
    tiles in one launch, bitwise equal to the store Gram with the same slabs."""
    import ctypes
    from rocket_amd import _lib, indexspace
    n, count = 120 * 97, 600
    p = patterns(count, n, 5, seed=4242)
    st = csi.PatternStore(n, count, "split_bf16")
    st.put(p)
    gram = csi.GramNCC(8, 8)
    host = p.cpu().numpy()
    out_a = torch.full((300, 300), -9.0, dtype=torch.float64, device="cuda")
    big = torch.full((64, 400), -9.0, dtype=torch.float64, device="cuda")
    plan = csi.GramPlan(gram, blocks=[(st, 5, 300, 100, 300, out_a), (st, 500, 64, 0, 333, big[:, 50:383])])
    plan.run(use_graph=True)
    plan.close()
    full = O.ncc_matrix(host)
    assert np.abs(out_a.cpu().numpy() - full[5:305, 100:400]).max() <= NCC_TOL
    b = big.cpu().numpy()
    assert np.abs(b[:, 50:383] - full[500:564, 0:333]).max() <= NCC_TOL
    assert np.all(b[:, :50] == -9.0) and np.all(b[:, 383:] == -9.0)      # nothing outside the view
    # index-space schedule vs the store Gram, same K slabs (S = 7)
    nkb = (n + 31) // 32
    S = 7
    kcb = (nkb + S - 1) // S
    cfg = (ctypes.c_int32 * 3)(S, 1, 0)
    _lib.check(_lib.load().prnu_gram_set_config(gram._ctx.handle, _lib.PREC_SPLIT_BF16, cfg))
    ref = st.gram(gram, count).cpu().numpy()
    _lib.check(_lib.load().prnu_gram_set_config(gram._ctx.handle, _lib.PREC_SPLIT_BF16, None))
    got, blocks = indexspace.ncc_by_index_space(gram, st, count, slab_values=kcb * 32)
    assert len(blocks) == 3 * 4 // 2
    iu = np.triu_indices(count, 1)
    assert np.array_equal(got.cpu().numpy()[iu], ref[iu])
    st.close()
    gram.close()


@pytest.mark.parametrize("fmt", ["split_bf16", "fp32", "fp8", "fp16"])
def test_persistent_launch_bitwise_equal(fmt):
    """The persistent launch (one workgroup per CU over per-XCD queues) gives
    the per-item launch's results bitwise, for triangles of several stores and
    rectangle blocks, eager and replayed from a graph twice (the queue
    counters are reset inside the graph)."""
    specs = [(64 * 64, 300), (100 * 90, 257)]
    stores = []
    for i, (n, count) in enumerate(specs):
        p = patterns(count, n, 3, seed=990 + i)
        st = csi.PatternStore(n, count, fmt)
        st.put(p)
        stores.append(st)
    gram = csi.GramNCC(8, 8)
    rect = torch.zeros((100, 200), dtype=torch.float64, device="cuda")
    tri = torch.zeros((257, 257), dtype=torch.float64, device="cuda")
    blocks = [(stores[0], 10, 100, 50, 200, rect), (stores[1], 0, 257, 0, 257, tri)]
    plans = [csi.GramPlan(gram, [(st, 0, c) for st, (_, c) in zip(stores, specs)]), csi.GramPlan(gram, blocks=blocks)]
    for plan in plans:
        assert not plan.persistent
        want = [o.clone() for o in plan.run(use_graph=False)]
        plan.persistent = True
        assert plan.persistent and plan.info()["workgroups"] % 8 == 0
        for use_graph in (False, True, True):
            for o in plan.outputs:
                o.fill_(float("nan"))
            got = plan.run(use_graph=use_graph)
            for g, (a, b) in enumerate(zip(got, want)):
                assert torch.equal(a, b), f"{fmt} graph={use_graph} block {g}"
        plan.close()
    from rocket_amd import _lib
    small = csi.GramPlan(gram, [(stores[0], 0, 8)])
    with pytest.raises(ValueError):                  # launch forms are 0 and 1
        _lib.call("prnu_gram_plan_set_launch", small.handle, 2)
    small.close()
    for st in stores:
        st.close()
    gram.close()


def fp16_bound(n):
    """FP16 rounding (unit roundoff u = 2^-11) of both operands: the NCC error
    of unit-variance patterns has standard deviation about u*sqrt(2/3)/sqrt(n)
    off the diagonal and 2u/sqrt(3n)*sqrt(3) on it; 6 sigma, plus 2e-6 for
    the FP32 accumulation inside the MFMA."""
    u = 2.0 ** -11
    return 6 * u * np.sqrt(2.0 / 3.0) / np.sqrt(n) + 2e-6, 6 * 2 * u / np.sqrt(n) + 2e-6


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_fp16_store_gram_within_rounding_bound(variant):
    """FP16 store (blocks of 64 values, power-of-two scale per pattern): n not
    a multiple of 64, ragged count, offset slots, every load variant; within
    the FP16 rounding bound of the FP64 oracle and of the split-BF16 store."""
    import ctypes
    from rocket_amd import _lib
    n, count, cap, first = 200 * 321, 300, 310, 3
    p = patterns(count, n, 4, seed=1601)
    st = csi.PatternStore(n, cap, "fp16")
    st.put(p, first_slot=first)
    gram = csi.GramNCC(200, 321)
    cfg = (ctypes.c_int32 * 3)(0, 1, variant)
    _lib.check(_lib.load().prnu_gram_set_config(gram._ctx.handle, _lib.PREC_FP16, cfg))
    got = st.gram(gram, count, first).cpu().numpy()
    _lib.check(_lib.load().prnu_gram_set_config(gram._ctx.handle, _lib.PREC_FP16, None))
    want = O.ncc_matrix(p.cpu().numpy())
    off = ~np.eye(count, dtype=bool)
    b_off, b_diag = fp16_bound(n)
    assert np.abs(got - want)[off].max() <= b_off
    assert np.abs(np.diag(got) - np.diag(want)).max() <= b_diag
    assert np.array_equal(got, got.T)
    split = csi.PatternStore(n, cap, "split_bf16")
    split.put(p, first_slot=first)
    ref = split.gram(gram, count, first).cpu().numpy()
    assert np.abs(got - ref)[off].max() <= b_off
    s, ss, am = st.stats(count, first)
    assert np.array_equal(am, np.abs(p.cpu().numpy()).max(axis=1))
    split.close()
    st.close()
    gram.close()


def test_fp16_store_scale_is_exact():
    """The per-pattern scale is a power of two, so patterns multiplied by 2^-20
    or 2^40 (exact in FP32, and their FP32 squares stay normal) store the same
    FP16 values and give bitwise the same NCC."""
    n, count = 96 * 70, 40
    p = patterns(count, n, 2, seed=1602)
    gram = csi.GramNCC(8, 8)
    outs = []
    for k in (0, -20, 40):
        st = csi.PatternStore(n, count, "fp16")
        st.put(p * (2.0 ** k))
        outs.append(st.gram(gram, count).cpu().numpy())
        st.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    gram.close()
