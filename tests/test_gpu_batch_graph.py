"""Many H-matrices in one launch, HIP-graph replay, plan configurations and
the autotuner, against the FP64 oracle."""
import numpy as np
import pytest

from tests.helpers import build, c1, c4_members, cauchy, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def members(hm):
    return c4_members(hm, nmat=12, nlo=300, nhi=6000, seed=6)


def test_batch_one_launch_matches_oracle(hm, orc, members):
    """HM_ONE_LAUNCH: the single-right-hand-side product of a batch runs as
    one launch (projections, reductions and row tiles ordered largest member
    first, row tiles waiting on per-member counters); every member equals the oracle,
    for H and H', with alpha/beta, over products in a row (the counters
    reset themselves), and no bounded wait ran out."""
    import torch
    d = hm.DeviceHmat.batch(members, "f32", {"one_launch": True})
    assert d.info()["fused"] == 1
    rng = np.random.default_rng(10)
    for rep in range(3):
        xs = [rng.standard_normal(m.ncols) for m in members]
        x = torch.from_numpy(np.concatenate(xs)).float().cuda()
        y = d.matvec(x).cpu().numpy()
        for i, m in enumerate(members):
            yi = y[d.row_offsets[i]:d.row_offsets[i + 1]]
            assert relerr(yi, orc.matvec(m, xs[i])) <= 2e-6, (rep, i)
    # H' over the same members, and y = 2 H x - 0.5 y
    zs = [rng.standard_normal(m.nrows) for m in members]
    z = torch.from_numpy(np.concatenate(zs)).float().cuda()
    w = d.matvec(z, transpose=True).cpu().numpy()
    for i, m in enumerate(members):
        assert relerr(w[d.col_offsets[i]:d.col_offsets[i + 1]], orc.matvec_t(m, zs[i])) <= 2e-6, i
    y0 = [rng.standard_normal(m.nrows) for m in members]
    yd = torch.from_numpy(np.concatenate(y0)).float().cuda()
    d.matvec(x, yd, alpha=2.0, beta=-0.5)
    yv = yd.cpu().numpy()
    for i, m in enumerate(members):
        ref = 2.0 * orc.matvec(m, xs[i]) - 0.5 * y0[i]
        assert relerr(yv[d.row_offsets[i]:d.row_offsets[i + 1]], ref) <= 4e-6, i
    assert d.info()["wait_faults"] == 0


def test_batch_three_launch_form(hm, orc, members):
    """The default form: the same batch as projection, reduction and row
    tile launches; every member equals the oracle, H and H'."""
    import torch
    d = hm.DeviceHmat.batch(members, "f32")
    assert d.info()["fused"] == 0
    rng = np.random.default_rng(13)
    xs = [rng.standard_normal(m.ncols) for m in members]
    y = d.matvec(torch.from_numpy(np.concatenate(xs)).float().cuda()).cpu().numpy()
    zs = [rng.standard_normal(m.nrows) for m in members]
    w = d.matvec(torch.from_numpy(np.concatenate(zs)).float().cuda(), transpose=True).cpu().numpy()
    for i, m in enumerate(members):
        assert relerr(y[d.row_offsets[i]:d.row_offsets[i + 1]], orc.matvec(m, xs[i])) <= 2e-6, i
        assert relerr(w[d.col_offsets[i]:d.col_offsets[i + 1]], orc.matvec_t(m, zs[i])) <= 2e-6, i


def test_batch_f64_bitwise_close(hm, orc, members):
    import torch
    d = hm.DeviceHmat.batch(members[:5], "f64")
    rng = np.random.default_rng(11)
    xs = [rng.standard_normal(m.ncols) for m in members[:5]]
    y = d.matvec(torch.from_numpy(np.concatenate(xs)).cuda()).cpu().numpy()
    for i, m in enumerate(members[:5]):
        assert relerr(y[d.row_offsets[i]:d.row_offsets[i + 1]], orc.matvec(m, xs[i])) <= 1e-13


@pytest.mark.parametrize("one_launch", [False, True])
def test_graph_replay_matches_eager(hm, orc, members, one_launch):
    import torch
    d = hm.DeviceHmat.batch(members, "f32", {"one_launch": one_launch})
    x = torch.zeros(d.ncols, dtype=torch.float32, device="cuda")
    y = torch.zeros(d.nrows, dtype=torch.float32, device="cuda")
    g = d.graph(x, y)
    for rep in range(4):
        x.copy_(torch.randn(d.ncols, generator=torch.Generator().manual_seed(rep)).cuda())
        ref = d.matvec(x).clone()            # eager product on the same input
        torch.cuda.synchronize()
        y.fill_(float("nan"))
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(y, ref), rep
    g.close()


@pytest.mark.parametrize("shape", [0, 2])
def test_graph_batched_rhs_and_transpose(hm, shape):
    """Graph replays of the batched H X and H' X equal eager products
    bitwise, three times in a row, for the automatic row-pass shape and the
    wide one (1024-thread workgroups)."""
    import torch
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    d = hm.DeviceHmat(f, "bf16x2", {"strip_shape": shape})
    X = d.new_cols(f.ncols, 32)
    Y = d.new_cols(f.nrows, 32)
    gN = d.graph(X, Y)
    Xt = d.new_cols(f.nrows, 32)
    Yt = d.new_cols(f.ncols, 32)
    gT = d.graph(Xt, Yt, transpose=True)
    for rep in range(3):
        X.copy_(torch.randn(f.ncols, 32))
        Xt.copy_(torch.randn(f.nrows, 32))
        refN = d.matmat(X).clone()
        refT = d.matmat_t(Xt).clone()
        gN.launch()
        gT.launch()
        torch.cuda.synchronize()
        assert torch.equal(Y, refN) and torch.equal(Yt, refT), rep


@pytest.mark.parametrize("cfg", [{"proj_chunk": 256, "row_waves": 2}, {"proj_chunk": 4096, "row_waves": 8},
                                 {"proj_chunk": 64, "row_waves": 4}])
def test_configs_keep_fp64_parity(hm, orc, cfg):
    import torch
    par, c, H = build(hm, np.sort(np.random.default_rng(1).uniform(0, 100, 20000)), cauchy, 1024, 64, 16)
    f = hm.flatten(H)
    d = hm.DeviceHmat(f, "f64", cfg)
    assert d.info()["proj_chunk"] == cfg["proj_chunk"] and d.info()["row_waves"] == cfg["row_waves"]
    x = np.random.default_rng(2).standard_normal(f.ncols)
    assert relerr(d.matvec(torch.from_numpy(x).cuda()).cpu().numpy(), orc.matvec(f, x)) <= 1e-13
    v = np.random.default_rng(3).standard_normal(f.nrows)
    assert relerr(d.matvec(torch.from_numpy(v).cuda(), transpose=True).cpu().numpy(), orc.matvec_t(f, v)) <= 1e-13


def test_no_transpose_flag(hm):
    import torch
    par, c, H = c1(hm, "aca2")
    d = hm.DeviceHmat(hm.flatten(H), "f64", {"no_transpose": True})
    from hmatrix_amd._native import HmError
    with pytest.raises(HmError) as e:
        d.matvec(torch.zeros(2000, dtype=torch.float64, device="cuda"), transpose=True)
    assert e.value.code == -4


def test_products_recorded_in_a_caller_graph(hm, orc, members):
    """The product entry points record into a caller's stream capture
    (torch.cuda.graph): four products of the batch (H twice, H', and
    y = 2 H x - 0.5 y) in one graph, for both launch forms.  Every replay,
    with eager work and synchronizations between replays, equals the eager
    products bitwise, and the eager products equal the oracle."""
    import torch
    rng = np.random.default_rng(21)
    for one in (False, True):
        d = hm.DeviceHmat.batch(members, "f32", {"one_launch": one})
        xh = [rng.standard_normal(d.ncols) for _ in range(2)]
        xs = [torch.from_numpy(v).float().cuda() for v in xh]
        z = torch.from_numpy(rng.standard_normal(d.nrows)).float().cuda()
        y0 = torch.from_numpy(rng.standard_normal(d.nrows)).float().cuda()
        ys = [torch.zeros(d.nrows, device="cuda") for _ in range(2)]
        w = torch.zeros(d.ncols, device="cuda")
        ya = torch.empty(d.nrows, device="cuda")
        outs = (ys[0], ys[1], w, ya)

        def products(stream):
            d.matvec(xs[0], ys[0], stream=stream)
            d.matvec(xs[1], ys[1], stream=stream)
            d.matvec(z, w, transpose=True, stream=stream)
            with torch.cuda.stream(stream):
                ya.copy_(y0)
            d.matvec(xs[0], ya, alpha=2.0, beta=-0.5, stream=stream)

        products(torch.cuda.current_stream())
        torch.cuda.synchronize()
        ref = [t.clone() for t in outs]
        y0h = y0.cpu().numpy().astype(np.float64)
        for i, m in enumerate(members):
            r0, r1 = d.row_offsets[i], d.row_offsets[i + 1]
            assert relerr(ref[1][r0:r1].cpu().numpy(), orc.matvec(m, xh[1][d.col_offsets[i]:d.col_offsets[i + 1]])) <= 2e-6
            want = 2.0 * orc.matvec(m, xh[0][d.col_offsets[i]:d.col_offsets[i + 1]]) - 0.5 * y0h[r0:r1]
            assert relerr(ref[3][r0:r1].cpu().numpy(), want) <= 4e-6
        s = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            products(s)
        for rep in range(3):
            for t in outs:
                t.fill_(float("nan"))          # eager work between the replays
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            for a, b in zip(outs, ref):
                assert torch.equal(a, b), (one, rep)
        del g
        assert d.info()["wait_faults"] == 0
        d.close()


@pytest.mark.parametrize("dtype", ["bf16x2", "f64"])
def test_multi_rhs_products_recorded_in_a_caller_graph(hm, orc, dtype):
    """The multi-right-hand-side path (split BF16 and the FP64 f64-MFMA
    layout) recorded into a caller's graph: H X1 (32 columns), H' X2 (13
    columns) and Y = 2 H X1 - 0.5 Y0 in one graph; every replay, with eager
    work between replays, equals the eager products bitwise, and the eager
    products equal the oracle."""
    import torch
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    d = hm.DeviceHmat(f, dtype)
    dt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-13 if dtype == "f64" else 2e-5
    rng = np.random.default_rng(23)
    X1h, X2h, Y0h = rng.standard_normal((f.ncols, 32)), rng.standard_normal((f.nrows, 13)), rng.standard_normal((f.nrows, 32))

    def cols(a, n, k):
        t = d.new_cols(n, k)
        t.copy_(torch.from_numpy(a).to(dt))
        return t

    X1, X2, Y0 = cols(X1h, f.ncols, 32), cols(X2h, f.nrows, 13), cols(Y0h, f.nrows, 32)
    Y1, W2, Ya = d.new_cols(f.nrows, 32), d.new_cols(f.ncols, 13), d.new_cols(f.nrows, 32)
    outs = (Y1, W2, Ya)

    def products(stream):
        d.matmat(X1, Y1, stream=stream)
        d.matmat(X2, W2, transpose=True, stream=stream)
        with torch.cuda.stream(stream):
            Ya.copy_(Y0)
        d.matmat(X1, Ya, alpha=2.0, beta=-0.5, stream=stream)

    products(torch.cuda.current_stream())
    torch.cuda.synchronize()
    ref = [t.clone() for t in outs]
    assert relerr(ref[0].cpu().numpy(), orc.matmat(f, X1h)) <= tol
    assert relerr(ref[1].cpu().numpy(), np.stack([orc.matvec_t(f, X2h[:, i]) for i in range(13)], 1)) <= tol
    assert relerr(ref[2].cpu().numpy(), 2.0 * orc.matmat(f, X1h) - 0.5 * Y0h) <= 2 * tol
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        products(s)
    for rep in range(3):
        for t in outs:
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(outs, ref):
            assert torch.equal(a, b), rep
    del g
    d.close()


def test_autotune_picks_a_config(hm, tmp_path):
    from hmatrix_amd.autotune import autotune
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    best, t = autotune(f, "f64", chunks=(256, 1024), waves=(2, 4), reps=3, cache=str(tmp_path / "at.json"))
    assert best["proj_chunk"] in (256, 1024) and best["row_waves"] in (2, 4) and len(t) == 4
    best2, t2 = autotune(f, "f64", chunks=(256, 1024), waves=(2, 4), reps=3, cache=str(tmp_path / "at.json"))
    assert best2 == best and t2 == t     # served from the cache


def test_batch_autotune_keeps_parity(hm, orc, members, tmp_path):
    """The many-matrix autotuner times proj_chunk x row_waves x (three
    launches, one launch), caches the result, and the plan it picks gives
    every member the oracle's product."""
    import torch
    from hmatrix_amd.autotune import autotune_batch
    cache = str(tmp_path / "atbatch.json")
    best, t = autotune_batch(members, "f32", chunks=(256, 1024), waves=(2, 4), reps=3, cache=cache)
    assert len(t) == 8 and best["proj_chunk"] in (256, 1024) and best["row_waves"] in (2, 4)
    assert autotune_batch(members, "f32", chunks=(256, 1024), waves=(2, 4), reps=3, cache=cache) == (best, t)
    d = hm.DeviceHmat.batch(members, "f32", dict(best))
    assert d.info()["fused"] == int(best["one_launch"])
    rng = np.random.default_rng(17)
    xs = [rng.standard_normal(m.ncols) for m in members]
    y = d.matvec(torch.from_numpy(np.concatenate(xs)).float().cuda()).cpu().numpy()
    for i, m in enumerate(members):
        assert relerr(y[d.row_offsets[i]:d.row_offsets[i + 1]], orc.matvec(m, xs[i])) <= 2e-6, i
    d.close()


def test_strip_shapes_agree_and_autotune(hm, orc, tmp_path):
    """Both multi-right-hand-side row-pass shapes, forced through
    hm_config.strip_shape, match the oracle (strong admissibility, so the
    narrow shape reads some X windows straight from memory); plan info
    reports the shape in use; the batched autotuner times both and caches."""
    import torch
    from hmatrix_amd.autotune import autotune_batched
    from tests.helpers import build, cauchy, sorted_uniform
    X = sorted_uniform(12000, 0.0, 40.0, 31)
    par, c, H = build(hm, X, cauchy, 2048, 128, 10, adj="strong")
    f = hm.flatten(H)
    R = np.random.default_rng(32).standard_normal((f.ncols, 32))
    ref = orc.matmat(f, R)
    for shape in (1, 2):
        d = hm.DeviceHmat(f, "bf16x2", {"strip_shape": shape})
        assert d.info()["strip_shape"] == shape
        Xd = d.new_cols(f.ncols, 32)
        Xd.copy_(torch.from_numpy(R).float())
        Y = d.matmat(Xd).cpu().numpy().astype(np.float64)
        assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) <= 2e-5
        d.close()
    auto = hm.DeviceHmat(f, "bf16x2")
    assert auto.info()["strip_shape"] in (1, 2)
    auto.close()
    best, t = autotune_batched(f, "bf16x2", reps=3, cache=str(tmp_path / "atb.json"))
    assert best["strip_shape"] in (1, 2) and len(t) == 2
    assert autotune_batched(f, "bf16x2", reps=3, cache=str(tmp_path / "atb.json")) == (best, t)


def test_native_autotune(hm, orc):
    """hm_autotune (the C entry point) picks a candidate for one and for 32
    right-hand sides, and the chosen configuration keeps parity."""
    import torch
    from hmatrix_amd.autotune import autotune_native
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    cfg, ms = autotune_native(f, "f64", nrhs=1, reps=3)
    assert cfg["proj_chunk"] in (256, 512, 1024, 2048, 4096) and cfg["row_waves"] in (2, 4, 8) and ms > 0
    d = hm.DeviceHmat(f, "f64", cfg)
    x = np.random.default_rng(5).standard_normal(f.ncols)
    assert relerr(d.matvec(torch.from_numpy(x).cuda()).cpu().numpy(), orc.matvec(f, x)) <= 1e-13
    d.close()
    cfg2, ms2 = autotune_native(f, "bf16x2", nrhs=32, reps=3)
    assert cfg2["strip_shape"] in (1, 2) and ms2 > 0
    # FP64 with several columns: the f64 MFMA layout or single-layout panels
    cfg3, ms3 = autotune_native(f, "f64", nrhs=32, reps=3)
    assert ms3 > 0 and (cfg3.get("strip_shape") == 0 or cfg3.get("no_batched")), cfg3
    d = hm.DeviceHmat(f, "f64", cfg3)
    X = np.random.default_rng(6).standard_normal((f.ncols, 32))
    Xd = d.new_cols(f.ncols, 32)
    Xd.copy_(torch.from_numpy(X))
    assert relerr(d.matmat(Xd).cpu().numpy(), orc.matmat(f, X)) <= 1e-13
    d.close()


def test_graph_fp64_multi_rhs(hm):
    """A captured FP64 H X of 10 columns (two panels on the single-right-hand-
    side layout) replays bitwise equal to the eager product, three times."""
    import torch
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    d = hm.DeviceHmat(f, "f64")
    X = d.new_cols(f.ncols, 10)
    Y = d.new_cols(f.nrows, 10)
    g = d.graph(X, Y)
    for rep in range(3):
        X.copy_(torch.randn(f.ncols, 10, dtype=torch.float64, generator=torch.Generator().manual_seed(rep)))
        ref = d.matmat(X).clone()
        torch.cuda.synchronize()
        Y.fill_(float("nan"))
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(Y, ref), rep
    g.close()
    d.close()
