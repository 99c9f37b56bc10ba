"""Parity at the full sizes of the benchmark workloads.

* N = 262144 sorted uniform points on [0, 400] (seed 3), exp(-|x-y|),
  MaxRank 24, MaxBlock N/4, MinBlock 64, weak admissibility, X = 32
  standard-normal columns (seed 4): the batched split-BF16 and FP32 H*X
  and H'X against the FP64 oracle on sampled columns (<= 2e-5), 4096
  seeded rows (seed 5) against the dense FP64 kernel (<= 1e-5; plain BF16
  storage reported at its own bound, <= 5e-3), FP64 H*y against the oracle
  (<= 1e-13), and H\\y on H + 10 I by refinement to a relative residual
  <= 1e-12.  The reference checks A*y ~ H*y against the dense kernel
  (test/harithm.jl:1-12) for the product of src/harithm.jl:317-329.
* 64 independent H-matrices, N log-uniform in [1000, 40000] (seed 6),
  1/(1+((x-y)/l)^2), MaxRank in {4, 8, 16}: one FP32 launch for all of them,
  every member's slice against the oracle (<= 2e-6), and graph replays
  bitwise equal to the eager product.
"""
import numpy as np
import pytest

from tests.helpers import build, c4_members, expk, relerr, sorted_uniform

pytestmark = pytest.mark.gpu

N3 = 262144
COLS = [0, 9, 22, 31]          # sampled right-hand sides for the oracle comparisons


@pytest.fixture(scope="module")
def c3(hm):
    par, c, H = build(hm, sorted_uniform(N3, 0.0, 400.0, 3), expk, -1, 64, 24)
    f = hm.flatten(H)
    R = np.random.default_rng(4).standard_normal((N3, 32))
    print(f"C3: {f.nblocks} leaves {f.counts()}", flush=True)
    return par, c, H, f, R


def _product(hm, f, dtype, R, transpose=False):
    import torch
    d = hm.DeviceHmat(f, dtype)
    assert d.info()["batched"] == 1
    X = d.new_cols(N3, R.shape[1])
    X.copy_(torch.from_numpy(R).to(X.device, torch.float32))
    Y = d.matmat(X, transpose=transpose).cpu().numpy().astype(np.float64)
    d.close()
    return Y


def _dense_rows(par, c, R, rows):
    """FP64 rows of the kernel matrix times R, in cluster order."""
    Xc = np.asarray(c.X)
    out = np.zeros((rows.size, R.shape[1]))
    for s in range(0, rows.size, 128):
        r = rows[s:s + 128]
        out[s:s + 128] = par.Kernel(Xc[r][:, None], Xc[None, :]) @ R
    return out


@pytest.mark.parametrize("dtype", ["bf16x2", "f32"])
def test_c3_batched_vs_oracle(hm, orc, c3, dtype):
    par, c, H, f, R = c3
    Y = _product(hm, f, dtype, R)
    assert np.isfinite(Y).all()
    ref = np.stack([orc.matvec(f, R[:, j]) for j in COLS], 1)
    err = relerr(Y[:, COLS], ref)
    print(f"C3 {dtype} H*X vs oracle (columns {COLS}): {err:.3e}", flush=True)
    assert err <= 2e-5


@pytest.mark.parametrize("dtype", ["bf16x2", "f32"])
def test_c3_batched_transpose_vs_oracle(hm, orc, c3, dtype):
    par, c, H, f, R = c3
    Y = _product(hm, f, dtype, R, transpose=True)
    ref = np.stack([orc.matvec_t(f, R[:, j]) for j in COLS], 1)
    err = relerr(Y[:, COLS], ref)
    print(f"C3 {dtype} H'X vs oracle: {err:.3e}", flush=True)
    assert err <= 2e-5


@pytest.fixture(scope="module")
def c3_dense(c3):
    par, c, H, f, R = c3
    rows = np.sort(np.random.default_rng(5).choice(N3, size=4096, replace=False))
    return rows, _dense_rows(par, c, R, rows)


@pytest.mark.parametrize("dtype,bound", [("bf16x2", 1e-5), ("f32", 1e-5), ("bf16", 5e-3)])
def test_c3_rows_vs_dense_kernel(hm, c3, c3_dense, dtype, bound):
    """4096 seeded rows of H*X against the dense FP64 kernel matrix: the
    error of the whole pipeline (ACA at 1e-10 plus the storage rounding)."""
    par, c, H, f, R = c3
    rows, ref = c3_dense
    Y = _product(hm, f, dtype, R)
    err = relerr(Y[rows], ref)
    print(f"C3 {dtype} H*X vs dense FP64 kernel (4096 rows): {err:.3e}", flush=True)
    assert err <= bound


def test_c3_fp64_matvec_vs_oracle(hm, orc, c3):
    import torch
    par, c, H, f, R = c3
    d = hm.DeviceHmat(f, "f64", {"no_batched": True})
    x = torch.from_numpy(R[:, 3].copy()).to(d.device)
    y = d.matvec(x).cpu().numpy()
    err = relerr(y, orc.matvec(f, R[:, 3]))
    d.close()
    print(f"C3 FP64 H*y vs oracle: {err:.3e}", flush=True)
    assert err <= 1e-13


def test_c3_fp64_batched_vs_oracle(hm, orc, c3):
    """FP64 H*X and H'X of the 32 columns on the f64 MFMA path (one pass
    over the FP64 factors per 32 columns) against the oracle on the sampled
    columns (<= 1e-13)."""
    import torch
    par, c, H, f, R = c3
    d = hm.DeviceHmat(f, "f64")
    assert d.info()["batched"] == 1
    X = d.new_cols(N3, 32)
    X.copy_(torch.from_numpy(R).to(X.device))
    Y = d.matmat(X).cpu().numpy()
    Yt = d.matmat(X, transpose=True).cpu().numpy()
    d.close()
    e = relerr(Y[:, COLS], np.stack([orc.matvec(f, R[:, j]) for j in COLS], 1))
    et = relerr(Yt[:, COLS], np.stack([orc.matvec_t(f, R[:, j]) for j in COLS], 1))
    print(f"C3 FP64 batched H*X vs oracle: {e:.3e}; H'X: {et:.3e}", flush=True)
    assert e <= 1e-13 and et <= 1e-13


def test_c3_solve_ir(hm, c3):
    """H\\y on H + 10 I (the shift of the reference's solve test,
    test/lu.jl:25) by refinement with FP64 GPU residuals."""
    par, c, H, f, R = c3
    G = H + 10 * hm.I
    y = R[:, 5].copy()
    x, hist = hm.solve_ir(G, y, tol=1e-12, inner="f32")
    for dv in G._device_cache.values():
        dv.close()
    print(f"C3+10I solve: {len(hist) - 1} sweeps, residual history {hist}", flush=True)
    assert hist[-1] <= 1e-12


@pytest.fixture(scope="module")
def c4(hm):
    members = c4_members(hm, nmat=64, nlo=1000, nhi=40000, seed=6)
    print(f"C4: {len(members)} members, total N = {sum(m.nrows for m in members)}", flush=True)
    return members


@pytest.mark.parametrize("one_launch", [False, True])
def test_c4_every_member_vs_oracle(hm, orc, c4, one_launch):
    import torch
    d = hm.DeviceHmat.batch(c4, "f32", {"no_transpose": True, "no_batched": True, "one_launch": one_launch})
    rng = np.random.default_rng(12)
    xs = [rng.standard_normal(m.ncols) for m in c4]
    x = torch.from_numpy(np.concatenate(xs)).float().to(d.device)
    y = d.matvec(x).cpu().numpy()
    worst = 0.0
    for i, m in enumerate(c4):
        e = relerr(y[d.row_offsets[i]:d.row_offsets[i + 1]], orc.matvec(m, xs[i]))
        worst = max(worst, e)
        assert e <= 2e-6, (i, m.nrows, e)
    info = d.info()
    assert info["fused"] == int(one_launch) and info["wait_faults"] == 0
    d.close()
    print(f"C4: worst member relerr {worst:.3e}", flush=True)


@pytest.mark.parametrize("one_launch", [False, True])
def test_c4_graph_replay_bitwise(hm, c4, one_launch):
    import torch
    d = hm.DeviceHmat.batch(c4, "f32", {"no_transpose": True, "no_batched": True, "one_launch": one_launch})
    x = torch.zeros(d.ncols, dtype=torch.float32, device=d.device)
    y = torch.zeros(d.nrows, dtype=torch.float32, device=d.device)
    g = d.graph(x, y)
    for rep in range(3):
        x.copy_(torch.randn(d.ncols, generator=torch.Generator().manual_seed(20 + rep)).to(d.device))
        ref = d.matvec(x).clone()
        torch.cuda.synchronize()
        y.fill_(float("nan"))
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(y, ref), rep
    g.close()
    assert d.info()["wait_faults"] == 0
    d.close()
