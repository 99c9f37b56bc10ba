"""Parity at four times the benchmark size: N = 2^20 sorted uniform points on
[0, 400] (seed 3), exp(-|x-y|), MaxRank 24, MaxBlock N/4, MinBlock 64, weak
admissibility (72061 leaves, a 1.5 GB FP64 arena), X = 32 standard-normal
columns (seed 4).  The layouts' 32-bit record, fragment and slice offsets
and the projection's super-tile count grow with N; this checks them beyond
the benchmark's size: the batched split-BF16 H*X against the FP64 oracle on
sampled columns (<= 2e-5) and 1024 seeded rows against the dense FP64
kernel (<= 1e-5), and FP64 H*y against the oracle (<= 1e-13).  The
reference checks A*y ~ H*y against the dense kernel (test/harithm.jl:1-12)."""
import numpy as np
import pytest

from tests.helpers import build, expk, relerr, sorted_uniform

pytestmark = pytest.mark.gpu

N = 1 << 20
COLS = [0, 17, 31]


@pytest.fixture(scope="module")
def big(hm):
    par, c, H = build(hm, sorted_uniform(N, 0.0, 400.0, 3), expk, -1, 64, 24)
    f = hm.flatten(H)
    R = np.random.default_rng(4).standard_normal((N, 32))
    print(f"N=2^20: {f.nblocks} leaves", flush=True)
    return c, f, R


def test_scale_batched_split_bf16(hm, orc, big):
    import torch
    c, f, R = big
    d = hm.DeviceHmat(f, "bf16x2")
    assert d.info()["batched"] == 1
    X = d.new_cols(N, 32)
    X.copy_(torch.from_numpy(R).float())
    Y = d.matmat(X).cpu().numpy().astype(np.float64)
    d.close()
    assert np.isfinite(Y).all()
    err = relerr(Y[:, COLS], np.stack([orc.matvec(f, R[:, j]) for j in COLS], 1))
    print(f"N=2^20 split BF16 H*X vs oracle: {err:.3e}", flush=True)
    assert err <= 2e-5
    # 1024 rows of the dense FP64 kernel matrix times X, on the device in
    # FP64 (torch as a calculator, in row blocks), in cluster order
    rows = np.sort(np.random.default_rng(5).choice(N, 1024, replace=False))
    xc = torch.from_numpy(np.asarray(c.X)).cuda()
    Rd = torch.from_numpy(R).cuda()
    ref = torch.empty((rows.size, 32), dtype=torch.float64, device="cuda")
    for s in range(0, rows.size, 64):
        r = torch.from_numpy(rows[s:s + 64]).cuda()
        ref[s:s + 64] = torch.exp(-torch.abs(xc[r][:, None] - xc[None, :])) @ Rd
    err_d = relerr(Y[rows], ref.cpu().numpy())
    print(f"N=2^20 split BF16 H*X vs dense FP64 kernel (1024 rows): {err_d:.3e}", flush=True)
    assert err_d <= 1e-5


def test_scale_fp64_matvec(hm, orc, big):
    import torch
    c, f, R = big
    d = hm.DeviceHmat(f, "f64", {"no_batched": True, "no_transpose": True})
    y = d.matvec(torch.from_numpy(R[:, 0]).cuda()).cpu().numpy()
    d.close()
    err = relerr(y, orc.matvec(f, R[:, 0]))
    print(f"N=2^20 FP64 H*y vs oracle: {err:.3e}", flush=True)
    assert err <= 1e-13
