"""hm_matvec_static (the entry point that allocates nothing: t and partial
sums in the library's static device array) against hm_matvec on the same
plans, through the C ABI: two calls in a row, the second on a smaller
matrix, bitwise equal to hm_matvec and within 1e-13 of the FP64 oracle (FP32
plan: 1e-6); a plan needing more scratch than the static maximum is rejected
and still runs through hm_matvec."""
import ctypes

import numpy as np
import pytest

from tests.helpers import build, c1, cauchy, relerr, sorted_uniform

pytestmark = pytest.mark.gpu


def _call(L, fn, h, x, y, stream):
    rc = getattr(L, fn)(h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), 1.0, 0.0,
                        ctypes.c_void_p(stream.cuda_stream))
    assert rc == 0, (fn, L.hm_last_error().decode())


def test_static_scratch_matches_plan_scratch(hm, orc):
    import torch
    from hmatrix_amd import _native
    L = _native.lib()
    stream = torch.cuda.current_stream()
    par, c, H = build(hm, sorted_uniform(65536, 0.0, 100.0, 1), cauchy, 1024, 64, 16)
    big = hm.flatten(H)
    par, c, H = c1(hm, "aca2")
    small = hm.flatten(H)
    # C2 (N = 65536) FP64, then the smaller C1 (N = 2000) in FP64 and FP32:
    # each call leaves the shared array holding its own t
    for f, dtype, tol in ((big, "f64", 1e-13), (small, "f64", 1e-13), (small, "f32", 1e-6)):
        print(f"hm_matvec_static: N={f.nrows} {dtype}", flush=True)
        d = hm.DeviceHmat(f, dtype, {"no_transpose": True, "no_batched": True})
        x64 = np.random.default_rng(f.nrows).standard_normal(f.ncols)
        x = torch.from_numpy(x64).to(d.device, d.vec_dtype)
        y_static = torch.full((f.nrows,), float("nan"), dtype=d.vec_dtype, device=d.device)
        y_plan = torch.full((f.nrows,), float("nan"), dtype=d.vec_dtype, device=d.device)
        _call(L, "hm_matvec_static", d._h, x, y_static, stream)
        _call(L, "hm_matvec", d._h, x, y_plan, stream)
        torch.cuda.synchronize()
        assert torch.equal(y_static, y_plan)
        assert relerr(y_static.cpu().numpy().astype(np.float64), orc.matvec(f, x64)) <= tol
        d.close()


def test_static_scratch_rejects_oversized_plan(hm):
    """Four 1x1 low-rank leaves of rank 150000: t alone takes 4.8 MB in FP64,
    more than HM_STATIC_SCRATCH_BYTES (4 MiB)."""
    import torch
    from hmatrix_amd import _native
    L = _native.lib()
    k = 150000
    rng = np.random.default_rng(3)
    blocks, chunks, off = [], [], 0
    for (r0, c0) in ((0, 0), (0, 1), (1, 0), (1, 1)):
        A, B = rng.standard_normal(k) / k, rng.standard_normal(k)
        blocks.append((1, r0, c0, 1, 1, k, off, off + k))
        chunks += [A, B]
        off += 2 * k
    f = hm.FlatHmat(2, 2, np.array(blocks, dtype=np.int64), np.concatenate(chunks))
    d = hm.DeviceHmat(f, "f64", {"no_transpose": True, "no_batched": True})
    x = torch.tensor([0.5, -2.0], dtype=torch.float64, device=d.device)
    y = torch.zeros(2, dtype=torch.float64, device=d.device)
    stream = torch.cuda.current_stream()
    rc = L.hm_matvec_static(d._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), 1.0, 0.0,
                            ctypes.c_void_p(stream.cuda_stream))
    assert rc == -4 and "HM_STATIC_SCRATCH_BYTES" in L.hm_last_error().decode()
    _call(L, "hm_matvec", d._h, x, y, stream)
    ref = np.zeros(2)
    for kd, r0, c0, m, n, kk, oa, ob in f.blocks:
        ref[r0] += (f.data[oa:oa + kk] @ f.data[ob:ob + kk]) * (0.5 if c0 == 0 else -2.0)
    assert relerr(y.cpu().numpy(), ref) <= 1e-13
    d.close()
