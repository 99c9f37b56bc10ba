"""The C-ABI library loads, exports every symbol include/hmat_amd.h declares,
and rejects malformed block tables before touching the GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "hmat_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(hm_\w+)\s*\(", txt, re.M)))


def test_exports_match_header(hm):
    from hmatrix_amd import _native
    L = _native.lib()
    names = _declared()
    assert len(names) >= 9
    for n in names:
        assert hasattr(L, n), n
        assert n in _native.SIGNATURES, n
    assert L.hm_version().decode().startswith("hmat_amd")


def _create(blocks, data, nrows, ncols, storage=0):
    from hmatrix_amd import _native
    L = _native.lib()
    b = np.ascontiguousarray(np.asarray(blocks, dtype=np.int64).reshape(-1, 8))
    d = np.ascontiguousarray(np.asarray(data, dtype=np.float64))
    h = ctypes.c_void_p()
    rc = L.hm_plan_create(nrows, ncols, b.shape[0], b.ctypes.data_as(ctypes.c_void_p),
                          d.ctypes.data_as(ctypes.c_void_p), d.size, storage, ctypes.byref(h))
    return rc, L.hm_last_error().decode(), h


@pytest.mark.parametrize("blocks,data,shape,msg", [
    ([[2, 0, 0, 4, 4, 0, 0, 0]], np.zeros(16), (4, 4), "unknown kind"),
    ([[0, 0, 0, 5, 4, 0, 0, 0]], np.zeros(20), (4, 4), "outside"),
    ([[0, 0, 0, 4, 4, 0, 1, 0]], np.zeros(16), (4, 4), "arena"),
    ([[1, 0, 0, 4, 4, 2, 0, 5]], np.zeros(12), (4, 4), "arena"),
    ([[0, 0, 0, 4, 2, 0, 0, 0]], np.zeros(16), (4, 4), "tile"),
    ([[0, 0, 0, 2, 4, 0, 0, 0], [0, 0, 0, 2, 4, 0, 0, 0]], np.zeros(8), (4, 4), "covered"),
])
def test_plan_create_rejects(hm, blocks, data, shape, msg):
    rc, err, h = _create(blocks, data, *shape)
    assert rc == -1 and msg in err and not h.value


@pytest.mark.parametrize("seed", [101, 105, 106])
def test_plan_create_rejects_corrupted_random_tables(hm, seed):
    """A seeded random exact tiling passes validation (where there is no GPU
    the call then stops at the device with HM_EHIP; on a GPU it builds a
    plan), and every corruption of it is rejected with HM_EINVAL and its
    message before any device work: a leaf dropped (gap), a leaf repeated
    (overlap), a leaf moved down one row, a dense leaf or low-rank factor
    reaching past the arena, offsets near 2^63 (no overflow in the bounds
    check), a negative rank, a negative offset, and the wrong matrix size."""
    from tests.helpers import random_partition_flat
    rng = np.random.default_rng(seed)
    nr, nc = int(rng.integers(300, 3000)), int(rng.integers(300, 3000))
    side, p_dense, kmax = {101: ((24, 400), 0.4, 40), 105: ((8, 128), 0.7, 8), 106: ((64, 1500), 0.2, 40)}[seed]
    f = random_partition_flat(hm, seed, nr, nc, side, p_dense, kmax)
    b, d = f.blocks, f.data
    rc, err, h = _create(b, d, nr, nc)
    assert rc in (0, -2), err
    if rc == 0:
        from hmatrix_amd import _native
        _native.lib().hm_plan_destroy(h)
    de = int(np.where(b[:, 0] == 0)[0][0])
    lr = int(np.where(b[:, 0] == 1)[0][0])
    big = (1 << 63) - 8

    def edit(i, col, v):
        c = b.copy()
        c[i, col] = v
        return c

    cases = [
        (b[1:], (nr, nc), "tile"),
        (np.vstack([b, b[3:4]]), (nr, nc), "tile"),
        (edit(de, 1, b[de, 1] + 1), (nr, nc), ""),                    # overlap or past the last row
        (edit(de, 6, d.size - b[de, 3] * b[de, 4] + 1), (nr, nc), "dense leaf outside the arena"),
        (edit(lr, 7, d.size), (nr, nc), "low-rank factors outside the arena"),
        (edit(de, 6, big), (nr, nc), "dense leaf outside the arena"),
        (edit(lr, 6, big), (nr, nc), "low-rank factors outside the arena"),
        (edit(lr, 5, -1), (nr, nc), "negative size or rank"),
        (edit(de, 6, -3), (nr, nc), "negative arena offset"),
        (b, (nr + 1, nc), "tile"),
    ]
    for blocks, (m, n), msg in cases:
        rc, err, h = _create(blocks, d, m, n)
        assert rc == -1 and not h.value and msg in err and err, (msg, err)


def test_plan_create_rejects_bad_storage(hm):
    rc, err, h = _create([[0, 0, 0, 1, 1, 0, 0, 0]], [1.0], 1, 1, storage=9)
    assert rc == -1 and "storage" in err


def test_plan_create_ex_rejects_bad_config(hm):
    from hmatrix_amd import _native
    L = _native.lib()
    b = np.array([[0, 0, 0, 1, 1, 0, 0, 0]], dtype=np.int64)
    d = np.ones(1)
    for cfg in [_native.HmConfig(100, 4, 0, 0), _native.HmConfig(1024, 3, 0, 0), _native.HmConfig(1024, 4, 8, 0),
                _native.HmConfig(1024, 4, 0, 3), _native.HmConfig(1024, 4, 0, -1)]:
        h = ctypes.c_void_p()
        rc = L.hm_plan_create_ex(1, 1, 1, b.ctypes.data_as(ctypes.c_void_p), d.ctypes.data_as(ctypes.c_void_p),
                                 1, 0, ctypes.byref(cfg), ctypes.byref(h))
        assert rc == -1 and not h.value


def test_batch_create_checks_each_member(hm):
    from hmatrix_amd import _native
    L = _native.lib()
    good = np.array([[0, 0, 0, 2, 2, 0, 0, 0]], dtype=np.int64)
    bad = np.array([[0, 0, 0, 2, 1, 0, 0, 0]], dtype=np.int64)     # covers half of its 2x2 matrix
    d = np.ones(4)
    for tables, msg in [((good, bad), "member 1"), ((bad, good), "member 0")]:
        pb = (ctypes.c_void_p * 2)(*[t.ctypes.data for t in tables])
        pd = (ctypes.c_void_p * 2)(d.ctypes.data, d.ctypes.data)
        two = lambda v: (ctypes.c_int64 * 2)(v, v)
        h = ctypes.c_void_p()
        rc = L.hm_plan_create_batch(2, two(2), two(2), two(1), pb, pd, two(4), 0, None, ctypes.byref(h))
        assert rc == -1 and msg in L.hm_last_error().decode() and not h.value


def test_batch_create_rejects_oversized_and_overflowing_members(hm):
    """hm_plan_create_batch refuses members whose sizes would take the
    concatenated plan past 2^31 rows (before allocating anything for them)
    and leaves whose offsets or extents overflow int64 arithmetic."""
    from hmatrix_amd import _native
    L = _native.lib()
    good = np.array([[0, 0, 0, 2, 2, 0, 0, 0]], dtype=np.int64)
    d = np.ones(4)
    big = (1 << 63) - 8

    def create(tables, nr, nc):
        pb = (ctypes.c_void_p * 2)(*[t.ctypes.data for t in tables])
        pd = (ctypes.c_void_p * 2)(d.ctypes.data, d.ctypes.data)
        two = lambda a, b: (ctypes.c_int64 * 2)(a, b)
        h = ctypes.c_void_p()
        rc = L.hm_plan_create_batch(2, two(*nr), two(*nc), two(1, 1), pb, pd, two(4, 4), 0, None, ctypes.byref(h))
        return rc, L.hm_last_error().decode(), h

    rc, err, h = create((good, good), (2, 1 << 40), (2, 2))
    assert rc == -1 and "too large" in err and not h.value
    for leaf in ([0, 0, 0, 2, 2, 0, big, 0], [1, 0, 0, 2, 2, 1, 0, big], [0, big, 0, 2, 2, 0, 0, 0]):
        rc, err, h = create((good, np.array([leaf], dtype=np.int64)), (2, 2), (2, 2))
        assert rc == -1 and "member 1" in err and not h.value, (leaf, err)


def test_timing_rejects_null_plan(hm):
    from hmatrix_amd import _native
    L = _native.lib()
    ms = (ctypes.c_double * 3)()
    assert L.hm_plan_timing(None, 1) == -1
    assert L.hm_plan_timing_read(None, ms, 3, None) == -1


def _header_arity():
    """Parameter count of every hm_* function the header declares."""
    txt = open(os.path.join(ROOT, "include", "hmat_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for m in re.finditer(r"(?:int|const char \*)\s*(hm_\w+)\s*\(([^)]*)\)\s*;", txt):
        params = [p for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        out[m.group(1)] = len(params)
    return out


def test_julia_binding_matches_header():
    """Every entry point the Julia binding names is declared, and each ccall
    passes as many arguments as the header's declaration takes (also for
    the calls whose function symbol is chosen at run time)."""
    ar = _header_arity()
    jl = open(os.path.join(ROOT, "hmatrix.jl_amd", "julia", "HmatrixAMD.jl")).read()
    for name in set(re.findall(r":(hm_\w+)", jl)):
        assert name in ar, name
    calls = re.findall(r"ccall\(\((:?\w+), libhmat\),\s*\w+,\s*\(([^()]*)\)", jl)
    assert len(calls) >= 5
    for fn, types in calls:
        n = len([t for t in types.split(",") if t.strip()])
        if fn.startswith(":"):
            assert ar[fn[1:]] == n, (fn, n, ar[fn[1:]])
        else:
            # fn is one of the product entry points with n parameters
            assert n in (ar["hm_matvec"], ar["hm_matmat"]), (fn, n)
    assert ar["hm_matvec"] == ar["hm_matvec_t"] and ar["hm_matmat"] == ar["hm_matmat_t"]


def test_autotune_rejects_before_touching_the_gpu(hm):
    """hm_autotune validates its arguments first: NULL output, reps < 1,
    nrhs outside [1, 32], an unknown storage."""
    from hmatrix_amd import _native
    L = _native.lib()
    b = np.array([[0, 0, 0, 1, 1, 0, 0, 0]], dtype=np.int64)
    d = np.ones(1)
    bp, dp = b.ctypes.data_as(ctypes.c_void_p), d.ctypes.data_as(ctypes.c_void_p)
    cfg = _native.HmConfig()
    assert L.hm_autotune(1, 1, 1, bp, dp, 1, 0, 1, 5, None, None) == -1
    assert L.hm_autotune(1, 1, 1, bp, dp, 1, 0, 1, 0, ctypes.byref(cfg), None) == -1
    assert L.hm_autotune(1, 1, 1, bp, dp, 1, 3, 33, 5, ctypes.byref(cfg), None) == -1
    assert L.hm_autotune(1, 1, 1, bp, dp, 1, 7, 8, 5, ctypes.byref(cfg), None) == -1
    assert "storage" in L.hm_last_error().decode()
