"""Host-side construction (cluster tree, admissibility, partition, flattening)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from tests.helpers import build, cauchy, expk, sorted_uniform, relerr


def _walk(c):
    yield c
    if not c.isleaf:
        yield from _walk(c.left)
        yield from _walk(c.right)


@pytest.mark.parametrize("X", [np.linspace(0, 10, 1000), sorted_uniform(777, -3, 3, 1),
                               np.random.default_rng(2).uniform(0, 1, (500, 2))])
def test_cluster_tree_invariants(hm, X):
    n, c = hm.construct_cluster(X, 32)
    N = X.shape[0]
    assert sorted(c.P.tolist()) == list(range(N))
    assert np.array_equal(np.asarray(c.X), hm.cluster._as_points(X)[c.P])
    ids = []
    for x in _walk(c):
        ids.append(x.id)
        assert x.e - x.s + 1 == x.N
        if x.isleaf:
            assert x.N <= 32
        else:
            assert x.N > 32
            assert x.left.s == x.s and x.right.e == x.e and x.left.e + 1 == x.right.s
            assert x.m == x.left.N and x.n == x.right.N
    assert ids == list(range(1, n + 1))     # pre-order ids (cluster.jl:223-225)


def test_bisect_1d_is_optimal_2means(hm):
    x = np.random.default_rng(3).uniform(0, 1, 200)
    _, P1, _, P2 = hm.bisect_cluster(x)
    best = min(((np.sort(x)[:i] - np.sort(x)[:i].mean()) ** 2).sum()
               + ((np.sort(x)[i:] - np.sort(x)[i:].mean()) ** 2).sum() for i in range(1, 200))
    got = ((x[P1] - x[P1].mean()) ** 2).sum() + ((x[P2] - x[P2].mean()) ** 2).sum()
    assert got <= best + 1e-12
    assert x[P1].max() < x[P2].min()


def test_admissibility_matches_reference_neighbors(hm):
    """Range query == Neighbors built as in compute_geom_info (cluster.jl:140-169):
    every cluster containing i paired with every cluster containing j."""
    X = sorted_uniform(300, 0, 1, 4)
    rng = np.random.default_rng(5)
    Adj = sp.random(300, 300, density=0.003, random_state=6, format="csr") + sp.identity(300)
    n, c = hm.construct_cluster(X, 16)
    adj = hm.cluster._Adjacency(Adj, c.P)
    Ap = sp.csr_matrix(Adj)[c.P][:, c.P].tocoo()
    p2c = {i: set() for i in range(300)}
    for x in _walk(c):
        for i in range(x.s, x.e + 1):
            p2c[i].add(x.id)
    nb = set()
    for i, j in zip(Ap.row, Ap.col):
        for k1 in p2c[i]:
            for k2 in p2c[j]:
                nb.add((k1, k2))
    nodes = list(_walk(c))
    for s in nodes[::3]:
        for t in nodes[::2]:
            assert adj.neighbours(s, t) == ((s.id, t.id) in nb)


def test_partition_rules_weak(hm):
    """Leaves are full iff a side is <= MinBlock or a leaf cluster; weak
    admissibility makes every other off-diagonal leaf low rank."""
    par, c, H = build(hm, np.linspace(0, 1, 1024), cauchy, 256, 32, 8)
    stack = [H]
    while stack:
        h = stack.pop()
        if h.is_fullmatrix:
            assert h.m <= 32 or h.n <= 32 or h.s.isleaf or h.t.isleaf
        elif h.is_rkmatrix:
            assert h.m <= 256 and h.n <= 256 and not h.s.same_range(h.t)
            assert h.A.shape[1] <= 8
        else:
            stack += [x for row in h.children for x in row]


def test_adj_required(hm):
    par = hm.Params()
    par.Geom = np.linspace(0, 1, 100)
    par.Kernel = cauchy
    par.ConsMethod = "aca2"
    with pytest.raises(ValueError, match="Adj"):
        hm.NewHmat(par)


def test_params_not_mutated(hm):
    par, c, H = build(hm, np.linspace(0, 1, 400), cauchy, -1, 32, 8)
    assert par.MaxBlock == -1 and par.Neighbors is None


def test_aca_accuracy_and_stop(hm):
    x = np.linspace(0, 1, 100)
    y = np.linspace(5, 6, 120)
    U, V = hm.aca2(cauchy, x, y, 20, 1e-10)
    A = hm.FullMat(cauchy, x, y)
    assert U.shape[1] <= 21
    assert np.abs(A - U @ V.T).max() < 1e-9 * np.abs(A).max()
    U1, V1 = hm.aca2(expk, x, y, 20, 1e-10)      # exactly rank 1 (+ a rounding column)
    assert U1.shape[1] <= 2
    assert np.abs(hm.FullMat(expk, x, y) - U1 @ V1.T).max() < 1e-15


def test_compress_zero_block_and_svd(hm):
    par = hm.Params()
    U, V = hm.compress(np.zeros((5, 7)), par)
    assert U.shape == (5, 1) and V.shape == (7, 1) and not U.any()
    M = np.outer(np.arange(1, 6), np.arange(1, 8)) * 1.0
    U, V = hm.compress(M, par)
    assert U.shape[1] == 1 and np.allclose(U @ V.T, M)


def test_bbfmm_and_separate(hm):
    x = np.linspace(0, 1, 50)
    y = np.linspace(3, 4, 60)
    U, V = hm.bbfmm(cauchy, x, y, 10)
    assert np.abs(hm.FullMat(cauchy, x, y) - U @ V.T).max() < 1e-6
    a = [lambda t, i=i: t ** i for i in range(3)]
    U, V = hm.lowrank._rk_matrix(a, a, x, y)
    assert U.shape == (50, 3) and np.allclose(V[:, 1], y - x.mean())


def test_flatten_roundtrip_and_info(hm, tmp_path):
    par, c, H = build(hm, sorted_uniform(900, 0, 9, 7), cauchy, 225, 32, 8)
    f = hm.flatten(H)
    assert np.allclose(hm.to_dense(f), hm.to_fmat(H), atol=1e-15, rtol=0)
    nd, nr, depth, ratio = hm.info(H)
    assert (nd, nr) == f.counts()
    assert ratio == pytest.approx(f.stored_elements() / 900 ** 2)
    p = os.path.join(tmp_path, "f.npz")
    hm.save_factors(p, f)
    g = hm.load_factors(p)
    assert np.array_equal(g.blocks, f.blocks) and np.array_equal(g.data, f.data)
    assert np.array_equal(g.perm, c.P)


def test_flat_validate_rejects(hm):
    f = hm.FlatHmat(4, 4, np.array([[0, 0, 0, 4, 4, 0, 0, 0]], dtype=np.int64), np.zeros(16))
    f.validate()
    with pytest.raises(ValueError):
        hm.FlatHmat(4, 4, np.array([[0, 0, 0, 4, 3, 0, 0, 0]], dtype=np.int64), np.zeros(16)).validate()
    with pytest.raises(ValueError):
        hm.FlatHmat(4, 4, np.array([[0, 0, 0, 4, 4, 0, 1, 0]], dtype=np.int64), np.zeros(16)).validate()


def test_product_fails_loudly_without_gpu(hm):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    par, c, H = build(hm, np.linspace(0, 1, 200), cauchy, 50, 16, 4)
    with pytest.raises(RuntimeError):
        H * np.ones(200)


def test_add_identity(hm):
    par, c, H = build(hm, np.linspace(0, 5, 500), cauchy, 125, 32, 6)
    G = H + 10 * hm.I
    assert np.allclose(hm.to_fmat(G), hm.to_fmat(H) + 10 * np.eye(500), rtol=0, atol=1e-14)
    assert np.allclose(hm.to_fmat(H), hm.to_fmat(hm.copy_hmat(H)))      # H untouched


@pytest.mark.parametrize("restart,dtype", [(60, "float64"), (7, "float64"), (4, "float32")])
def test_device_gmres_on_cpu_tensors(hm, restart, dtype):
    """The refinement's GMRES (basis on the tensor's device) run on CPU
    tensors with a dense operator: it reaches the requested residual with and
    without restarts, and its answer matches a direct solve."""
    import torch
    from hmatrix_amd.solve import _gmres_device
    rng = np.random.default_rng(restart)
    n = 300
    A = 4.0 * np.eye(n) + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    td = getattr(torch, dtype)
    At = torch.from_numpy(A).to(td)
    calls = []

    def mv(v):
        calls.append(1)
        return At @ v

    tol = 1e-10 if dtype == "float64" else 1e-5
    x, ok = _gmres_device(mv, torch.from_numpy(b).to(td), tol, restart, 200)
    assert ok
    x = x.double().numpy()
    assert np.linalg.norm(b - A @ x) <= 2 * tol * np.linalg.norm(b)
    assert relerr(x, np.linalg.solve(A, b)) <= 10 * tol
    if restart < 20:
        assert len(calls) > restart               # it had to restart


def test_autotune_cache_roundtrip(hm, tmp_path):
    """The autotuners' JSON cache: entries of several workloads coexist, a
    missing or unreadable file reads as a miss, and the key separates
    storage, shape and leaf statistics."""
    from hmatrix_amd.autotune import _load_cache, _store_cache, workload_key
    path = str(tmp_path / "sub" / "cache.json")
    assert _load_cache(path, "k") is None
    _store_cache(path, "a", {"config": {"proj_chunk": 512}, "timings": {"x": 1.0}})
    _store_cache(path, "b", {"config": {"strip_shape": 2}, "timings": {}})
    assert _load_cache(path, "a")["config"] == {"proj_chunk": 512}
    assert _load_cache(path, "b")["config"] == {"strip_shape": 2}
    with open(path, "w") as fh:
        fh.write("{not json")
    assert _load_cache(path, "a") is None
    par, c, H = build(hm, sorted_uniform(700, 0, 5, 2), cauchy, 200, 32, 6)
    f = hm.flatten(H)
    assert workload_key(f, "f64") != workload_key(f, "f32")
    assert workload_key(f, "f64").startswith("f64:700x700:")
