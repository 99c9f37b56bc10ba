"""H-matrix product benchmark on one MI355X.

Workload (default): N = 262144 sorted uniform points on [0, 400] (seed 3),
kernel exp(-|x-y|), MaxRank 24, MaxBlock N/4, MinBlock 64, weak
admissibility (Adj = identity), ACA construction on the host in FP64;
X = 32 standard-normal right-hand sides (seed 4).

One step = one product Y = H*X of the 32 right-hand sides, inputs resident
in HBM.  Steps are independent products; by default two are in flight at a
time on two default-priority HIP streams (--streams), which is the
throughput reported; the one-at-a-time latency is reported beside it.  The batched path streams the factors once per step in the chosen
storage (default: split BF16, FP32 accumulation).  The plain FP64 path
(one matvec per right-hand side, the reference's H*X semantics) is timed
beside it.  Several plan replicas are rotated so the working set exceeds the
256 MB Infinity Cache.

Prints one JSON line (see the repository's measurement notes for fields).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(hm, N: int, nrhs: int):
    import scipy.sparse as sp
    par = hm.Params()
    X = np.sort(np.random.default_rng(3).uniform(0.0, 400.0, N))
    par.Geom = X
    par.Kernel = lambda x, y: np.exp(-np.abs(x - y))
    par.MaxRank, par.MaxBlock, par.MinBlock = 24, -1, 64
    par.ConsMethod = "aca2"
    par.Adj = sp.identity(N, format="csr")
    t0 = time.time()
    c, H = hm.NewHmat(par)
    f = hm.flatten(H)
    log(f"construction N={N}: {time.time() - t0:.1f}s, {f.nblocks} leaves {f.counts()}")
    RHS = np.random.default_rng(4).standard_normal((N, nrhs))
    return par, c, f, RHS


def algorithmic_bytes(f, storage: str, nrhs: int, vec_bytes: int) -> int:
    """b_f * (sum_dense m n + sum_rk (m+n) k) + (b_x + b_y) N nrhs."""
    b = f.blocks
    dense = b[:, 0] == 0
    de = int((b[dense, 3] * b[dense, 4]).sum())
    re = int(((b[~dense, 3] + b[~dense, 4]) * b[~dense, 5]).sum())
    fb = {"f64": 8, "f32": 4, "bf16": 2, "bf16x2": 4}[storage]
    return fb * (de + re) + 2 * vec_bytes * f.nrows * nrhs


def kernel_bytes(f, storage: str, nrhs: int, vec_bytes: int) -> dict:
    """Algorithmic bytes of one launch of each kernel of a product:
    projection  = b_f * sum_rk n k + b_x N nrhs + b_t T nrhs   (reads B, X; writes t)
    row pass    = b_f * (sum_dense m n + sum_rk m k) + (b_x + b_y) N nrhs + b_t T nrhs
                  (reads dense leaves, A, X, t; writes Y)
    reduction   = 0 (folds the projection's partial sums; no algorithmic bytes)
    with T = sum_rk k the length of t."""
    b = f.blocks
    dense = b[:, 0] == 0
    de = int((b[dense, 3] * b[dense, 4]).sum())
    ae = int((b[~dense, 3] * b[~dense, 5]).sum())
    be = int((b[~dense, 4] * b[~dense, 5]).sum())
    T = int(b[~dense, 5].sum())
    fb = {"f64": 8, "f32": 4, "bf16": 2, "bf16x2": 4}[storage]
    nv = vec_bytes * nrhs
    return {"proj": fb * be + nv * (f.ncols + T),
            "reduce": 0,
            "rows": fb * (de + ae) + nv * (f.ncols + T + f.nrows)}


def dense_rows_reference(par, c, RHS, rows):
    """FP64 dense kernel rows: (K(x_rows, x) @ RHS) in cluster order."""
    Xc = np.asarray(c.X)
    out = np.zeros((rows.size, RHS.shape[1]))
    for s in range(0, rows.size, 128):
        r = rows[s:s + 128]
        K = par.Kernel(Xc[r][:, None], Xc[None, :])
        out[s:s + 128] = K @ RHS
    return out


def timed(torch, fn, steps: int, warmup: int, stream, after_warmup=None):
    for i in range(warmup):
        fn(i)
    torch.cuda.synchronize()
    if after_warmup is not None:
        after_warmup()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for i in range(steps):
        fn(i)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps   # ms per step


def timed_streams(torch, fn, steps: int, warmup: int, streams):
    """Like timed(), with step i issued on streams[i % len(streams)]: the
    steps are independent products, so with several streams consecutive
    products overlap on the GPU.  Bracketed by synchronizations; the clock
    runs on streams[0], which joins the others before the end event."""
    for i in range(warmup):
        fn(i, streams[i % len(streams)])
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(streams[0])
    for s in streams[1:]:
        s.wait_event(e0)
    for i in range(steps):
        fn(i, streams[i % len(streams)])
    for s in streams[1:]:
        ev = torch.cuda.Event()
        ev.record(s)
        streams[0].wait_event(ev)
    e1.record(streams[0])
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps   # ms per step


def cpu_baseline(f, seconds: float, threads: int = 1):
    """Oracle FP64 matvec on the same factors, bounded sample: 1 thread (the
    serial restatement) or `threads` OpenMP threads (row ownership, the same
    sums bitwise)."""
    from oracle import oracle
    x = np.random.default_rng(9).standard_normal(f.ncols)
    mv = (lambda: oracle.matvec(f, x)) if threads == 1 else (lambda: oracle.matvec_mt(f, x, threads))
    mv()        # warm
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        mv()
        n += 1
    dt = time.perf_counter() - t0
    return n / dt, n, dt


def host_threads() -> int:
    """Threads for the all-core CPU leg: OMP_NUM_THREADS when set (the GPU
    host's share), else the visible cores, at most 16."""
    try:
        return max(1, int(os.environ["OMP_NUM_THREADS"]))
    except (KeyError, ValueError):
        return max(1, min(16, os.cpu_count() or 1))


def profiled_traffic(storage: str, kernel: str):
    """HBM bytes per launch of one kernel from the committed counter summary
    (profiles/traffic_<storage>.json: FETCH_SIZE and WRITE_SIZE per dispatch
    on this workload, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{storage}.json")
    try:
        with open(path) as fh:
            return json.load(fh)["bytes_per_launch"][kernel]
    except (OSError, ValueError, KeyError):
        return None


def measure_batched(hm, torch, f, RHS, storage, args, streams, rows, Yd):
    """Time the batched product in one storage: S products in flight on S
    streams (the throughput), one at a time (latency), and one at a time
    with the library's per-kernel events; relative error of 4096 rows
    against the dense FP64 kernel.  Returns a dict (plans kept under
    "_plans" for the caller)."""
    N, nrhs = RHS.shape
    dev = streams[0].device
    stream = streams[0]
    R = args.replicas
    t0 = time.time()
    plans = [hm.DeviceHmat(f, storage) for _ in range(R)]
    info = plans[0].info()
    log(f"plans ({storage} x{R}) created in {time.time() - t0:.1f}s; info {info}")
    if info["batched"] != 1:
        raise RuntimeError("the batched layout was not built for this matrix")
    pad = getattr(args, "ld_pad", 0)

    def cols():
        # column-major N x nrhs with leading dimension N + pad
        return torch.zeros((nrhs, N + pad), dtype=torch.float32, device=dev)[:, :N].t()

    Xs = [cols() for _ in range(R)]
    Ys = [cols() for _ in range(R)]
    for Xd in Xs:
        Xd.copy_(torch.from_numpy(RHS).to(dev, torch.float32))

    def step_batched(i):
        plans[i % R].matmat(Xs[i % R], Ys[i % R], stream=stream)

    def start_kernel_timing():
        for p in plans:
            p.timing(True)
            p.timing_read()

    ms = timed_streams(torch, lambda i, st: plans[i % R].matmat(Xs[i % R], Ys[i % R], stream=st),
                       args.steps, args.warmup, streams)
    ms1 = timed(torch, step_batched, args.steps, args.warmup, stream)
    # kernel durations: single-stream steps again, with HIP events recorded by
    # the library on `stream` around each kernel (the events add a few
    # microseconds of gaps per step, so the step times above are taken without)
    ms_ev = timed(torch, step_batched, args.steps, 0, stream, start_kernel_timing)
    ph = {"proj_ms": 0.0, "reduce_ms": 0.0, "rows_ms": 0.0, "launches": 0}
    for p in plans:
        r = p.timing_read()
        p.timing(False)
        for k in ph:
            ph[k] += r[k]
    nl = max(ph["launches"], 1)
    kus = {k: ph[f"{k}_ms"] * 1e3 / nl for k in ("proj", "reduce", "rows")}
    log(f"batched {storage}: {ms * 1e3:.1f} us per product of {nrhs} rhs with {len(streams)} products in "
        f"flight, {ms1 * 1e3:.1f} us one at a time ({ms_ev * 1e3:.1f} with kernel events); kernels (us/launch, "
        f"{ph['launches']} launches): proj {kus['proj']:.1f} reduce {kus['reduce']:.1f} rows {kus['rows']:.1f}")
    Yb = Ys[0].cpu().numpy().astype(np.float64)
    relerr_dense = float(np.linalg.norm(Yb[rows] - Yd) / np.linalg.norm(Yd))
    log(f"batched {storage}: relerr vs dense FP64 ({rows.size} rows) {relerr_dense:.3e}")
    return {"ms": ms, "ms1": ms1, "ms_ev": ms_ev, "kus": kus, "relerr_dense": relerr_dense,
            "_plans": (plans, Xs, Ys)}


def scale_report(hm, torch, args, streams) -> dict:
    """The headline workload at four times its size (N = 2^20 points on
    [0, 400], the same kernel and parameters; --extras): split-BF16 H*X, 32
    right-hand sides, with the headline's schedule (plan replicas, products
    in flight), its kb_rows rate and the relative error of 256 rows against
    the dense FP64 kernel."""
    import copy
    N = 1 << 20
    par, c, f, RHS = build_workload(hm, N, 32)
    rows = np.sort(np.random.default_rng(5).choice(N, size=256, replace=False))
    Yd = dense_rows_reference(par, c, RHS, rows)
    a = copy.copy(args)
    a.steps, a.warmup, a.ld_pad = 50, 5, 0
    r = measure_batched(hm, torch, f, RHS, "bf16x2", a, streams, rows, Yd)
    for p in r["_plans"][0]:
        p.close()
    nbytes = algorithmic_bytes(f, "bf16x2", 32, 4)
    kb = kernel_bytes(f, "bf16x2", 32, 4)
    out = {"scale_n": N, "scale_leaves": int(f.nblocks), "scale_us_per_product": r["ms"] * 1e3,
           "scale_single_stream_us": r["ms1"] * 1e3, "scale_H-matvec/s": 32 / (r["ms"] * 1e-3),
           "scale_step_frac": nbytes / (r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "scale_kb_rows_us": r["kus"]["rows"],
           "scale_kb_rows_frac": kb["rows"] / (r["kus"]["rows"] * 1e-6) / 1e9 / HBM_PEAK_GBS,
           "scale_relerr_vs_dense_fp64": r["relerr_dense"]}
    log(f"N=2^20: {out['scale_us_per_product']:.1f} us per product ({out['scale_H-matvec/s']:.0f} H-matvec/s, step "
        f"{out['scale_step_frac']:.3f} of 8 TB/s), kb_rows {out['scale_kb_rows_frac']:.3f}")
    return out


def c4_report(hm, torch, stream) -> dict:
    """64 H-matrices, N log-uniform in [1000, 40000], 1/(1+((x-y)/l)^2),
    l ~ U[0.2, 5], MaxRank in {4, 8, 16}; FP32; all members per launch.
    The plan shape (proj_chunk x row_waves, three launches per product or
    one) comes from the batch autotuner; both forms are then timed eager,
    as a captured HIP graph of one product (hm_graph_create) and as one
    caller graph (torch.cuda.graph) holding four products, with the tuned
    proj_chunk and row_waves."""
    from hmatrix_amd.autotune import autotune_batch
    from hmatrix_amd.workloads import uneven_batch
    t0 = time.time()
    members = uneven_batch(hm, nmat=64, nlo=1000, nhi=40000, seed=6)
    log(f"C4 construction: {time.time() - t0:.1f}s, total N = {sum(m.nrows for m in members)}")
    t0 = time.time()
    best, tune = autotune_batch(members, "f32", chunks=(256, 512, 1024), waves=(2, 4), reps=20, cache=None)
    log(f"C4 autotune ({len(tune)} shapes, {time.time() - t0:.1f}s): best {best}")
    res = {}
    for one in (False, True):
        cfg = {"no_transpose": True, "no_batched": True, "proj_chunk": best["proj_chunk"],
               "row_waves": best["row_waves"], "one_launch": one}
        d = hm.DeviceHmat.batch(members, "f32", cfg)
        x = torch.randn(d.ncols, dtype=torch.float32, device=d.device)
        y = torch.zeros(d.nrows, dtype=torch.float32, device=d.device)
        g = d.graph(x, y)
        ms_g = timed(torch, lambda i: g.launch(stream), 50, 5, stream)
        ms_e = timed(torch, lambda i: d.matvec(x, y, stream=stream), 50, 5, stream)
        g.close()
        # four products recorded in one caller graph: the runtime's fixed
        # cost per graph launch is shared by the four
        cs = torch.cuda.Stream(device=x.device)
        g4 = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g4, stream=cs):
            for _ in range(4):
                d.matvec(x, y, stream=cs)
        with torch.cuda.stream(stream):              # replay() launches on the current stream
            ms_g4 = timed(torch, lambda i: g4.replay(), 20, 3, stream) / 4
        del g4
        if one:
            assert d.info()["fused"] == 1 and d.info()["wait_faults"] == 0
        nrows = d.nrows
        d.close()
        res[one] = (ms_g, ms_e, ms_g4)
    elems = sum(m.stored_elements() for m in members)
    nbytes = 4 * elems + 8 * nrows
    ms_best = min(res[False][1], res[True][1])
    log(f"C4 batch (64 matrices, FP32, 1 rhs each, proj_chunk {best['proj_chunk']}, row_waves "
        f"{best['row_waves']}), three launches: graph {res[False][0] * 1e3:.1f} us, eager {res[False][1] * 1e3:.1f} "
        f"us, graph of 4 {res[False][2] * 1e3:.1f} us per product; one launch: graph {res[True][0] * 1e3:.1f} us, "
        f"eager {res[True][1] * 1e3:.1f} us, graph of 4 {res[True][2] * 1e3:.1f} us; "
        f"{64 / (ms_best * 1e-3):.0f} H-matvec/s; {nbytes / (ms_best * 1e-3) / 1e9:.0f} GB/s")
    return {"workload": "64 H-matrices, N log-uniform in [1000, 40000] (seed 6), 1/(1+((x-y)/l)^2), "
                        "MaxRank in {4,8,16}, FP32, one right-hand side each, all members per launch",
            "total_N": int(nrows), "tuned_config": best, "autotune_ms": tune,
            "graph_us": res[False][0] * 1e3, "eager_us": res[False][1] * 1e3,
            "graph_of_4_us_per_product": res[False][2] * 1e3,
            "one_launch_graph_us": res[True][0] * 1e3, "one_launch_eager_us": res[True][1] * 1e3,
            "one_launch_graph_of_4_us_per_product": res[True][2] * 1e3,
            "H-matvec/s": 64 / (ms_best * 1e-3), "algorithmic_bytes": nbytes,
            "GB/s": nbytes / (ms_best * 1e-3) / 1e9, "frac": nbytes / (ms_best * 1e-3) / 1e9 / HBM_PEAK_GBS}


def extras_report(hm, torch, stream, plan_b, Xb, Yb) -> dict:
    """Timings of the other product forms (stderr and the JSON line's
    "extras"): the transposed batched product H'X on the same plan replicas
    (rotated); the plain FP64 H*y, r -= H*y and H'*y at N=65536 (sorted
    uniform on [0,100], seed 1, Cauchy kernel, MaxBlock 1024, MinBlock 64,
    MaxRank 16); H\\y on that matrix + 10 I and on the N=262144 matrix
    + 10 I by iterative refinement; the batched product on the N=262144
    points with strong admissibility; the FP64 adjoint of
    L = 0.5||Hx - b||^2 at N=65536 (seed 7, same kernel and parameters)."""
    import scipy.sparse as sp
    out = {}
    R = len(plan_b)
    ms = timed(torch, lambda i: plan_b[i % R].matmat(Xb[i % R], Yb[i % R], stream=stream, transpose=True),
               20, 3, stream)
    out["batched_transpose_us"] = ms * 1e3
    log(f"H'X (batched, same plans, one at a time): {ms * 1e3:.1f} us per product of {Xb[0].shape[1]} rhs")

    def cauchy_plan(seed):
        par = hm.Params()
        par.Geom = np.sort(np.random.default_rng(seed).uniform(0.0, 100.0, 65536))
        par.Kernel = lambda x, y: 1.0 / (1.0 + (x - y) ** 2)
        par.MaxBlock, par.MinBlock, par.MaxRank = 1024, 64, 16
        par.ConsMethod = "aca2"
        par.Adj = sp.identity(65536, format="csr")
        c, H = hm.NewHmat(par)
        return hm.flatten(H)

    f2 = cauchy_plan(1)
    d2 = hm.DeviceHmat(f2, "f64", {"no_batched": True})
    x = torch.from_numpy(np.random.default_rng(2).standard_normal(65536)).to(d2.device)
    y = torch.zeros(65536, dtype=torch.float64, device=d2.device)
    ms = timed(torch, lambda i: d2.matvec(x, y, stream=stream), 50, 5, stream)
    b = f2.blocks
    dense = b[:, 0] == 0
    nbytes = 8 * int((b[dense, 3] * b[dense, 4]).sum() + ((b[~dense, 3] + b[~dense, 4]) * b[~dense, 5]).sum()) \
        + 16 * 65536
    out["c2_fp64_matvec_us"] = ms * 1e3
    out["c2_fp64_GBs"] = nbytes / (ms * 1e-3) / 1e9
    log(f"C2 FP64 H*y (N=65536, {f2.nblocks} leaves): {ms * 1e3:.1f} us, {out['c2_fp64_GBs']:.0f} GB/s")
    # hmat_matvec!(r, H, v, s): r += s*H*v (reads r as well), and v'*H (the transposed layout)
    ms_acc = timed(torch, lambda i: d2.matvec(x, y, alpha=-1.0, beta=1.0, stream=stream), 50, 5, stream)
    ms_t = timed(torch, lambda i: d2.matvec(x, y, transpose=True, stream=stream), 50, 5, stream)
    out["c2_fp64_accumulate_us"] = ms_acc * 1e3
    out["c2_fp64_matvec_t_us"] = ms_t * 1e3
    log(f"C2 FP64 r -= H*y: {ms_acc * 1e3:.1f} us; H'*y: {ms_t * 1e3:.1f} us")
    d2.close()

    # H\y by iterative refinement on the C2 matrix shifted by 10 I (the
    # reference's solve test shifts its matrix the same way, test/lu.jl:25):
    # FP64 GPU residuals, GMRES corrections on an FP32 copy
    par = hm.Params()
    par.Geom = np.sort(np.random.default_rng(1).uniform(0.0, 100.0, 65536))
    par.Kernel = lambda x, y: 1.0 / (1.0 + (x - y) ** 2)
    par.MaxBlock, par.MinBlock, par.MaxRank = 1024, 64, 16
    par.ConsMethod = "aca2"
    par.Adj = sp.identity(65536, format="csr")
    c, H = hm.NewHmat(par)
    G = H + 10 * hm.I
    yv = np.random.default_rng(2).standard_normal(65536)
    hm.solve_ir(G, yv, tol=1e-12, inner="f32")           # plans and warm-up
    t0 = time.perf_counter()
    xs, hist = hm.solve_ir(G, yv, tol=1e-12, inner="f32")
    dt = time.perf_counter() - t0
    out["c2_solve_ir_ms"] = dt * 1e3
    out["c2_solve_ir_sweeps"] = len(hist) - 1
    out["c2_solve_ir_relres"] = hist[-1]
    log(f"C2+10I solve (FP64 GPU residual, FP32 GMRES corrections; N=65536): {dt * 1e3:.1f} ms, "
        f"{len(hist) - 1} refinement sweeps, relative residual {hist[-1]:.2e}")
    for dv in G._device_cache.values():
        dv.close()
    # the same at N=262144: exp(-|x-y|) workload + 10 I
    par3 = hm.Params()
    par3.Geom = np.sort(np.random.default_rng(3).uniform(0.0, 400.0, 262144))
    par3.Kernel = lambda x, y: np.exp(-np.abs(x - y))
    par3.MaxRank, par3.MaxBlock, par3.MinBlock = 24, -1, 64
    par3.ConsMethod = "aca2"
    par3.Adj = sp.identity(262144, format="csr")
    c3, H3 = hm.NewHmat(par3)
    G3 = H3 + 10 * hm.I
    y3 = np.random.default_rng(4).standard_normal(262144)
    hm.solve_ir(G3, y3, tol=1e-12, inner="f32")          # plans and warm-up
    t0 = time.perf_counter()
    xs3, hist3 = hm.solve_ir(G3, y3, tol=1e-12, inner="f32")
    dt3 = time.perf_counter() - t0
    out["c3_solve_ir_ms"] = dt3 * 1e3
    out["c3_solve_ir_sweeps"] = len(hist3) - 1
    out["c3_solve_ir_relres"] = hist3[-1]
    log(f"C3+10I solve (N=262144, FP64 GPU residual, FP32 GMRES corrections): {dt3 * 1e3:.1f} ms, "
        f"{len(hist3) - 1} refinement sweeps, relative residual {hist3[-1]:.2e}")
    for dv in G3._device_cache.values():
        dv.close()
    del G3, H3, c3

    # the same N=262144 exp(-|x-y|) points with strong (tridiagonal)
    # admissibility: three dense leaves per block row, the batched path with
    # a larger dense share; one plan per stream, two products in flight
    par3 = hm.Params()
    par3.Geom = np.sort(np.random.default_rng(3).uniform(0.0, 400.0, 262144))
    par3.Kernel = lambda x, y: np.exp(-np.abs(x - y))
    par3.MaxRank, par3.MaxBlock, par3.MinBlock = 24, -1, 64
    par3.ConsMethod = "aca2"
    par3.Adj = sp.diags([1, 1, 1], [-1, 0, 1], shape=(262144, 262144), format="csr")
    c3, H3 = hm.NewHmat(par3)
    f3 = hm.flatten(H3)
    del H3, c3
    S3 = [stream, torch.cuda.Stream(device=stream.device)]
    p3 = [hm.DeviceHmat(f3, "bf16x2") for _ in S3]
    X3 = [p3[0].new_cols(262144, 32) for _ in S3]
    Y3 = [p3[0].new_cols(262144, 32) for _ in S3]
    R3 = np.random.default_rng(4).standard_normal((262144, 32))
    for Xd in X3:
        Xd.copy_(torch.from_numpy(R3).to(Xd.device, torch.float32))
    ms1 = timed(torch, lambda i: p3[i % 2].matmat(X3[i % 2], Y3[i % 2], stream=stream), 20, 3, stream)
    ms2 = timed_streams(torch, lambda i, st: p3[i % 2].matmat(X3[i % 2], Y3[i % 2], stream=st), 20, 3, S3)
    from oracle import oracle
    cols = [0, 17]
    ref = np.stack([oracle.matvec(f3, R3[:, j]) for j in cols], 1)
    got = Y3[0].cpu().numpy()[:, cols].astype(np.float64)
    err3 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    nb3 = algorithmic_bytes(f3, "bf16x2", 32, 4)
    out["c3_strong_us"] = ms2 * 1e3
    out["c3_strong_single_us"] = ms1 * 1e3
    out["c3_strong_GBs"] = nb3 / (ms2 * 1e-3) / 1e9
    out["c3_strong_relerr_vs_fp64_hmatrix"] = err3
    log(f"C3 strong admissibility ({f3.nblocks} leaves, {nb3 / 1e6:.0f} MB per product): {ms2 * 1e3:.1f} us per "
        f"32-rhs product with two in flight ({32 / (ms2 * 1e-3):.0f} H-matvec/s, {out['c3_strong_GBs']:.0f} GB/s), "
        f"{ms1 * 1e3:.1f} us one at a time; relerr vs FP64 H-matrix (2 columns) {err3:.2e}")
    for q in p3:
        q.close()
    del X3, Y3

    f5 = cauchy_plan(7)
    d5 = hm.DeviceHmat(f5, "f64", {"no_batched": True})
    rng = np.random.default_rng(8)
    x5 = torch.from_numpy(rng.standard_normal(65536)).to(d5.device)
    b5 = torch.from_numpy(rng.standard_normal(65536)).to(d5.device)
    bufs = d5.adjoint(x5, b5, stream=stream)[1:]       # res, dx, g reused below
    ms = timed(torch, lambda i: d5.adjoint(x5, b5, stream=stream, out=bufs), 20, 3, stream)
    ms_nl = timed(torch, lambda i: d5.adjoint(x5, b5, loss=False, stream=stream, out=bufs), 20, 3, stream)
    out["c5_adjoint_us"] = ms * 1e3
    out["c5_adjoint_no_loss_us"] = ms_nl * 1e3
    log(f"C5 FP64 adjoint (residual, H' residual, factor gradients; N=65536): {ms_nl * 1e3:.1f} us; "
        f"with the loss read back to the host: {ms * 1e3:.1f} us")
    d5.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--nrhs", type=int, default=32)
    ap.add_argument("--storage", default="bf16x2", choices=["f32", "bf16", "bf16x2"])
    ap.add_argument("--replicas", type=int, default=4)
    ap.add_argument("--streams", type=int, default=2,
                    help="independent products in flight: step i runs on stream i %% S (default "
                         "priority streams; replica r always on stream r %% S)")
    ap.add_argument("--plain-steps", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--sample-rows", type=int, default=4096)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="after the measurement, run 3 device copies of this many MB (the plain copy "
                         "kernel the FETCH_SIZE/WRITE_SIZE counters are scaled against)")
    ap.add_argument("--no-c4", action="store_true",
                    help="skip the uneven 64-matrix batch (one launch; reported in the line's \"c4\" field)")
    ap.add_argument("--no-bf16", action="store_true",
                    help="skip the plain BF16 storage line (the line's \"bf16\" field)")
    ap.add_argument("--ld-pad", type=int, default=0,
                    help="leading dimension of X and Y = N + this many rows (default 0)")
    ap.add_argument("--extras", action="store_true",
                    help="also time H'X, the FP64 H*y at N=65536 and the FP64 adjoint at N=65536")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the product is "
              "not sharded across GPUs", file=sys.stderr)
        return 2

    import torch
    import __graft_entry__
    hm = __graft_entry__.package()
    N, nrhs = args.n, args.nrhs
    par, c, f, RHS = build_workload(hm, N, nrhs)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    S = max(1, args.streams)
    if args.replicas % S:
        print(f"--replicas ({args.replicas}) must be a multiple of --streams ({S})", file=sys.stderr)
        return 2
    streams = [stream] + [torch.cuda.Stream(device=dev) for _ in range(S - 1)]

    R = args.replicas
    rows = np.sort(np.random.default_rng(5).choice(N, size=min(args.sample_rows, N), replace=False))
    t0 = time.time()
    Yd = dense_rows_reference(par, c, RHS, rows)
    log(f"dense FP64 sample ({rows.size} rows) in {time.time() - t0:.1f}s")
    main_run = measure_batched(hm, torch, f, RHS, args.storage, args, streams, rows, Yd)
    plans, Xs, Ys = main_run.pop("_plans")
    ms, ms1, ms_ev, kus = main_run["ms"], main_run["ms1"], main_run["ms_ev"], main_run["kus"]
    relerr_dense = main_run["relerr_dense"]
    Yb = Ys[0].cpu().numpy().astype(np.float64)
    # plain BF16 storage (half the factor bytes of the split form), reported
    # at its own error against the dense FP64 kernel
    bf16_line = None
    if args.storage != "bf16" and not args.no_bf16:
        r16 = measure_batched(hm, torch, f, RHS, "bf16", args, streams, rows, Yd)
        for p in r16.pop("_plans")[0]:
            p.close()
        kb16 = kernel_bytes(f, "bf16", nrhs, 4)
        g16 = kb16["rows"] / (r16["kus"]["rows"] * 1e-6) / 1e9
        bf16_line = {"H-matvec/s": nrhs / (r16["ms"] * 1e-3), "us_per_product": r16["ms"] * 1e3,
                     "single_stream_us": r16["ms1"] * 1e3,
                     "step_frac": algorithmic_bytes(f, "bf16", nrhs, 4) / (r16["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "kb_rows_us": r16["kus"]["rows"], "kb_rows_GBs": g16, "kb_rows_frac": g16 / HBM_PEAK_GBS,
                     "relerr_vs_dense_fp64": r16["relerr_dense"]}

    # plain FP64 path: one matvec per right-hand side.  The same S-stream
    # schedule: the 32 columns of a step are 32 independent FP64 matvecs,
    # column j on stream j % S with its own plan (a plan's scratch serves one
    # stream at a time)
    p64s = [hm.DeviceHmat(f, "f64", {"no_batched": True}) for _ in range(S)]
    p64 = p64s[0]
    X64 = p64.new_cols(N, nrhs)
    X64.copy_(torch.from_numpy(RHS).to(dev))
    Y64 = p64.new_cols(N, nrhs)
    cols_x = [X64[:, j] for j in range(nrhs)]
    cols_y = [Y64[:, j] for j in range(nrhs)]

    def step_plain(i, st):
        for j in range(nrhs):
            p64s[j % S].matvec(cols_x[j], cols_y[j], stream=streams[j % S])

    ms64 = timed_streams(torch, step_plain, args.plain_steps, 1, streams)
    Yp = Y64.cpu().numpy()
    relerr_h64 = float(np.linalg.norm(Yb - Yp) / np.linalg.norm(Yp))
    relerr_dense64 = float(np.linalg.norm(Yp[rows] - Yd) / np.linalg.norm(Yd))
    log(f"plain f64 ({S} streams): {ms64 * 1e3:.1f} us per {nrhs} rhs; relerr vs dense {relerr_dense64:.3e}; "
        f"batched vs plain {relerr_h64:.3e}")
    # FP64 H*X through hm_matmat (the reference's own precision for a matrix
    # of right-hand sides): one pass over the FP64 factors per 32 columns on
    # the f64 MFMA, two plan replicas rotated; beside it the same plans'
    # single-column layout (panels of 8 columns on VALU)
    pm = [hm.DeviceHmat(f, "f64") for _ in range(2)]
    Xm = [X64, pm[0].new_cols(N, nrhs)]
    Xm[1].copy_(X64)
    Ym = [pm[0].new_cols(N, nrhs) for _ in range(2)]
    ms64m = timed(torch, lambda i: pm[i % 2].matmat(Xm[i % 2], Ym[i % 2], stream=stream), args.plain_steps * 5,
                  2, stream)
    for q in pm:
        q.timing(True)
        q.timing_read()
    timed(torch, lambda i: pm[i % 2].matmat(Xm[i % 2], Ym[i % 2], stream=stream), args.plain_steps * 5, 0, stream)
    k64 = {"proj_ms": 0.0, "reduce_ms": 0.0, "rows_ms": 0.0, "launches": 0}
    for q in pm:
        r = q.timing_read()
        for k in k64:
            k64[k] += r[k]
    k64us = {k: k64[f"{k}_ms"] * 1e3 / max(k64["launches"], 1) for k in ("proj", "reduce", "rows")}
    relerr_m64 = float(np.linalg.norm(Ym[0].cpu().numpy() - Yp) / np.linalg.norm(Yp))
    ms64p = timed(torch, lambda i: p64.matmat(X64, Y64, stream=stream), args.plain_steps, 1, stream)
    log(f"FP64 H*X ({nrhs} rhs, hm_matmat, f64 MFMA, one stream): {ms64m * 1e3:.1f} us (proj {k64us['proj']:.1f}, "
        f"reduce {k64us['reduce']:.1f}, rows {k64us['rows']:.1f}); vs per-column {relerr_m64:.1e}; "
        f"panels of 8 on the single-column layout: {ms64p * 1e3:.1f} us")
    kb64 = kernel_bytes(f, "f64", nrhs, 8)
    for q in pm:
        q.close()

    # drop-in path with host buffers (PCIe-inclusive; never the reported value)
    t0 = time.perf_counter()
    plans[0].mul(RHS)
    torch.cuda.synchronize()
    pcie_s = time.perf_counter() - t0
    log(f"host-buffer H*X (upload + product + download): {pcie_s * 1e3:.1f} ms")

    c4 = None if args.no_c4 else c4_report(hm, torch, stream)
    extras = extras_report(hm, torch, stream, plans, Xs, Ys) if args.extras else None
    if args.extras:
        for p in plans:             # the N = 2^20 replicas take their place
            p.close()
        extras.update(scale_report(hm, torch, args, streams))

    if args.calibrate_copy_mb > 0:
        n = args.calibrate_copy_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device=dev)
        b_ = torch.empty_like(a)
        for _ in range(3):
            b_.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration: 3 copies of {args.calibrate_copy_mb} MB done")

    bytes_b = algorithmic_bytes(f, args.storage, nrhs, 4)
    bytes_64 = algorithmic_bytes(f, "f64", 1, 8)
    kb = kernel_bytes(f, args.storage, nrhs, 4)
    kgbs = {k: (kb[k] / (kus[k] * 1e-6) / 1e9 if kus[k] > 0 else None) for k in kb}
    achieved = kgbs["rows"]
    step_gbs = bytes_b / (ms * 1e-3) / 1e9
    value = nrhs / (ms * 1e-3)
    plain_mv_s = nrhs / (ms64 * 1e-3)
    cpu = cpu_mt = None
    if args.cpu_seconds > 0:
        rate, n, dt = cpu_baseline(f, args.cpu_seconds)
        cpu = {"value": rate, "unit": "H-matvec/s", "cores": 1, "kind": "port",
               "sample": f"{n} FP64 oracle matvecs (1 RHS each) on the same factors in {dt:.1f}s"}
        log(f"cpu baseline: {rate:.3f} matvec/s")
        nt = host_threads()
        rate_mt, n_mt, dt_mt = cpu_baseline(f, args.cpu_seconds / 2, nt)
        cpu_mt = {"value": rate_mt, "unit": "H-matvec/s", "cores": nt, "kind": "port",
                  "sample": f"{n_mt} FP64 oracle matvecs (1 RHS each, {nt} OpenMP threads) in {dt_mt:.1f}s"}
        log(f"cpu baseline, {nt} threads: {rate_mt:.3f} matvec/s")
    line = {
        "metric": "H-matvec/s (N=262144, 32 RHS, BF16-split batched, relerr<=1e-5)",
        "value": value,
        "unit": "H-matvec/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x2": "bf16x2-storage/f32-acc", "bf16": "bf16-storage/f32-acc",
                  "f32": "f32"}[args.storage],
        "data": "synthetic",
        "config": {"workload": f"H*X exp(-|x-y|) N={N} nrhs={nrhs} MaxRank=24 MinBlock=64 "
                               f"MaxBlock=N/4 weak-adm, {args.storage} storage",
                   "N": N, "nrhs": nrhs, "leaves": f.nblocks, "dense_leaves": f.counts()[0],
                   "rk_leaves": f.counts()[1], "replicas": R, "streams": S},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(args.storage, "kb_rows"),
                     "kernel": "kb_rows (batched row pass)", "algorithmic_bytes": kb["rows"],
                     "us_per_launch": kus["rows"]},
        "kernels": {"kb_proj": {"us": kus["proj"], "algorithmic_bytes": kb["proj"], "GB/s": kgbs["proj"],
                                "traffic": profiled_traffic(args.storage, "kb_proj")},
                    "kb_reduce": {"us": kus["reduce"], "algorithmic_bytes": 0,
                                  "traffic": profiled_traffic(args.storage, "kb_reduce")},
                    "kb_rows": {"us": kus["rows"], "algorithmic_bytes": kb["rows"], "GB/s": kgbs["rows"],
                                "traffic": profiled_traffic(args.storage, "kb_rows")}},
        "products_in_flight": S,
        "single_stream_ms_per_step": ms1,
        "single_stream_H-matvec/s": nrhs / (ms1 * 1e-3),
        "ms_per_step_with_kernel_events": ms_ev,
        "step_GBs": step_gbs,
        "step_frac": step_gbs / HBM_PEAK_GBS,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_mt,
        "relerr_vs_dense_fp64": relerr_dense,
        "relerr_vs_fp64_hmatrix": relerr_h64,
        "plain_fp64": {"H-matvec/s": plain_mv_s, "us_per_matvec": ms64 * 1e3 / nrhs,
                       "GB/s": bytes_64 / (ms64 * 1e-3 / nrhs) / 1e9,
                       "relerr_vs_dense_fp64": relerr_dense64},
        "speedup_vs_per_column_fp64": value / plain_mv_s,
        "speedup_vs_fp64_matmat": value / (nrhs / (ms64m * 1e-3)),
        "fp64_matmat": {"us_per_product": ms64m * 1e3, "H-matvec/s": nrhs / (ms64m * 1e-3),
                        "relerr_vs_per_column": relerr_m64,
                        "GB/s": algorithmic_bytes(f, "f64", nrhs, 8) / (ms64m * 1e-3) / 1e9,
                        "kernels_us": k64us,
                        "rows_GBs": kb64["rows"] / (k64us["rows"] * 1e-6) / 1e9 if k64us["rows"] > 0 else None,
                        "proj_GBs": kb64["proj"] / (k64us["proj"] * 1e-6) / 1e9 if k64us["proj"] > 0 else None,
                        "single_layout_panels_us": ms64p * 1e3,
                        "traffic": {k: profiled_traffic("f64", k) for k in ("kb_proj64", "kb_reduce", "kb_rows64")}},
        "host_buffer_product_ms": pcie_s * 1e3,
        "algorithmic_bytes_per_step": bytes_b,
    }
    if bf16_line is not None:
        line["bf16"] = bf16_line
    if c4 is not None:
        line["c4"] = c4
    if extras is not None:
        line["extras"] = extras
    print(json.dumps(line), flush=True)
    for p in plans + p64s:
        p.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
