"""The host layout builder (hmatrix.jl_amd/csrc/hm_layout.cpp) built for the
CPU alone with AddressSanitizer and UndefinedBehaviorSanitizer, run on the
block tables of the GPU tests: every single-right-hand-side layout (both
directions, four storages, projection chunks of 64 and 1024 rows, with and
without the one-launch member order) and every batched layout (four
storages; three row-pass shapes for FP32 and BF16, FP64's own) is built and checked offset by offset
against the arena it indexes (check_direction / check_batched).  A sanitizer
report or a violated check fails the test.

The tables: the narrow-leaf comb tables, the wide-rank table and the
random rectangular partitions of the GPU tests, C1 (src/cluster.jl:48-133 partition, weak
admissibility), a strong-admissibility exp(-|x-y|) matrix, an uneven batch
of five matrices concatenated block-diagonally as hm_plan_create_batch does,
and a 5-column table, for which the batched layout must be refused (the
plan then reports ``batched`` 0 and multiplies panels on the single-column
layout)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.helpers import build, c1, c4_members, comb_flat, expk, random_partition_flat, sorted_uniform, wide_rank_flat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "hmatrix.jl_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    out = str(tmp_path_factory.mktemp("layout") / "layout_check")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-Werror",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           "-I", CSRC, os.path.join(CSRC, "hm_layout.cpp"), os.path.join(ROOT, "tests", "native", "layout_check.cpp"),
           "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def _write(path, nrows, ncols, blocks, data):
    blocks = np.ascontiguousarray(blocks, dtype=np.int64).reshape(-1, 8)
    data = np.ascontiguousarray(data, dtype=np.float64)
    with open(path, "wb") as f:
        np.array([nrows, ncols, blocks.shape[0], data.size], dtype=np.int64).tofile(f)
        blocks.tofile(f)
        data.tofile(f)


def _run(checker, tmp_path, nrows, ncols, blocks, data, members=None, plant=True):
    """Check every layout of the table; with plant, also require the checks
    to report each mistake planted in a copy of each layout."""
    table = str(tmp_path / "table.bin")
    _write(table, nrows, ncols, blocks, data)
    args = [checker, table]
    if members is not None:
        mr, mc = members
        mpath = str(tmp_path / "members.bin")
        np.concatenate([[len(mr) - 1], mr, mc]).astype(np.int64).tofile(mpath)
        args.append(mpath)
    else:
        args.append("-")
    if plant:
        args.append("plant")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "all layouts checked" in r.stdout
    if plant:
        line = [x for x in r.stdout.splitlines() if x.startswith("planted")][0]
        n = int(line.split()[1])
        assert n >= 8 and line.endswith(f"caught {n}"), line
    return r.stdout


def _batched_flags(out):
    return [line for line in out.splitlines() if "batched H" in line]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_comb_tables(hm, checker, tmp_path, seed):
    f = comb_flat(hm, seed, nblk=70)
    out = _run(checker, tmp_path, f.nrows, f.ncols, f.blocks, f.data)
    assert all("batched H 1" in line for line in _batched_flags(out))


def test_wide_rank_table(hm, checker, tmp_path):
    f = wide_rank_flat(hm, 0)
    _run(checker, tmp_path, f.nrows, f.ncols, f.blocks, f.data)


@pytest.mark.parametrize("seed,side,p_dense,kmax", [(101, (24, 400), 0.4, 40), (102, (24, 400), 0.4, 40),
                                                    (103, (24, 400), 0.4, 40), (104, (24, 400), 0.4, 40),
                                                    (105, (8, 128), 0.7, 8), (106, (64, 1500), 0.2, 40)])
def test_random_partitions(hm, checker, tmp_path, seed, side, p_dense, kmax):
    """Seeded random exact tilings of rectangular matrices (the tables of
    the GPU random-partition parity test): every layout checks, and every
    planted mistake is caught."""
    rng = np.random.default_rng(seed)
    nr, nc = int(rng.integers(300, 3000)), int(rng.integers(300, 3000))
    f = random_partition_flat(hm, seed, nr, nc, side, p_dense, kmax)
    out = _run(checker, tmp_path, f.nrows, f.ncols, f.blocks, f.data)
    assert all("batched H 1" in line for line in _batched_flags(out))


def test_c1_weak(hm, checker, tmp_path):
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    _run(checker, tmp_path, f.nrows, f.ncols, f.blocks, f.data)


def test_strong_admissibility(hm, checker, tmp_path):
    """Leaves far from the diagonal: the automatic shape must pick the wide
    row pass and its X windows must still hold every dense piece."""
    par, c, H = build(hm, sorted_uniform(4096, 0.0, 400.0, 3), expk, 512, 64, 12, adj="strong")
    f = hm.flatten(H)
    out = _run(checker, tmp_path, f.nrows, f.ncols, f.blocks, f.data, plant=False)
    assert any("batched H 1" in line for line in _batched_flags(out))


def test_uneven_batch_members(hm, checker, tmp_path):
    """Five matrices on a block diagonal (offsets as hm_plan_create_batch
    concatenates them), with the one-launch member order: each member's
    row tiles read only t entries of its own projections."""
    ms = c4_members(hm, nmat=5, nlo=300, nhi=3000, seed=6)
    blocks, data, r0, c0, d0 = [], [], 0, 0, 0
    mr, mc = [0], [0]
    for m in ms:
        b = np.array(m.blocks, dtype=np.int64).copy()
        b[:, 1] += r0
        b[:, 2] += c0
        b[:, 6] += d0
        b[:, 7] += d0
        blocks.append(b)
        data.append(np.asarray(m.data))
        r0, c0, d0 = r0 + m.nrows, c0 + m.ncols, d0 + len(m.data)
        mr.append(r0)
        mc.append(c0)
    _run(checker, tmp_path, r0, c0, np.concatenate(blocks), np.concatenate(data), (mr, mc))


def test_narrow_matrix_refuses_batched_layout(hm, checker, tmp_path):
    """Fewer than 8 columns: no batched layout for H (an octet of X rows
    must fit in X); H' (5 x 40) still gets one."""
    rng = np.random.default_rng(3)
    C = rng.standard_normal((40, 5))
    out = _run(checker, tmp_path, 40, 5, [(0, 0, 0, 40, 5, 0, 0, 0)], C.ravel(order="F"), plant=False)
    flags = _batched_flags(out)
    assert flags and all("batched H 0" in line and "H' 1" in line for line in flags), flags
