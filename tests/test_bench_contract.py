"""bench.py on the CPU: the one-GPU contract (--gpus above 1 exits with a
message and status 2) and the algorithmic-byte formulas behind
roofline.achieved, checked against the per-iteration figures worked out by
hand for the 256x256 workload (131 070 velocity equations, 2 343 944 K
entries, 65 536 pressure unknowns)."""
import os
import subprocess
import sys

import bench

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NV, NNZ, NP = 131070, 2343944, 65536


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_fp64_bytes_per_iteration():
    sizes = [(NV, NNZ, NP)]
    # one inner Jacobi-PCG iteration: nnz (8 + 4) + 11 n_v 8 = 28.13 MB + 11.53 MB
    b_cg = bench.fp64_bytes({"inner": 1, "uzawa": 0}, sizes)
    assert b_cg == NNZ * 12 + 11 * NV * 8
    assert abs(b_cg / 1e6 - 39.66) < 0.01
    # one Uzawa iteration: G s and G^T u (nnz_G = 4 n_v, 12 B each, twice), 7 n_v and 12 n_p doubles
    b_uz = bench.fp64_bytes({"inner": 0, "uzawa": 1}, sizes)
    assert b_uz == 2 * 4 * NV * 12 + 7 * NV * 8 + 12 * NP * 8
    assert abs(b_uz / 1e6 - 26.2) < 0.05
    # linear in the counts
    assert bench.fp64_bytes({"inner": 3, "uzawa": 2}, sizes) == 3 * b_cg + 2 * b_uz


def test_mixed_bytes_per_iteration():
    # finest level only (no coarse levels): FP64 part nnz 12 + n_v 88 and no
    # V-cycle traffic (a single level is preconditioned by Jacobi)
    sizes = [(NV, NNZ, NP)]
    b = bench.mixed_bytes({"inner": 1, "uzawa": 0}, sizes)
    fp64 = NNZ * 12 + NV * 88
    assert b == fp64
    # a level above the coarsest adds its K once per leg (two legs), 28 B of
    # vectors per row and 8 B per row of the level below
    n2, z2 = 32768, 580000
    b2 = bench.mixed_bytes({"inner": 1, "uzawa": 0}, [(n2, z2, NP // 4), (NV, NNZ, NP)])
    assert b2 == fp64 + 2 * NNZ * 8 + 28 * NV + 8 * n2
    b3 = bench.mixed_bytes({"inner": 1, "uzawa": 0}, [(8, 64, 4), (n2, z2, NP // 4), (NV, NNZ, NP)])
    assert b3 == b2 + 2 * z2 * 8 + 28 * n2 + 8 * 8
    # 256x256 with its seven coarser levels: about 95 MB per inner iteration
    # (39.7 MB FP64 + 55 MB V-cycle)
    levels = [(NV // 4 ** k, NNZ // 4 ** k, NP // 4 ** k) for k in range(7, -1, -1)]
    per = bench.mixed_bytes({"inner": 1, "uzawa": 0}, levels)
    assert 93e6 < per < 97e6
