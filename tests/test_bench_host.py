"""Host-side pieces of bench.py (no GPU): the refusal of several GPUs, the
byte counts the roofline uses, and the traffic figure read from the committed
counter summary."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_several_gpus_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_spmv_byte_counts():
    # the n=2 64^3 workload: 6,340,608 rows, 88,064,000 nonzeros
    assert bench.spmv_bytes(6340608, 88064000, True) == 780599300
    assert bench.spmv_bytes(6340608, 88064000, False) == 12 * 88064000 + 4 * 6340609 + 16 * 6340608
    # block SELL, n=2: per face 6 blocks of 4 FP32 values + index, 2 diagonal-block entries, x and y
    faces = 3170304
    got = bench.block_sell_bytes(2, faces, 6.0, True)
    assert got == 6 * faces * (16 + 4) + faces * 2 * 4 + 2 * 2 * faces * 4
    # the iteration's launches also read the fused dot products' operands
    assert bench.block_sell_bytes(2, faces, 6.0, True, dot_operands=1.0) == got + 2 * faces * 4
    assert bench.block_sell_bytes(2, faces, 6.0, False, dot_operands=0.5) == (
        bench.block_sell_bytes(2, faces, 6.0, False) + faces * 8)
    assert bench.block_sell_bytes(2, faces, 6.0, False) == 6 * faces * (32 + 4) + faces * 2 * 8 + 2 * 2 * faces * 8


def test_traffic_from_profile():
    t = bench.profiled_traffic()
    assert t is not None and 4e8 < t < 6e8


def test_assembly_byte_count():
    # n=2 tetrahedra, 64^3 cubes: 1,572,864 cells, 3,170,304 faces, 88,064,000 nonzeros
    C, F, nnz = 1572864, 3170304, 88064000
    got = bench.assembly_bytes(2, 4, C, F, nnz)
    # values + rhs written; per cell b (n^2 * 14 packed entries and row sums) and R_ijKe
    # (n^2 * 4), R_iK and f (2n), w (n * 4); incidence
    per_cell = 8 * (4 * 14 + 4 * 4 + 2 * 2 + 2 * 4)
    assert got == 8 * nnz + 8 * 2 * F + C * per_cell + 8 * F + 16 * C
    assert abs(got - 1.863e9) < 1e7
