"""bench.py contract checks that need no GPU: the single-GPU refusal and the algorithmic
byte count of the SpMV that its roofline line divides by."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


@pytest.mark.parametrize("n", [2, 8])
def test_bench_refuses_several_gpus(n):
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_spmv_algorithmic_bytes_of_the_128_contact_system():
    import bench

    # 128^3 contact system: 6 291 456 DOF in 2 097 152 block rows, 30 838 620 stored 3 x 3
    # blocks; per block 9 doubles + one int32 column, per block row one int32 pointer (+1),
    # x read and y written once
    info = {"block_size": 3, "nnz_blocks": 30_838_620, "n": 6_291_456}
    assert bench.spmv_bytes(info) == 2_452_787_028
    # FP32 values: 4-byte entries, the vectors stay FP64
    assert bench.spmv_bytes(info, value_bytes=4) == 2_452_787_028 - 30_838_620 * 9 * 4
