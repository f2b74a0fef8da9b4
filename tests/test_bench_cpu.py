"""bench.py's host-side contract, checked without a GPU: the one-GPU rule,
the profiled HBM traffic it reports for the default workload, and the
algorithmic byte count behind roofline.achieved."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from spx import workloads as W  # noqa: E402

M = N = 5_558_326  # the circuit-like matrix of the default workload
NNZ = 59_543_688


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one MI355X" in msg["error"]


def test_default_workload_has_profiled_traffic():
    """The default run (kernel 4, 2 chunks per unit, +-1024 band) finds its
    FETCH_SIZE + WRITE_SIZE figure in profiles/; it must be a plausible
    share of the algorithmic bytes (the 12-bit columns save ~149 MB)."""
    t = bench.profiled_traffic(4, 2, 1024)
    B = W.bytes_model(M, N, NNZ)
    assert t is not None and 0.7 * B < t < 1.1 * B, (t, B)
    assert bench.profiled_traffic(4, 2, 0) is None  # no profile for the uniform-column variant


def test_algorithmic_bytes_closed_form():
    """B = 4(m+1) + 12 nnz + 8n + 8m for FP64 and beta = 0 (825,690,780 at the
    default workload); FP32 values take 4 bytes; beta != 0 reads y too;
    SpMM scales the x and y terms by k."""
    assert W.bytes_model(M, N, NNZ) == 825_690_780
    assert W.bytes_model(M, N, NNZ, value_bytes=4) == 825_690_780 - 4 * NNZ
    assert W.bytes_model(M, N, NNZ, beta_nonzero=True) == 825_690_780 + 8 * M
    assert W.bytes_model(M, N, NNZ, k=8) == 4 * (M + 1) + 12 * NNZ + 8 * N * 8 + 8 * M * 8
