"""Pieces of bench.py's contract that need no GPU: the algorithmic byte
count of one SpMM launch against the arrays it reads, the recorded traffic
of the headline workload, and the refusal of more than one GPU."""

import json
import os
import subprocess
import sys

import numpy

import bench
from imate_amd.sample_matrices import sparse_correlation_matrix

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_spmm_bytes_are_the_arrays_one_launch_reads_and_writes():
    """CSR values, int32 column indices and int64 row pointers once, the
    64-probe operand block read once and the product block written once."""
    A = sparse_correlation_matrix(3000, nnz_per_row=20, seed=1)
    n = A.shape[0]
    X = numpy.zeros((n, 64))
    want = (A.data.nbytes + A.indices.astype(numpy.int32).nbytes +
            numpy.zeros(n + 1, dtype=numpy.int64).nbytes + 2 * X.nbytes)
    assert bench.spmm_bytes(n, A.nnz) == want
    Xf = numpy.zeros((n, 64), dtype=numpy.float32)
    want32 = (A.data.astype(numpy.float32).nbytes +
              A.indices.astype(numpy.int32).nbytes + 8 * (n + 1) + 2 * Xf.nbytes)
    assert bench.spmm_bytes(n, A.nnz, vbytes=4) == want32


def test_headline_figures_are_consistent():
    """The headline config's algorithmic count (1.722 GB, plus v_{j-1}'s
    block in 29 of 30 launches = 2.241 GB) and its recorded traffic (the two
    counter runs, each scaled by its copy-kernel factor) add up."""
    n, nnz, m = 1 << 20, 53307360, 30
    b = bench.spmm_bytes(n, nnz)
    assert b == 1721818760
    assert abs(b + n * 64 * 8 * (m - 1) / m - 2.2408e9) < 1e6
    with open(os.path.join(ROOT, 'profiles', 'spmm_traffic.json')) as f:
        rec = json.load(f)
    assert bench.measured_traffic(n, nnz) == rec['bytes_per_launch']
    assert abs(rec['fetch_bytes'] + rec['write_bytes'] - rec['bytes_per_launch']) < 1.0
    assert bench.measured_traffic(n, nnz + 1) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert 'one GPU only' in r.stdout
