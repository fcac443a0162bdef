"""Calibrates the CPU baseline: the oracle's C restatement of imate's CPU
Lanczos loop (libslq_oracle.so, what bench.py times on the GPU machine's
host as ``cpu_baseline``) against imate's own CPU sources compiled into
oracle/_ref (``make -C oracle ref``; only where /root/reference exists, so
this runs in the build container only).

Both run the same work on the same host cores: T threads, each running P
probes of lanczos_degree Lanczos steps on the bench's n = 2^20 correlation
matrix (imate parallelises its Monte-Carlo loop the same way: one sample
per OpenMP thread at a time, c_trace_estimator.cpp:191-307). imate's loop
runs from T Python threads calling its one-thread driver (the ctypes call
releases the interpreter lock), each with its own seed; the restatement
runs its OpenMP driver with T threads. Writes oracle/cpu_calibration.json:
steps/s of each and their ratio, which bench.py reports beside its
``cpu_baseline``.

Usage: python oracle/calibrate_cpu_baseline.py [--n 1048576] [--threads 8]
"""

import argparse
import ctypes
import json
import os
import platform
import sys
import threading
import time

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle  # noqa: E402


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def reference_rate(R, A, m, threads, per_thread, seed):
    dat = numpy.ascontiguousarray(A.data, dtype=numpy.float64)
    idx = A.indices.astype(numpy.int32)
    ip = A.indptr.astype(numpy.int32)
    n = A.shape[0]
    out = [None] * threads
    f = R.ref_slq_recurrences_f64
    f.restype = ctypes.c_int
    tol = float(numpy.finfo(numpy.float64).eps)

    def work(t):
        a = numpy.zeros((per_thread, m))
        b = numpy.zeros((per_thread, m))
        sz = numpy.zeros(per_thread, numpy.int32)
        f(_ptr(dat), _ptr(idx), _ptr(ip), n, n, 0, ctypes.c_double(0.0),
          None, None, None, 0, ctypes.c_longlong(seed + t), per_thread, m,
          ctypes.c_double(tol), 0, _ptr(a), _ptr(b), _ptr(sz), None)
        out[t] = int(sz.sum())

    ths = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    t0 = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    dt = time.perf_counter() - t0
    return sum(out) / dt, dt, sum(out)


def port_rate(A, m, threads, per_thread, seed):
    op = oracle.Operator(A)
    tol = float(numpy.finfo(numpy.float64).eps)
    t0 = time.perf_counter()
    a, b, sizes = oracle.recurrences(op, seed, per_thread * threads, m, tol, 0,
                                     num_threads=threads)
    dt = time.perf_counter() - t0
    return int(sizes.sum()) / dt, dt, int(sizes.sum())


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def main():
    p = argparse.ArgumentParser()
    p.add_argument('--n', type=int, default=1 << 20)
    p.add_argument('--threads', type=int, default=8)
    p.add_argument('--per-thread', type=int, default=2)
    p.add_argument('--lanczos-degree', type=int, default=30)
    p.add_argument('--repeats', type=int, default=2)
    args = p.parse_args()
    R = oracle.ref_lib()
    if R is None:
        sys.exit('oracle/_ref is not built (make -C oracle ref)')
    from imate_amd.sample_matrices import sparse_correlation_matrix
    A = sparse_correlation_matrix(args.n, nnz_per_row=50.0, seed=0)
    m, T, P = args.lanczos_degree, args.threads, args.per_thread
    # warm-up: page in both libraries and the matrix
    reference_rate(R, A, 2, T, 1, 0)
    port_rate(A, 2, T, 1, 0)
    ref, port = [], []
    for r in range(args.repeats):
        ref.append(reference_rate(R, A, m, T, P, 100 * r))
        port.append(port_rate(A, m, T, P, 100 * r))
        print('repeat %d: reference %.1f steps/s (%.1fs), port %.1f steps/s '
              '(%.1fs)' % (r, ref[-1][0], ref[-1][1], port[-1][0], port[-1][1]),
              flush=True)
    rr = max(x[0] for x in ref)
    pr = max(x[0] for x in port)
    out = {'n': args.n, 'nnz': int(A.nnz), 'lanczos_degree': m,
           'threads': T, 'probes_per_thread': P,
           'reference_steps_per_s': rr, 'port_steps_per_s': pr,
           'port_over_reference': pr / rr,
           'host': cpu_model(),
           'cpu_count': os.cpu_count(),
           'note': 'best of %d repeats; reference = imate CPU sources built '
                   'into oracle/_ref, port = libslq_oracle.so' % args.repeats}
    with open(os.path.join(HERE, 'cpu_calibration.json'), 'w') as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
