"""Benchmark: plain FP64 GPU SLQ at scale on one MI355X.

Workload (one "step"): ``imate_amd.logdet(A, method='slq', gpu=True,
lanczos_degree=30, min_num_samples=256, max_num_samples=256)`` on a seeded
sparse SPD correlation matrix with n = 2^20 rows and about 50 non-zeros per
row (FP64 CSR), i.e. 256 Rademacher probes in four 64-probe blocks,
256 x 30 = 7680 Lanczos matvec-steps per step. The matrix is uploaded once
(``imate_amd.Matrix``) before the timed region, so inputs are resident in HBM.

Printed (one JSON line): matvec-steps/s, the SpMM kernel's roofline figures
from HIP events recorded around every SpMM launch of the timed steps, the
logdet error against a Cholesky reference when one is recorded for this exact
matrix, and the CPU baseline: the oracle's C restatement of imate's CPU SLQ
(long-double CSR matvec, OpenMP over samples) timed on this host's cores on a
bounded sample of the same workload.
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print('[bench] ' + msg, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=5)
    p.add_argument('--warmup', type=int, default=1)
    p.add_argument('--n', type=int, default=2**20)
    p.add_argument('--nnz-per-row', type=float, default=50.0)
    p.add_argument('--lanczos-degree', type=int, default=30)
    p.add_argument('--samples', type=int, default=256)
    p.add_argument('--seed', type=int, default=0)
    p.add_argument('--func', default='logdet', choices=['logdet', 'traceinv'])
    p.add_argument('--no-cpu', action='store_true',
                   help='skip the CPU baseline leg')
    p.add_argument('--cpu-threads', type=int, default=0)
    p.add_argument('--copy-calibration', action='store_true',
                   help='only run the plain copy kernel (rocprof calibration)')
    return p.parse_args()


def spmm_bytes(n, nnz, b=64, vbytes=8):
    """Algorithmic HBM bytes of one fused Lanczos SpMM launch: the CSR arrays
    once (values, int32 columns, int64 row pointers), the gathered operand
    block once, the normalised basis block written and the product written."""
    return nnz * (vbytes + 4) + (n + 1) * 8 + 3 * n * b * vbytes


def cholesky_reference(n, k, seed):
    path = os.path.join(HERE, 'tests', 'golden', 'cholesky_logdet.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    for row in table.get('entries', []):
        if row['n'] == n and abs(row['nnz_per_row'] - k) < 1e-9 and \
                row['seed'] == seed:
            return row['logdet']
    return None


def cpu_baseline(A, m, threads, seed):
    """Oracle restatement of imate's CPU SLQ recurrences on `threads` host
    threads, two probes per thread (same matrix, same lanczos_degree)."""
    from oracle import oracle
    op = oracle.Operator(A)
    eps = float(numpy.finfo(numpy.float64).eps)
    # one short warm-up to page in the library and matrix
    oracle.recurrences(op, seed, 1, 2, eps, 0, num_threads=1)
    t = time.perf_counter()
    a, b, sizes = oracle.recurrences(op, seed, 2 * threads, m, eps, 0,
                                     num_threads=threads)
    dt = time.perf_counter() - t
    steps = int(numpy.sum(sizes))
    return steps / dt, dt, steps


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU '
              'path (probes would shard across GPUs with one gather of the '
              'per-sample results; not built). Use --gpus 1.', flush=True)
        sys.exit(2)

    import imate_amd
    from imate_amd import _lib
    lib = _lib.load()

    if args.copy_calibration:
        ms = ctypes.c_double(0)
        nbytes = 1 << 30
        _lib.check(lib.slq_copy_bandwidth(nbytes, 20, ctypes.byref(ms)),
                   'slq_copy_bandwidth')
        print(json.dumps({'copy_bytes': nbytes, 'ms_per_launch': ms.value,
                          'GBps': 2 * nbytes / ms.value / 1e6}), flush=True)
        return

    from imate_amd.sample_matrices import sparse_correlation_matrix
    n, k, m, s = args.n, args.nnz_per_row, args.lanczos_degree, args.samples
    t = time.perf_counter()
    A = sparse_correlation_matrix(n, nnz_per_row=k, seed=args.seed)
    log('matrix n=%d nnz=%d (%.1f/row) built in %.1fs'
        % (n, A.nnz, A.nnz / n, time.perf_counter() - t))

    Aop = imate_amd.Matrix(A)
    fn = imate_amd.logdet if args.func == 'logdet' else imate_amd.traceinv
    kw = dict(method='slq', gpu=True, lanczos_degree=m, min_num_samples=s,
              max_num_samples=s, seed=args.seed, return_info=True)
    t = time.perf_counter()
    Aop.get_linear_operator(gpu=True)
    log('upload %.2fs' % (time.perf_counter() - t))

    for i in range(args.warmup):
        val, info = fn(Aop, **kw)
        log('warmup %d: %s=%.6f alg %.3fs' % (i, args.func, val,
                                             info['time']['alg_wall_time']))

    handle = Aop.gpu_Aop.handle
    _lib.check(lib.slq_profile_enable(handle, 1), 'slq_profile_enable')
    vals = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        val, info = fn(Aop, **kw)
        vals.append(val)
    elapsed = time.perf_counter() - t0
    spmm_ms = ctypes.c_double(0)
    spmm_n = ctypes.c_int64(0)
    upd_ms = ctypes.c_double(0)
    upd_n = ctypes.c_int64(0)
    _lib.check(lib.slq_profile_read(handle, ctypes.byref(spmm_ms),
                                    ctypes.byref(spmm_n), ctypes.byref(upd_ms),
                                    ctypes.byref(upd_n)), 'slq_profile_read')
    lib.slq_profile_enable(handle, 0)

    steps_per_call = s * m
    value = steps_per_call * args.steps / elapsed
    ms_per_step = elapsed / args.steps * 1e3
    avg_spmm = spmm_ms.value / max(spmm_n.value, 1)
    avg_upd = upd_ms.value / max(upd_n.value, 1)
    bspmm = spmm_bytes(n, A.nnz)
    achieved = bspmm / (avg_spmm * 1e-3) / 1e9
    upd_bytes = 4 * n * 64 * 8
    log('timed %d steps: %.1f ms/step, %.0f matvec-steps/s; SpMM %.3f ms '
        '(%d launches), update %.3f ms (%d)' % (args.steps, ms_per_step, value,
                                                avg_spmm, spmm_n.value,
                                                avg_upd, upd_n.value))

    est = float(numpy.mean(vals))
    ref = cholesky_reference(n, k, args.seed) if args.func == 'logdet' else None
    accuracy = {'estimate': est,
                'mc_error_95': float(info['error']['absolute_error'])}
    if ref is not None:
        accuracy['cholesky_logdet'] = ref
        accuracy['rel_error_vs_cholesky'] = abs(est - ref) / abs(ref)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        rate, dt, steps = cpu_baseline(A, m, threads, args.seed)
        cpu = {'value': rate, 'unit': 'matvec-steps/s', 'cores': threads,
               'kind': 'port',
               'sample': '%d probes x %d Lanczos steps (two probes per '
                         'thread) on the same n=%d matrix: %d steps in %.1fs'
                         % (2 * threads, m, n, steps, dt)}
        log('cpu baseline: %.1f matvec-steps/s on %d threads' % (rate, threads))

    out = {
        'metric': 'Lanczos matvec-steps/s (plain FP64 GPU SLQ logdet)',
        'value': value,
        'unit': 'matvec-steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (seeded sparse correlation matrix)',
        'config': {'workload': 'slq-%s-correlation-n%d-nnz%.0f-m%d-s%d-b64'
                               % (args.func, n, k, m, s),
                   'n': n, 'nnz': int(A.nnz), 'lanczos_degree': m,
                   'samples': s, 'block': 64, 'parallelism': 'single-gpu'},
        'roofline': {'bound': 'hbm', 'achieved': achieved,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': None,
                     'kernel': 'spmm_kernel<double> (fused Lanczos SpMM)',
                     'bytes_per_launch': bspmm,
                     'avg_launch_ms': avg_spmm},
        'update_kernel': {'avg_launch_ms': avg_upd,
                          'GBps': upd_bytes / (avg_upd * 1e-3) / 1e9},
        'accuracy': accuracy,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
