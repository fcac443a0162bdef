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
import scipy.sparse

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
    p.add_argument('--no-forms', action='store_true',
                   help='skip the compact precision / batched / surrogate '
                        'legs of the default run')
    p.add_argument('--cpu-threads', type=int, default=0)
    p.add_argument('--copy-calibration', action='store_true',
                   help='only run the plain copy kernel (rocprof calibration)')
    p.add_argument('--dump', default=None,
                   help='surrogate form: write the held-out and validation '
                        'points to this .npz file')
    p.add_argument('--form', default='plain',
                   choices=['plain', 'precision', 'batched', 'surrogate',
                            'gram', 'dense', 'general_b'],
                   help='plain: the headline workload (default); the others '
                        'measure the other SLQ forms and print their own '
                        'JSON lines')
    return p.parse_args()


def spmm_bytes(n, nnz, b=64, vbytes=8):
    """Algorithmic HBM bytes of one Lanczos SpMM launch: the CSR arrays once
    (values, int32 columns, int64 row pointers), the gathered operand block
    read once and the product block written once:
    nnz (s_v + 4) + 8 (n + 1) + 2 n b s_v. (The epilogue's read of
    v_{j-1}, n b s_v more after a block's first step, is added where the
    launch does it.)"""
    return nnz * (vbytes + 4) + (n + 1) * 8 + 2 * n * b * vbytes


def measured_traffic(n, nnz):
    """HBM bytes per SpMM launch measured by rocprofv3 --pmc for this exact
    workload (profiles/spmm_traffic.json: FETCH_SIZE and WRITE_SIZE of the
    SpMM kernels, each scaled by the factor the same counter reads for a
    plain copy kernel of known size), or None when none is recorded."""
    path = os.path.join(HERE, 'profiles', 'spmm_traffic.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get('n') != n or rec.get('nnz') != nnz:
        return None
    return rec['bytes_per_launch']


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


def cpu_baseline(A, m, threads, seed, per_thread=8):
    """Oracle restatement of imate's CPU SLQ recurrences on `threads` host
    threads, `per_thread` probes per thread (same matrix, same
    lanczos_degree): about 11 s at n = 2^20 on 16 threads."""
    from oracle import oracle
    op = oracle.Operator(A)
    eps = float(numpy.finfo(numpy.float64).eps)
    # one short warm-up to page in the library and matrix
    oracle.recurrences(op, seed, 1, 2, eps, 0, num_threads=1)
    t = time.perf_counter()
    a, b, sizes = oracle.recurrences(op, seed, per_thread * threads, m, eps,
                                     0, num_threads=threads)
    dt = time.perf_counter() - t
    steps = int(numpy.sum(sizes))
    return steps / dt, dt, steps


def cpu_path(A, m, threads, seed, probes=64):
    """This library's own CPU path (imate_amd with gpu=False, the host
    implementation slq_estimate_host) on the same matrix and cores: probes
    in blocks of 8 per pass over A. Not the baseline (that is the oracle
    above); reported beside it."""
    import imate_amd
    M = imate_amd.Matrix(A)
    _, info = imate_amd.logdet(M, method='slq', gpu=False, lanczos_degree=m,
                               min_num_samples=probes, max_num_samples=probes,
                               seed=seed, num_threads=threads,
                               return_info=True)
    dt = info['time']['alg_wall_time']
    return {'value': probes * m / dt, 'unit': 'matvec-steps/s',
            'cores': threads,
            'sample': '%d probes x %d Lanczos steps in %.1fs (sampling loop)'
                      % (probes, m, dt)}


def cpu_calibration():
    """The port's speed relative to imate's own CPU Lanczos loop (imate's
    sources compiled into oracle/_ref), measured on the same cores in the
    build container by oracle/calibrate_cpu_baseline.py."""
    path = os.path.join(HERE, 'oracle', 'cpu_calibration.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        c = json.load(f)
    return {'port_over_imate_cpu': c['port_over_reference'],
            'imate_cpu_steps_per_s': c['reference_steps_per_s'],
            'port_steps_per_s': c['port_steps_per_s'],
            'threads': c['threads'], 'host': c['host'],
            'workload': 'n=%d, nnz=%d, lanczos_degree %d, %d probes per thread'
                        % (c['n'], c['nnz'], c['lanczos_degree'],
                           c['probes_per_thread'])}


def precision_form(args, compact=False):
    """Precision form: n = 65536 correlation matrix with kernel scale 0.01
    (~247 non-zeros per row, cond ~1e4), lanczos_degree 60, full
    re-orthogonalisation, 64-probe blocks. FP32 SpMM / basis / Gram-Schmidt
    with FP64 recurrences and eigensolve ('mixed') against the plain FP64 GPU
    path; errors of logdet, traceinv and Schatten p=2 against exact values
    (dense FP64 Cholesky, tests/golden/exact_precision_form.json).
    compact: logdet only, full re-orthogonalisation, both precisions."""
    import imate_amd
    from imate_amd.sample_matrices import sparse_correlation_matrix
    n, m, s = 65536, 60, args.samples
    A = sparse_correlation_matrix(n, scale=0.01, seed=0)
    exact = {}
    path = os.path.join(HERE, 'tests', 'golden', 'exact_precision_form.json')
    if os.path.exists(path):
        with open(path) as f:
            exact = json.load(f)
    Aop = imate_amd.Matrix(A)
    Aop.get_linear_operator(gpu=True)
    rows = []
    for prec in ('native', 'mixed'):
        for ortho in ((-1,) if compact else (-1, 0)):
            kw = dict(method='slq', gpu=True, lanczos_degree=m,
                      min_num_samples=s, max_num_samples=s, seed=args.seed,
                      orthogonalize=ortho, precision=prec, return_info=True)
            imate_amd.logdet(Aop, **kw)   # warm-up (workspace, FP32 copies)
            t0 = time.perf_counter()
            ld, i1 = imate_amd.logdet(Aop, **kw)
            el1 = time.perf_counter() - t0
            res = {'precision': prec, 'orthogonalize': ortho,
                   'logdet_matvec_steps_per_s': s * m / el1,
                   'logdet': ld,
                   'logdet_mc_err': float(i1['error']['absolute_error'])}
            if exact:
                res['logdet_rel_err'] = abs(ld - exact['logdet']) / abs(exact['logdet'])
            if not compact:
                t0 = time.perf_counter()
                ti, i2 = imate_amd.traceinv(Aop, **kw)
                t2, i3 = imate_amd.trace(Aop, p=2, **kw)
                el = time.perf_counter() - t0 + el1
                sch = (t2 / n) ** 0.5
                res.update({'seconds_3_calls': el,
                            'matvec_steps_per_s': 3 * s * m / el,
                            'traceinv': ti,
                            'traceinv_mc_err': float(i2['error']['absolute_error']),
                            'schatten2': sch})
                if exact:
                    res['traceinv_rel_err'] = abs(ti - exact['traceinv']) / abs(exact['traceinv'])
                    res['schatten2_rel_err'] = abs(sch - exact['schatten2']) / exact['schatten2']
            log('precision form: %s' % json.dumps(res))
            rows.append(res)
    out = {'form': 'precision', 'n': n, 'nnz': int(A.nnz),
           'lanczos_degree': m, 'samples': s, 'runs': rows,
           'reorthogonalization': {
               'mfma_utilization': 0.0,
               'why': 'every probe has its own Krylov basis, so h[k][p] = '
                      'sum_i V[i][k][p] r[i][p] is a batched GEMV with no '
                      'operand shared across probes: 2 flops per basis '
                      'value read (0.5 flop/B FP32, 0.25 flop/B FP64), '
                      'bound by HBM; the Gram-Schmidt kernels run on the '
                      'vector ALUs at 3.8-5.5 TB/s'}}
    out['reorthogonalization']['block_krylov'] = block_krylov_leg(
        Aop, s, m, args.seed, exact, ('mixed',) if compact else ('mixed', 'native'))
    if not compact:
        out['exact'] = exact
    else:
        nat, mix = rows[0], rows[1]
        out['mixed_over_native_speed'] = (mix['logdet_matvec_steps_per_s'] /
                                          nat['logdet_matvec_steps_per_s'])
    return out


def block_krylov_leg(Aop, s, m, seed, exact, precisions):
    """The block Lanczos form (krylov='block') on the precision-form matrix:
    the 64 probes of a block share one basis, so its full
    re-orthogonalisation is a GEMM against the stored basis, run on the
    matrix cores (bk_tn_kernel: Q^T W; bk_nn_kernel: W -= Q H). Reports the
    call time (unprofiled run), the logdet error, and the GEMM kernels'
    device time and flops from HIP events around every launch (profiled
    run): achieved TFLOP/s and their fraction of the MI355X dense matrix
    peak of the GEMM's type (FP32 157.3, FP64 78.6 TFLOP/s)."""
    import ctypes
    import imate_amd
    from imate_amd import _lib
    lib = _lib.load()
    h = Aop.get_linear_operator(gpu=True).handle
    rows = []
    for prec in precisions:
        kw = dict(method='slq', gpu=True, krylov='block', lanczos_degree=m,
                  min_num_samples=s, max_num_samples=s, seed=seed,
                  precision=prec, return_info=True)
        imate_amd.logdet(Aop, **kw)     # warm-up: workspace, rocSOLVER load
        t0 = time.perf_counter()
        ld, info = imate_amd.logdet(Aop, **kw)
        el = time.perf_counter() - t0
        _lib.check(lib.slq_profile_enable(h, 1), 'slq_profile_enable')
        imate_amd.logdet(Aop, **kw)
        ms = ctypes.c_double(0.0)
        nl = ctypes.c_int64(0)
        fl = ctypes.c_double(0.0)
        _lib.check(lib.slq_profile_gemm(h, ctypes.byref(ms), ctypes.byref(nl),
                                        ctypes.byref(fl)), 'slq_profile_gemm')
        _lib.check(lib.slq_profile_enable(h, 0), 'slq_profile_enable')
        peak = 157.3 if prec == 'mixed' else 78.6
        tflops = fl.value / (ms.value * 1e-3) / 1e12 if ms.value > 0 else None
        r = {'precision': prec, 'gemm_dtype': 'f32' if prec == 'mixed' else 'f64',
             'seconds': el, 'logdet_matvec_steps_per_s': s * m / el,
             'logdet': ld, 'logdet_mc_err': float(info['error']['absolute_error']),
             'gemm_ms': ms.value, 'gemm_launches': nl.value,
             'gemm_flops': fl.value, 'gemm_tflops': tflops,
             'mfma_peak_tflops': peak,
             'gemm_mfma_fraction': tflops / peak if tflops else None}
        if exact:
            r['logdet_rel_err'] = abs(ld - exact['logdet']) / abs(exact['logdet'])
        log('block Krylov form: %s' % json.dumps(r))
        rows.append(r)
    return {'lanczos_degree': m, 'block': 64,
            'tridiagonal_size': 64 * m,
            'note': 'opt-in form (its samples differ from imate\'s per-probe '
                    'Lanczos); the block tridiagonal eigensolves (64 m x 64 m, '
                    'one batched rocSOLVER dsyevd call per group of blocks) '
                    'take the largest part of the call',
            'runs': rows}


def batched_form(args, compact=False):
    """Batched form. (i) A + tI with 256 log-spaced t in [1e-4, 1e2] on an
    n = 65536 correlation matrix (~50 non-zeros per row, FP64): one Lanczos
    run per probe with Ritz shifts for every t, against a measured loop of
    one SLQ run per t (all 256). (ii) 32 correlation matrices with sizes
    log-uniform in [1e3, 2e5], ~30 non-zeros per row, FP32: one hipGraph
    launch over all problems (with autotuned launch settings) against a
    per-problem launch loop of the same kernels and against plain FP64 GPU
    SLQ (one logdet call per matrix, the same probes: the FP32 values'
    distance from the FP64 ones is reported in units of the Monte-Carlo
    error, i.e. the accuracy is matched when it is far below 1). (iii, not
    compact) A + tB with a general B, 8 t, n = 262144: the tiled
    union-pattern SpMM against the two-stream CSR SpMM (general_b_leg)."""
    import imate_amd
    from imate_amd.batch import slq_batch
    from imate_amd.sample_matrices import sparse_correlation_matrix
    m, s = 30, 64
    out = {'form': 'batched', 'lanczos_degree': m, 'samples': s}
    # (i) many t of A + tI
    n = 65536
    A = sparse_correlation_matrix(n, nnz_per_row=50, seed=0)
    ts = numpy.logspace(-4, 2, 256)
    Aop = imate_amd.AffineMatrixFunction(A)
    kw = dict(method='slq', gpu=True, lanczos_degree=m, min_num_samples=s,
              max_num_samples=s, seed=args.seed, return_info=True)
    imate_amd.logdet(Aop, parameters=ts, **kw)
    t0 = time.perf_counter()
    vals, info = imate_amd.logdet(Aop, parameters=ts, **kw)
    t_shift = time.perf_counter() - t0
    imate_amd.logdet(Aop, parameters=float(ts[0]), **kw)
    t0 = time.perf_counter()
    direct = []
    for t in ts:
        v, _ = imate_amd.logdet(Aop, parameters=float(t), **kw)
        direct.append(v)
    t_loop = time.perf_counter() - t0
    direct = numpy.array(direct)
    rel = numpy.max(numpy.abs(direct - vals) / numpy.abs(direct))
    out['affine_t256'] = {
        'n': n, 'nnz': int(A.nnz), 'num_t': len(ts),
        'shift_seconds': t_shift, 'per_t_loop_seconds': t_loop,
        'speedup': t_loop / t_shift,
        't_evaluations_per_s': len(ts) * s / t_shift,
        'max_rel_diff_shift_vs_direct': float(rel)}
    log('affine: %s' % json.dumps(out['affine_t256']))
    if not compact:
        # (iii) many t of A + tB, general B: one recurrence per t on the
        # tiles of the union pattern, against the two-stream CSR SpMM
        out['affine_general_b'] = general_b_leg(args.seed, m, s)
    # (ii) 32 uneven matrices
    rng = numpy.random.default_rng(args.seed)
    sizes = numpy.round(numpy.exp(rng.uniform(numpy.log(1e3), numpy.log(2e5),
                                              32))).astype(int)
    mats = [sparse_correlation_matrix(int(k), nnz_per_row=30, seed=i,
                                      dtype='float32')
            for i, k in enumerate(sizes)]
    nnz = sum(M.nnz for M in mats)
    tg, ig = slq_batch(mats, lanczos_degree=m, num_samples=s, seed=args.seed,
                       use_graph=True)
    tg, ig = slq_batch([ig['operators'][i] for i in range(32)],
                       lanczos_degree=m, num_samples=s, seed=args.seed,
                       use_graph=True)
    tl, il = slq_batch(ig['operators'], lanczos_degree=m, num_samples=s,
                       seed=args.seed, use_graph=False, autotune=False)
    # plain FP64 GPU SLQ, one call per matrix (uploaded beforehand)
    m64 = [imate_amd.Matrix(M.astype(numpy.float64)) for M in mats]
    for M in m64:
        M.get_linear_operator(gpu=True)
    kw64 = dict(method='slq', gpu=True, lanczos_degree=m, min_num_samples=s,
                max_num_samples=s, seed=args.seed, return_info=True)
    imate_amd.logdet(m64[0], **kw64)
    t0 = time.perf_counter()
    r64 = [imate_amd.logdet(M, **kw64) for M in m64]
    t64 = time.perf_counter() - t0
    v64 = numpy.array([r[0] for r in r64])
    e64 = numpy.array([float(r[1]['error']['absolute_error']) for r in r64])
    steps = 32 * s * m
    out['matrices32'] = {
        'sizes': [int(k) for k in sizes], 'total_rows': int(sizes.sum()),
        'total_nnz': int(nnz), 'dtype': 'f32',
        'graph_seconds': ig['run_time'], 'graph_setup_seconds': ig['setup_time'],
        'loop_seconds': il['run_time'],
        'speedup': il['run_time'] / ig['run_time'],
        'graph_matvec_steps_per_s': steps / ig['run_time'],
        'loop_matvec_steps_per_s': steps / il['run_time'],
        'plain_fp64_gpu_seconds': t64,
        'speedup_vs_plain_fp64_gpu': t64 / ig['run_time'],
        'max_abs_diff_fp32_vs_fp64_over_mc_error':
            float(numpy.max(numpy.abs(tg - v64) / e64)),
        'samples_identical_graph_vs_loop':
            bool(numpy.array_equal(ig['samples'], il['samples'])),
        'autotuned_grids': ig['grids']}
    log('matrices: %s' % json.dumps({k: v for k, v in out['matrices32'].items()
                                     if k not in ('sizes', 'autotuned_grids')}))
    if compact:
        out['matrices32'] = {k: v for k, v in out['matrices32'].items()
                             if k not in ('sizes', 'autotuned_grids')}
    return out


def general_b_leg(seed, m, s, n=262144, num_t=8):
    """A + tB for num_t log-spaced t in [1e-3, 10]: A an n x n correlation
    matrix (~50 non-zeros per row), B SPD on a wider pattern of the same
    points (~80 per row) plus about one random symmetric pair per row."""
    import scipy.sparse
    import imate_amd
    from imate_amd.sample_matrices import sparse_correlation_matrix
    A = sparse_correlation_matrix(n, nnz_per_row=50, seed=0)
    C = sparse_correlation_matrix(n, nnz_per_row=80, seed=0)
    rng = numpy.random.default_rng(1)
    k = n // 2
    R = scipy.sparse.csr_matrix((rng.random(k), (rng.integers(0, n, k),
                                                 rng.integers(0, n, k))),
                                shape=(n, n))
    S = 0.1 * (R + R.T)
    B = (C + S + scipy.sparse.diags(
        numpy.asarray(abs(S).sum(axis=1)).ravel())).tocsr()
    ts = numpy.logspace(-3, 1, num_t)
    Aop = imate_amd.AffineMatrixFunction(A, B)
    kw = dict(parameters=ts, method='slq', gpu=True, lanczos_degree=m,
              min_num_samples=s, max_num_samples=s, seed=seed)
    res = {'n': n, 'nnz_a': int(A.nnz), 'nnz_b': int(B.nnz), 'num_t': num_t}
    vals = {}
    # union tiles with the t side by side (default), one t after the other,
    # and the two-stream CSR SpMM
    for path, pair, conc in (('pair_tiles', None, None),
                             ('pair_tiles_sequential_t', None, '0'),
                             ('csr', '0', '0')):
        for var, val in (('SLQ_PAIR_TILES', pair), ('SLQ_CONCURRENT_T', conc)):
            if val is None:
                os.environ.pop(var, None)
            else:
                os.environ[var] = val
        imate_amd.logdet(Aop, **kw)            # builds, warms up
        t0 = time.perf_counter()
        vals[path] = numpy.asarray(imate_amd.logdet(Aop, **kw))
        sec = time.perf_counter() - t0
        res[path] = {'seconds': sec,
                     'matvec_steps_per_s': num_t * s * m / sec}
    os.environ.pop('SLQ_PAIR_TILES', None)
    os.environ.pop('SLQ_CONCURRENT_T', None)
    res['speedup'] = res['csr']['seconds'] / res['pair_tiles']['seconds']
    res['speedup_t_side_by_side'] = (res['pair_tiles_sequential_t']['seconds'] /
                                     res['pair_tiles']['seconds'])
    res['t_side_by_side_bitwise_equal'] = bool(numpy.array_equal(
        vals['pair_tiles'], vals['pair_tiles_sequential_t']))
    res['max_rel_diff'] = float(numpy.max(numpy.abs(vals['pair_tiles'] -
                                                    vals['csr']) /
                                          numpy.abs(vals['csr'])))
    log('affine general B: %s' % json.dumps(res))
    return res


def interpolator_baseline(fam, ls, ts, t_range, m, samples, slq, ci, pred,
                          hs, ht, n):
    """InterpolateLogdet (kind 'imbf', ortho2 basis) per held-out scale over
    the held-out t values: its errors against the SLQ judge and the dense
    Cholesky values, beside the learned surrogate's, and its cost per scale
    (upload of K(l), one SLQ call for the 6 points, the fit)."""
    from imate_amd import InterpolateLogdet
    ti = numpy.geomspace(t_range[0], t_range[1], 6)
    opts = dict(method='slq', gpu=True, lanczos_degree=m,
                min_num_samples=samples, max_num_samples=samples, seed=707)
    q = ts.shape[1]
    ipred = numpy.empty(len(ht))
    build = fit = 0.0
    for i, scale in enumerate(ls):
        t0 = time.perf_counter()
        K = fam.matrix(scale)
        t1 = time.perf_counter()
        L = InterpolateLogdet(K, kind='imbf', ti=ti, options=opts)
        ipred[i * q:(i + 1) * q] = L.interpolate(ts[i])
        t2 = time.perf_counter()
        build += t1 - t0
        fit += t2 - t1
    res = {'kind': 'imbf (ortho2)', 'points': len(ti),
           'matrix_build_seconds_per_scale': build / len(ls),
           'upload_slq_fit_seconds_per_scale': fit / len(ls),
           'inside_ci': float(numpy.mean(numpy.abs(ipred - slq) <= ci)),
           'median_abs_err_per_n_vs_slq':
               float(numpy.median(numpy.abs(ipred - slq)) / n),
           'max_abs_err_per_n_vs_slq':
               float(numpy.max(numpy.abs(ipred - slq)) / n),
           'median_abs_err_per_n_surrogate_vs_slq':
               float(numpy.median(numpy.abs(pred - slq)) / n)}
    path = os.path.join(HERE, 'tests', 'golden', 'surrogate_cholesky.json')
    if os.path.exists(path):
        with open(path) as f:
            ref = json.load(f)
        ei, es = [], []
        for e in ref['entries']:
            i = e['scale_index'] * q + e['t_index']
            assert abs(e['scale'] - hs[i]) < 1e-12 * hs[i]
            ei.append(abs(ipred[i] - e['logdet']))
            es.append(abs(pred[i] - e['logdet']))
        res['vs_cholesky'] = {
            'count': len(ei),
            'median_abs_err_per_n_interpolator': float(numpy.median(ei) / n),
            'median_abs_err_per_n_surrogate': float(numpy.median(es) / n),
            'max_abs_err_per_n_interpolator': float(numpy.max(ei) / n),
            'max_abs_err_per_n_surrogate': float(numpy.max(es) / n)}
    log('interpolator baseline: %s' % json.dumps(res))
    return res


def surrogate_form(args, compact=False):
    """Learned form. Family: n = 65536 correlation matrices K(l) over the
    length scale l in [0.01, 0.1] with the support radius fixed at
    0.01 ln(1/0.03) (one pattern, ~247 non-zeros per row), queries
    logdet(K(l) + tI), t in [1e-4, 1e2], both log-uniform. The surrogate is
    trained on 40 x 50 = 2000 GPU-SLQ estimates (lanczos_degree 60, 128
    probes; scales and t stratified over the box), calibrated on 10 x 40
    random validation estimates, judged on 10 x 20
    held-out queries against SLQ (fallback where it leaves the 95% interval)
    and against dense Cholesky values recorded for 12 of them, then swept
    over 10,000 random queries (trust mode: SLQ only where the surrogate is
    not trusted) and compared with plain FP64 GPU SLQ per query."""
    from imate_amd.sample_matrices import CorrelationFamily
    from imate_amd.surrogate import LogdetSurrogate, sample_queries
    n = 65536
    support = 0.01 * numpy.log(1.0 / 0.03)
    t0 = time.perf_counter()
    fam = CorrelationFamily(n, support, seed=0)
    log('family n=%d nnz=%d built in %.1fs' % (n, fam.nnz,
                                               time.perf_counter() - t0))
    m = 60
    sur = LogdetSurrogate(fam, lanczos_degree=m, num_samples=128, seed=0)
    out = {'form': 'surrogate', 'n': n, 'nnz': int(fam.nnz),
           'scale_range': sur.scale_range, 't_range': sur.t_range,
           'lanczos_degree': m, 'samples': 128, 'network': '2-128-128-1 SiLU',
           'train_dtype': 'f32', 'inference_dtype': 'bf16 (MFMA)'}
    sur.solve([0.03], [1.0])                      # first-use set-up
    train = sur.generate(40, 50, seed=101)
    log('training data: 2000 SLQ estimates in %.1fs' % train['seconds'])
    sec = sur.train(train)
    out['train'] = {'points': 2000, 'slq_seconds': train['seconds'],
                    'train_seconds': sec, 'adam_steps': int(len(sur.losses)),
                    'loss_first': float(sur.losses[0]),
                    'loss_last': float(sur.losses[-1])}
    log('train: %s' % json.dumps(out['train']))
    val = sur.generate(10, 40, seed=202, stratified=False)
    rho = sur.calibrate(val)
    out['validation'] = {'points': len(rho),
                         'inside_ci': float(numpy.mean(numpy.abs(rho) <= 1)),
                         'rms_rho': float(numpy.sqrt(numpy.mean(rho ** 2)))}
    # held-out queries, judged by SLQ
    ls, ts = sample_queries(10, 20, sur.scale_range, sur.t_range, seed=303)
    hs, ht = numpy.repeat(ls, 20), ts.ravel()
    y, info = sur.query(hs, ht, judge='slq', seed=303000)
    slq, ci, pred = info['slq'], info['ci'], info['surrogate']
    # errors per matrix row (|error| / n): logdet(K + tI) changes sign inside
    # the box, so errors relative to the value are meaningless near its zero
    held = {'points': len(ht),
            'surrogate_inside_ci': float(numpy.mean(numpy.abs(pred - slq) <= ci)),
            'fallback_fraction': float(info['fallback'].mean()),
            'median_abs_err_per_n_surrogate_vs_slq':
                float(numpy.median(numpy.abs(pred - slq)) / n),
            'max_abs_err_per_n_surrogate_vs_slq':
                float(numpy.max(numpy.abs(pred - slq)) / n),
            'median_err_over_ci': float(numpy.median(numpy.abs(pred - slq) / ci)),
            'max_err_over_ci': float(numpy.max(numpy.abs(pred - slq) / ci)),
            'median_ci_per_n': float(numpy.median(ci) / n)}
    # how the trust rule (used without SLQ in the sweep) fares on these
    trusted = sur.trusted(hs, ht)
    inside = numpy.abs(pred - slq) <= ci
    held['trusted_fraction'] = float(trusted.mean())
    held['inside_ci_among_trusted'] = float(inside[trusted].mean()) \
        if trusted.any() else None
    held['inside_ci_among_untrusted'] = float(inside[~trusted].mean()) \
        if (~trusted).any() else None
    path = os.path.join(HERE, 'tests', 'golden', 'surrogate_cholesky.json')
    if os.path.exists(path):
        with open(path) as f:
            ref = json.load(f)
        rows = []
        for e in ref['entries']:
            i = e['scale_index'] * 20 + e['t_index']
            ex = e['logdet']
            assert abs(e['scale'] - hs[i]) < 1e-12 * hs[i] and \
                abs(e['t'] - ht[i]) < 1e-12 * ht[i]
            rows.append({'surrogate': abs(pred[i] - ex), 'slq': abs(slq[i] - ex),
                         'returned': abs(y[i] - ex), 'ci': ci[i],
                         'rel_surrogate': abs(pred[i] - ex) / abs(ex),
                         'rel_slq': abs(slq[i] - ex) / abs(ex)})
        col = {k: numpy.array([r[k] for r in rows]) for k in rows[0]}
        held['vs_cholesky'] = {
            'count': len(rows),
            'median_abs_err_per_n_surrogate': float(numpy.median(col['surrogate']) / n),
            'median_abs_err_per_n_slq': float(numpy.median(col['slq']) / n),
            'max_abs_err_per_n_surrogate': float(col['surrogate'].max() / n),
            'max_abs_err_per_n_slq': float(col['slq'].max() / n),
            'max_abs_err_per_n_returned': float(col['returned'].max() / n),
            'max_rel_err_surrogate': float(col['rel_surrogate'].max()),
            'max_rel_err_slq': float(col['rel_slq'].max()),
            'surrogate_within_ci_of_exact':
                float(numpy.mean(col['surrogate'] <= col['ci'])),
            'slq_within_ci_of_exact': float(numpy.mean(col['slq'] <= col['ci']))}
    # the reference's own surrogate over t, one interpolator per scale
    # (imbf, ortho2 basis, 6 points over the t range, estimated in one call)
    held['interpolator'] = interpolator_baseline(
        fam, ls, ts, sur.t_range, m, 128, slq, ci, pred, hs, ht, n)
    out['held_out'] = held
    log('held-out: %s' % json.dumps(held))
    if args.dump:
        numpy.savez(args.dump, held_scale=hs, held_t=ht, held_pred=pred,
                    held_slq=slq, held_ci=ci, val_scale=val['scales'],
                    val_t=val['t'], val_y=val['y'], val_ci=val['ci'],
                    val_rho=rho, losses=sur.losses)
    # plain FP64 GPU SLQ per query: new l (values refreshed in place) + SLQ
    rng = numpy.random.default_rng(404)
    Q = 10000
    qs = numpy.exp(rng.uniform(numpy.log(0.01), numpy.log(0.1), Q))
    qt = numpy.exp(rng.uniform(numpy.log(1e-4), numpy.log(1e2), Q))
    k = 16
    _, _, plain_sec = sur.solve(qs[:k], qt[:k], seed=505)
    plain_per_query = plain_sec / k
    plain_split = dict(sur.last_solve)
    # surrogate sweep: BF16 inference of all queries, trust check, SLQ for
    # the untrusted ones (executed up to 64, the rest charged at the plain
    # per-query time)
    sur.predict(qs[:64], qt[:64])                 # warm-up
    t0 = time.perf_counter()
    pred_all, ksec = sur.predict(qs, qt, bf16=True)
    t_pred = time.perf_counter() - t0
    t0 = time.perf_counter()
    ok = sur.trusted(qs, qt)
    t_trust = time.perf_counter() - t0
    fb = numpy.nonzero(~ok)[0]
    run = fb[:64]
    fb_sec = 0.0
    if len(run):
        _, _, fb_sec = sur.solve(qs[run], qt[run], seed=606)
    fb_total = fb_sec + (len(fb) - len(run)) * plain_per_query
    p32, _ = sur.predict(qs, qt, bf16=False)
    total = t_pred + t_trust + fb_total
    out['sweep'] = {
        'queries': Q, 'inference_kernel_seconds': ksec,
        'inference_call_seconds': t_pred,
        'inference_queries_per_s_kernel': Q / ksec,
        'trust_seconds': t_trust, 'fallback_queries': int(len(fb)),
        'fallback_fraction': float(len(fb) / Q),
        'fallback_seconds_run': fb_sec, 'fallback_run': int(len(run)),
        'fallback_seconds_total': fb_total,
        'surrogate_queries_per_s': Q / total,
        'plain_slq_seconds_per_query': plain_per_query,
        'plain_slq_queries_per_s': 1.0 / plain_per_query,
        'plain_slq_values_seconds_per_query':
            plain_split['values_seconds'] / k,
        'plain_slq_gpu_seconds_per_query': plain_split['slq_seconds'] / k,
        'gain_vs_plain_slq': plain_per_query * Q / total,
        'bf16_vs_f32_max_abs_diff_per_n':
            float(numpy.max(numpy.abs(pred_all - p32)) / n)}
    log('sweep: %s' % json.dumps(out['sweep']))
    if compact:
        h = out['held_out']
        sw = out['sweep']
        keep = {'n': n, 'lanczos_degree': m, 'samples': 128,
                'train': out['train'],
                'held_out': {k: h[k] for k in (
                    'points', 'surrogate_inside_ci', 'fallback_fraction',
                    'median_abs_err_per_n_surrogate_vs_slq') if k in h},
                'sweep': {k: sw[k] for k in (
                    'queries', 'fallback_fraction', 'surrogate_queries_per_s',
                    'plain_slq_queries_per_s', 'gain_vs_plain_slq',
                    'inference_queries_per_s_kernel')}}
        if 'vs_cholesky' in h:
            keep['held_out']['vs_cholesky'] = {
                k: h['vs_cholesky'][k] for k in (
                    'count', 'median_abs_err_per_n_surrogate',
                    'median_abs_err_per_n_slq')}
        if 'interpolator' in h:
            keep['held_out']['interpolator_inside_ci'] = h['interpolator']['inside_ci']
        return keep
    return out


def gram_form(args):
    """Golub-Kahn (gram=True) at scale: logdet(A^T A) for a seeded sparse
    Gaussian A of 2^20 x 2^18 at 1 % density (2621 entries per row, 2.75e9
    non-zeros, FP64), generated on the device; lanczos_degree 30, 256 probes.
    The transpose (CSR of A^T) is built on the device in column chunks during
    an untimed warm-up call. Checked against the Marchenko-Pastur limit:
    for A with i.i.d. entries of variance s2 = k / ncols, E logdet(A^T A) =
    p log s2 + sum_{i<p} (psi((N - i) / 2) + log 2) (Wishart; the sparse
    matrix matches it up to O(1 / column degree))."""
    import imate_amd
    from scipy.special import digamma
    from imate_amd.sample_matrices import sparse_gaussian_device
    N, p, density = 2**20, 2**18, 0.01
    m, s = args.lanczos_degree, args.samples
    t0 = time.perf_counter()
    D = sparse_gaussian_device(N, p, density, seed=args.seed)
    t_gen = time.perf_counter() - t0
    log('generated %d x %d, nnz %d in %.2fs' % (N, p, D.get_nnz(), t_gen))
    kw = dict(gram=True, method='slq', gpu=True, seed=args.seed,
              return_info=True)
    t0 = time.perf_counter()
    imate_amd.logdet(D, lanczos_degree=2, min_num_samples=1,
                     max_num_samples=1, **kw)
    t_setup = time.perf_counter() - t0
    log('warm-up (transpose build) %.1fs' % t_setup)
    t0 = time.perf_counter()
    val, info = imate_amd.logdet(D, lanczos_degree=m, min_num_samples=s,
                                 max_num_samples=s, **kw)
    elapsed = time.perf_counter() - t0
    k = int(round(density * p))
    s2 = k / p
    i = numpy.arange(p, dtype=numpy.float64)
    mp = p * numpy.log(s2) + float(numpy.sum(digamma((N - i) / 2.0) + numpy.log(2.0)))
    steps = s * m
    bytes_step = 2 * D.get_nnz() * 12 + 2 * (N + p) * 64 * 8
    block_steps = (s // 64) * m
    out = {'form': 'gram', 'shape': [N, p], 'nnz': int(D.get_nnz()),
           'density': density, 'lanczos_degree': m, 'samples': s,
           'generate_seconds': t_gen, 'setup_seconds_incl_transpose': t_setup,
           'seconds': elapsed, 'gk_steps_per_s': steps / elapsed,
           'matvecs_per_s': 2 * steps / elapsed,
           'hbm_gb_per_block_step_algorithmic': bytes_step / 1e9,
           'achieved_GBps_algorithmic': bytes_step * block_steps / elapsed / 1e9,
           'logdet': float(val),
           'mc_error_95': float(info['error']['absolute_error']),
           'marchenko_pastur_wishart_mean': mp,
           'rel_diff_vs_wishart_mean': abs(float(val) - mp) / abs(mp)}
    # the precision form: FP32 copies of A and A^T and FP32 vectors, so each
    # gathered operand row is 256 B instead of 512 (FP64 recurrences)
    imate_amd.logdet(D, lanczos_degree=2, min_num_samples=1, max_num_samples=1,
                     precision='mixed', **kw)
    t0 = time.perf_counter()
    valm, infom = imate_amd.logdet(D, lanczos_degree=m, min_num_samples=s,
                                   max_num_samples=s, precision='mixed', **kw)
    el_m = time.perf_counter() - t0
    out['mixed'] = {'seconds': el_m, 'gk_steps_per_s': steps / el_m,
                    'speedup_vs_fp64': elapsed / el_m, 'logdet': float(valm),
                    'mc_error_95': float(infom['error']['absolute_error']),
                    'rel_diff_vs_fp64': abs(float(valm) - float(val)) / abs(float(val)),
                    'rel_diff_vs_wishart_mean': abs(float(valm) - mp) / abs(mp)}
    log('gram: %s' % json.dumps(out))
    return out


def dense_form(args):
    """Dense operator (imate's cuDenseMatrix): logdet of an SPD dense matrix
    A = (R + R^T)/2 + 1.5 sqrt(n) I, R uniform in [-1, 1], FP64, n = 32768
    (8.6 GB), lanczos_degree 30, 64 probes. The SpMM is a (n x n)(n x 64)
    product on the matrix cores: 2 n^2 64 flops and n^2 8 bytes of matrix per
    launch, timed with HIP events; against the same matrix through the CSR
    path (dense converted to CSR and 16-row tiles) at n = 8192."""
    import imate_amd
    from imate_amd import _lib
    lib = _lib.load()
    out = {'form': 'dense', 'lanczos_degree': args.lanczos_degree,
           'samples': 64}

    def spd(n, seed):
        rng = numpy.random.default_rng(seed)
        A = rng.uniform(-1.0, 1.0, (n, n))
        A += A.T
        A *= 0.5
        A[numpy.diag_indices(n)] += 1.5 * numpy.sqrt(n)
        return A

    def run(M, n, label):
        kw = dict(method='slq', gpu=True, lanczos_degree=args.lanczos_degree,
                  min_num_samples=64, max_num_samples=64, seed=args.seed,
                  return_info=True)
        imate_amd.logdet(M, **kw)                        # warm-up
        h = M.gpu_Aop.handle
        _lib.check(lib.slq_profile_enable(h, 1), 'slq_profile_enable')
        t0 = time.perf_counter()
        val, info = imate_amd.logdet(M, **kw)
        el = time.perf_counter() - t0
        ms = ctypes.c_double(0)
        cnt = ctypes.c_int64(0)
        ums = ctypes.c_double(0)
        ucnt = ctypes.c_int64(0)
        _lib.check(lib.slq_profile_read(h, ctypes.byref(ms), ctypes.byref(cnt),
                                        ctypes.byref(ums), ctypes.byref(ucnt)),
                   'slq_profile_read')
        lib.slq_profile_enable(h, 0)
        spmm = ms.value / max(cnt.value, 1)
        flops = 2.0 * n * n * 64
        res = {'path': label, 'n': n, 'seconds': el,
               'matvec_steps_per_s': 64 * args.lanczos_degree / el,
               'spmm_ms': spmm, 'spmm_tflops': flops / spmm / 1e9,
               'spmm_frac_fp64_mfma_peak': flops / spmm / 1e9 / 78.6,
               'spmm_matrix_GBps': 8.0 * n * n / spmm / 1e6,
               'logdet': float(val),
               'mc_error_95': float(info['error']['absolute_error'])}
        log('dense: %s' % json.dumps(res))
        return res

    n = 32768
    t0 = time.perf_counter()
    A = spd(n, args.seed)
    log('dense n=%d built in %.1fs' % (n, time.perf_counter() - t0))
    M = imate_amd.Matrix(A)
    M.get_linear_operator(gpu=True)
    out['dense_32768'] = run(M, n, 'dense')
    del M, A
    n = 8192
    A = spd(n, args.seed)
    sign, ld_exact = numpy.linalg.slogdet(A)
    out['dense_8192'] = run(imate_amd.Matrix(A), n, 'dense')
    t0 = time.perf_counter()
    Mc = imate_amd.Matrix(scipy.sparse.csr_matrix(A))
    Mc.get_linear_operator(gpu=True)
    out['csr_tiles_8192_upload_seconds'] = time.perf_counter() - t0
    out['csr_tiles_8192'] = run(Mc, n, 'csr (16-row tiles)')
    out['exact_logdet_8192'] = float(ld_exact)
    out['rel_err_8192'] = abs(out['dense_8192']['logdet'] - ld_exact) / abs(ld_exact)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this package has no multi-GPU '
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

    forms = {'precision': precision_form, 'batched': batched_form,
             'surrogate': surrogate_form, 'gram': gram_form,
             'dense': dense_form}
    if args.form in forms:
        print(json.dumps(forms[args.form](args)), flush=True)
        return
    if args.form == 'general_b':
        print(json.dumps({'form': 'general_b', 'lanczos_degree': 30,
                          'samples': 64,
                          'affine_general_b': general_b_leg(args.seed, 30, 64),
                          'affine_general_b_small': general_b_leg(
                              args.seed, 30, 64, n=16384, num_t=16)}),
              flush=True)
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
    done = ctypes.c_int64(0)
    _lib.check(lib.slq_profile_steps(handle, ctypes.byref(done)),
               'slq_profile_steps')
    lib.slq_profile_enable(handle, 0)

    # completed Lanczos steps of the processed probes (a probe whose
    # recurrence breaks down early counts only the steps it ran)
    value = done.value / elapsed
    ms_per_step = elapsed / args.steps * 1e3
    avg_spmm = spmm_ms.value / max(spmm_n.value, 1)
    avg_upd = upd_ms.value / max(upd_n.value, 1)
    # with the beta term in the SpMM epilogue (the library default) every
    # SpMM after a block's first also reads v_{j-1}'s block once and the
    # update reads two blocks and writes one; otherwise the update reads
    # three blocks after the first step
    beta_in_spmm = os.environ.get('SLQ_SPMM_BETA', '1').strip() != '0'
    block = n * 64 * 8
    bspmm = spmm_bytes(n, A.nnz) + (block * (m - 1) / m if beta_in_spmm else 0)
    achieved = bspmm / (avg_spmm * 1e-3) / 1e9
    upd_bytes = block * (3 if beta_in_spmm else 3 + (m - 1) / m)
    log('timed %d steps: %.1f ms/step, %.0f matvec-steps/s; SpMM %.3f ms '
        '(%d launches), update %.3f ms (%d)' % (args.steps, ms_per_step, value,
                                                avg_spmm, spmm_n.value,
                                                avg_upd, upd_n.value))

    # the achievable streaming rate on this device: a plain copy kernel
    # (1 GiB read + 1 GiB written per launch), for frac_copy_peak
    copy_ms = ctypes.c_double(0)
    copy_gbps = None
    if lib.slq_copy_bandwidth(1 << 30, 10, ctypes.byref(copy_ms)) == 0 and \
            copy_ms.value > 0:
        copy_gbps = 2 * (1 << 30) / (copy_ms.value * 1e-3) / 1e9
        log('copy kernel: %.0f GB/s' % copy_gbps)

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
               'sample': '%d probes x %d Lanczos steps (eight probes per '
                         'thread) on the same n=%d matrix: %d steps in %.1fs'
                         % (8 * threads, m, n, steps, dt)}
        cal = cpu_calibration()
        if cal is not None:
            cpu['calibration'] = cal
        log('cpu baseline: %.1f matvec-steps/s on %d threads' % (rate, threads))
        cpu['product_cpu_path'] = cpu_path(A, m, threads, args.seed)
        log('gpu=False path: %.1f matvec-steps/s on %d threads'
            % (cpu['product_cpu_path']['value'], threads))

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
        'data': 'synthetic (seeded sparse correlation matrix, points '
                'uniform in the unit square sorted along a Hilbert curve)',
        'config': {'workload': 'slq-%s-correlation-n%d-nnz%.0f-m%d-s%d-b64'
                               % (args.func, n, k, m, s),
                   'n': n, 'nnz': int(A.nnz), 'lanczos_degree': m,
                   'samples': s, 'block': 64, 'parallelism': 'single-gpu',
                   'completed_steps_per_call': done.value / args.steps},
        'roofline': {'bound': 'hbm', 'achieved': achieved,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS,
                     'traffic': measured_traffic(n, A.nnz),
                     'traffic_source': 'rocprofv3 FETCH_SIZE and WRITE_SIZE of '
                                       'this workload scaled by the copy '
                                       'kernel calibration, recorded in '
                                       'profiles/spmm_traffic.json (not '
                                       'measured in this run)',
                     'copy_peak_GBps': copy_gbps,
                     'frac_copy_peak': (achieved / copy_gbps
                                        if copy_gbps else None),
                     'kernel': 'mfma_spmm_kernel<double> + tile sums (one '
                               'Lanczos SpMM launch; its epilogue also '
                               'subtracts beta_{j-1} v_{j-1})'
                               if beta_in_spmm else
                               'mfma_spmm_kernel<double> + tile sums (one '
                               'Lanczos SpMM launch)',
                     'bytes_per_launch': bspmm,
                     'avg_launch_ms': avg_spmm},
        'update_kernel': {'avg_launch_ms': avg_upd,
                          'GBps': upd_bytes / (avg_upd * 1e-3) / 1e9},
        'accuracy': accuracy,
        'cpu_baseline': cpu,
    }
    if not args.no_forms:
        # the other forms, compact, each against plain FP64 GPU SLQ
        t0 = time.perf_counter()
        forms = {}
        for name, fn in (('precision', precision_form),
                         ('batched', batched_form),
                         ('surrogate', surrogate_form)):
            t1 = time.perf_counter()
            forms[name] = fn(args, compact=True)
            forms[name]['leg_seconds'] = time.perf_counter() - t1
            log('%s leg done in %.1fs' % (name, forms[name]['leg_seconds']))
        forms['seconds'] = time.perf_counter() - t0
        out['forms'] = forms
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
