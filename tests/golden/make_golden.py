"""Writes tests/golden/ref_vectors.npz from the reference's own CPU sources.

Run in a tree where /root/reference exists, after ``make -C oracle ref``
(which compiles imate's C++ RNG, vector/matrix operations, operators,
Gram-Schmidt, Lanczos, Golub-Kahn, convergence tools and erf_inv from the
reference sources into oracle/_ref/libimate_ref.so). Only inputs and outputs
are stored; the fixtures let the oracle be checked where the reference is
absent.

    python tests/golden/make_golden.py
"""

import ctypes
import os
import sys

import numpy
import scipy.sparse

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from imate_amd.sample_matrices import (toeplitz,  # noqa: E402
                                       sparse_correlation_matrix)

P = ctypes.c_void_p


def ptr(a):
    return a.ctypes.data_as(P) if a is not None else None


def ref_recurrences(R, A, gram, seed, ns, m, tol, ortho, kind=0, t=0.0,
                    B=None, dtype=numpy.float64):
    A = scipy.sparse.csr_matrix(A)
    A.sort_indices()
    dat = numpy.ascontiguousarray(A.data, dtype=dtype)
    idx = A.indices.astype(numpy.int32)
    ip = A.indptr.astype(numpy.int32)
    if B is not None:
        B = scipy.sparse.csr_matrix(B)
        B.sort_indices()
        bd = numpy.ascontiguousarray(B.data, dtype=dtype)
        bi = B.indices.astype(numpy.int32)
        bp = B.indptr.astype(numpy.int32)
    else:
        bd = bi = bp = None
    a = numpy.zeros((ns, m), dtype)
    b = numpy.zeros((ns, m), dtype)
    sz = numpy.zeros(ns, numpy.int32)
    f = R.ref_slq_recurrences_f64 if dtype == numpy.float64 else \
        R.ref_slq_recurrences_f32
    real = ctypes.c_double if dtype == numpy.float64 else ctypes.c_float
    f(ptr(dat), ptr(idx), ptr(ip), A.shape[0], A.shape[1], kind,
      ctypes.c_double(t), ptr(bd), ptr(bi), ptr(bp), int(gram),
      ctypes.c_longlong(seed), ns, m, real(tol), ortho, ptr(a), ptr(b),
      ptr(sz), None)
    return a, b, sz


def main():
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit('oracle/_ref/libimate_ref.so missing: make -C oracle ref')
    out = {}

    # RNG: words of thread streams
    cases = [(0, 1, 0), (1, 1, 0), (12345, 1, 0), (2**31 + 5, 1, 0),
             (7, 3, 0), (7, 3, 2)]
    for c, (seed, nt, tid) in enumerate(cases):
        w = numpy.zeros(40, numpy.uint64)
        R.ref_stream_words(ctypes.c_longlong(seed), nt, tid,
                           ctypes.c_longlong(40), ptr(w))
        out['rng_%d_case' % c] = numpy.array([seed, nt, tid], numpy.int64)
        out['rng_%d_words' % c] = w

    # erf_inv
    R.ref_erf_inv.restype = ctypes.c_double
    xs = numpy.array([-0.999, -0.5, -0.1, 0.0, 0.05, 0.3, 0.69, 0.7, 0.71,
                      0.9, 0.95, 0.99, 0.999, 0.999999])
    out['erfinv_x'] = xs
    out['erfinv_y'] = numpy.array([R.ref_erf_inv(ctypes.c_double(x))
                                   for x in xs])

    # recurrences
    corr = sparse_correlation_matrix(300, nnz_per_row=20, seed=0)
    tg = toeplitz(2.0, 1.0, 50, gram=True)
    tb = toeplitz(2.0, 1.0, 50, gram=False)
    bdiag = scipy.sparse.diags(numpy.linspace(1.0, 2.0, 300)).tocsr()
    mats = {'corr': corr, 'toep_gram': tg, 'toep_bidiag': tb, 'bdiag': bdiag}
    for name, M in mats.items():
        M = scipy.sparse.csr_matrix(M)
        out['mat_%s_indptr' % name] = M.indptr.astype(numpy.int64)
        out['mat_%s_indices' % name] = M.indices.astype(numpy.int32)
        out['mat_%s_data' % name] = M.data.astype(numpy.float64)
        out['mat_%s_shape' % name] = numpy.array(M.shape, numpy.int64)
    rec = [
        # (matrix, B, gram, ortho, kind, t, dtype)
        ('corr', None, 0, 0, 0, 0.0, 'float64'),
        ('corr', None, 0, -1, 0, 0.0, 'float64'),
        ('corr', None, 0, 5, 0, 0.0, 'float64'),
        ('corr', None, 0, 0, 1, 0.5, 'float64'),
        ('corr', 'bdiag', 0, 0, 2, 0.25, 'float64'),
        ('corr', None, 0, 0, 0, 0.0, 'float32'),
        ('corr', None, 0, -1, 0, 0.0, 'float32'),
        ('toep_gram', None, 0, 0, 0, 0.0, 'float64'),
        ('toep_gram', None, 0, -1, 0, 0.0, 'float64'),
        ('toep_bidiag', None, 1, 0, 0, 0.0, 'float64'),
        ('toep_bidiag', None, 1, -1, 0, 0.0, 'float64'),
        ('toep_bidiag', None, 1, 4, 0, 0.0, 'float64'),
        ('toep_bidiag', None, 1, 0, 1, 0.5, 'float64'),
        ('corr', None, 1, 0, 0, 0.0, 'float64'),
    ]
    m, ns, seed = 20, 4, 7
    for c, (mname, bname, gram, ortho, kind, t, dt) in enumerate(rec):
        dtype = numpy.dtype(dt)
        tol = float(numpy.finfo(dtype).eps)
        A = mats[mname]
        B = mats[bname] if bname else None
        a, b, sz = ref_recurrences(R, A, gram, seed, ns, m, tol, ortho, kind,
                                   t, B, dtype.type)
        out['rec_%d_spec' % c] = numpy.array(
            [mname, bname or '', str(gram), str(ortho), str(kind), repr(t),
             dt, str(m), str(ns), str(seed)])
        out['rec_%d_alpha' % c] = a
        out['rec_%d_beta' % c] = b
        out['rec_%d_sizes' % c] = sz

    # convergence tools on a samples array with outliers
    rng = numpy.random.default_rng(3)
    S = rng.standard_normal((60, 3)) + numpy.array([5.0, -2.0, 0.1])
    S[7, 0] += 40.0
    S[13, 2] -= 25.0
    processed = numpy.arange(60, dtype=numpy.int32)[::-1].copy()
    err = numpy.zeros(3)
    used = numpy.zeros(3, numpy.int32)
    conv = numpy.zeros(3, numpy.int32)
    R.ref_check_convergence.restype = ctypes.c_int
    flags = []
    for k in range(1, 61):
        flags.append(R.ref_check_convergence(
            ptr(S), 60, 3, 10, ptr(processed), k, ctypes.c_double(0.95),
            ctypes.c_double(0.0), ctypes.c_double(0.05), ptr(err), ptr(used),
            ptr(conv)))
    out['conv_samples'] = S
    out['conv_processed'] = processed
    out['conv_flags'] = numpy.array(flags, numpy.int32)
    out['conv_error'] = err.copy()
    out['conv_used'] = used.copy()
    out['conv_converged'] = conv.copy()
    nout = numpy.zeros(3, numpy.int32)
    trace = numpy.zeros(3)
    err2 = numpy.zeros(3)
    used_all = numpy.array([60, 60, 45], numpy.int32)
    R.ref_average_estimates(ptr(S), 60, 3, ctypes.c_double(0.95),
                            ctypes.c_double(0.01), ptr(used_all),
                            ptr(processed), ptr(nout), ptr(trace), ptr(err2))
    out['avg_used'] = used_all
    out['avg_num_outliers'] = nout
    out['avg_trace'] = trace
    out['avg_error'] = err2

    path = os.path.join(HERE, 'ref_vectors.npz')
    numpy.savez_compressed(path, **out)
    print('wrote', path, os.path.getsize(path), 'bytes')


if __name__ == '__main__':
    main()
