"""slq_estimate_noalloc (the entry point that allocates no device memory)
against slq_estimate through the C ABI: two calls in a row, the second on a
smaller input, plus an A + tI handle with several inquiries, Golub-Kahn
after its transpose exists, and the limits it reports as errors."""

import ctypes

import numpy
import pytest
import scipy.sparse

from imate_amd import _lib
from imate_amd.sample_matrices import sparse_correlation_matrix

pytestmark = pytest.mark.gpu


def _csr_handle(lib, A):
    A = A.tocsr()
    A.sort_indices()
    indptr = A.indptr.astype(numpy.int64)
    ix = A.indices.astype(numpy.int32)
    data = numpy.ascontiguousarray(A.data)
    dt = _lib.SLQ_F64 if data.dtype == numpy.float64 else _lib.SLQ_F32
    h = ctypes.c_void_p()
    _lib.check(lib.slq_op_create_csr(A.shape[0], A.shape[1], A.nnz,
                                     _lib.ptr(indptr), _lib.ptr(ix),
                                     _lib.ptr(data), dt, ctypes.byref(h)),
               'slq_op_create_csr')
    return h


def _estimate(lib, entry, h, m, s, params=None, gram=0, ortho=0, func=0,
              precision=0, seed=3):
    q = 1 if params is None else len(params)
    par = None if params is None else numpy.ascontiguousarray(params,
                                                             dtype=float)
    out = {'trace': numpy.zeros(q), 'error': numpy.zeros(q),
           'samples': numpy.zeros((s, q)),
           'processed': numpy.zeros(s, dtype=numpy.int32),
           'num_used': numpy.zeros(q, dtype=numpy.int32),
           'num_outliers': numpy.zeros(q, dtype=numpy.int32),
           'converged': numpy.zeros(q, dtype=numpy.int32)}
    allc = ctypes.c_int32(0)
    wall = ctypes.c_double(0)
    tol = float(numpy.finfo(float).eps)
    r = getattr(lib, entry)(
        h, None if par is None else _lib.ptr(par), q, func, None, gram, 1.0,
        ortho, seed, m, tol, s, s, 0.0, 1e-2, 0.95, 0.001, 0, precision,
        _lib.ptr(out['trace']), _lib.ptr(out['error']),
        _lib.ptr(out['samples']), _lib.ptr(out['processed']),
        _lib.ptr(out['num_used']), _lib.ptr(out['num_outliers']),
        _lib.ptr(out['converged']), ctypes.byref(allc), ctypes.byref(wall))
    return r, out


def _same(a, b):
    for k in a:
        assert numpy.array_equal(a[k], b[k], equal_nan=True), k


def test_noalloc_equals_estimate_twice_then_smaller(gpu_lib):
    lib = gpu_lib
    big = _csr_handle(lib, sparse_correlation_matrix(6000, nnz_per_row=30,
                                                     seed=1))
    small = _csr_handle(lib, sparse_correlation_matrix(900, nnz_per_row=20,
                                                       seed=2))
    try:
        for h, m, s in ((big, 20, 150), (small, 12, 70)):
            r0, want = _estimate(lib, 'slq_estimate', h, m, s)
            _lib.check(r0, 'slq_estimate')
            r1, got = _estimate(lib, 'slq_estimate_noalloc', h, m, s)
            _lib.check(r1, 'slq_estimate_noalloc')
            _same(want, got)
            assert numpy.all(numpy.isfinite(got['samples']))
        # the shared array keeps the last call's contents: the big handle
        # again after the small one gives its first result once more
        r2, again = _estimate(lib, 'slq_estimate_noalloc', big, 20, 150)
        _lib.check(r2, 'slq_estimate_noalloc')
        r3, want = _estimate(lib, 'slq_estimate', big, 20, 150)
        _same(want, again)
    finally:
        lib.slq_op_destroy(big)
        lib.slq_op_destroy(small)


def test_noalloc_affine_reorth_and_gram(gpu_lib):
    lib = gpu_lib
    A = sparse_correlation_matrix(3000, nnz_per_row=25, seed=4)
    h = _csr_handle(lib, A)
    aff = ctypes.c_void_p()
    _lib.check(lib.slq_op_create_affine(h, None, ctypes.byref(aff)),
               'slq_op_create_affine')
    rect = scipy.sparse.random(1500, 500, density=0.03, random_state=5,
                               format='csr')
    rect = (rect + scipy.sparse.eye(1500, 500, format='csr')).tocsr()
    g = _csr_handle(lib, rect)
    try:
        ts = [1e-3, 0.1, 2.0]
        for kw in (dict(params=ts), dict(params=ts, ortho=-1, func=1)):
            r0, want = _estimate(lib, 'slq_estimate', aff, 15, 90, **kw)
            _lib.check(r0, 'slq_estimate')
            r1, got = _estimate(lib, 'slq_estimate_noalloc', aff, 15, 90, **kw)
            _lib.check(r1, 'slq_estimate_noalloc')
            _same(want, got)
        # Golub-Kahn needs the transpose: refused before it exists
        r, _ = _estimate(lib, 'slq_estimate_noalloc', g, 10, 40, gram=1)
        assert r == _lib.SLQ_E_ARG
        r0, want = _estimate(lib, 'slq_estimate', g, 10, 40, gram=1)
        _lib.check(r0, 'slq_estimate')
        r1, got = _estimate(lib, 'slq_estimate_noalloc', g, 10, 40, gram=1)
        _lib.check(r1, 'slq_estimate_noalloc')
        _same(want, got)
    finally:
        lib.slq_op_destroy(aff)
        lib.slq_op_destroy(h)
        lib.slq_op_destroy(g)


def test_noalloc_limits(gpu_lib):
    lib = gpu_lib
    # 2^19 rows in FP64: four 256 MiB probe blocks exceed the fixed array
    n = 1 << 19
    D = scipy.sparse.diags(numpy.linspace(1.0, 2.0, n), format='csr')
    h = _csr_handle(lib, D)
    try:
        r, _ = _estimate(lib, 'slq_estimate_noalloc', h, 5, 4)
        assert r == _lib.SLQ_E_ARG
        buf = ctypes.create_string_buffer(512)
        lib.slq_last_error(buf, 512)
        assert b'256' in buf.value or b'workspace' in buf.value
        # profiling on the handle is refused as well
        small = _csr_handle(lib, sparse_correlation_matrix(500, nnz_per_row=10,
                                                           seed=6))
        try:
            _lib.check(lib.slq_profile_enable(small, 1), 'slq_profile_enable')
            r, _ = _estimate(lib, 'slq_estimate_noalloc', small, 5, 10)
            assert r == _lib.SLQ_E_ARG
            _lib.check(lib.slq_profile_enable(small, 0), 'slq_profile_enable')
            r, _ = _estimate(lib, 'slq_estimate_noalloc', small, 5, 10)
            _lib.check(r, 'slq_estimate_noalloc')
        finally:
            lib.slq_op_destroy(small)
    finally:
        lib.slq_op_destroy(h)
