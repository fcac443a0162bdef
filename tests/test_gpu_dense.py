"""GPU tests of the dense operator (imate's cuDenseMatrix counterpart): the
matrix-core dense SpMM against numpy products, and SLQ on dense matrices
against the oracle (which runs on the CSR copy of the same matrix) on the
same seeded probes.

Tolerances: products to 1e-13 of the largest entry (FP64) and 1e-5 (FP32);
SLQ per sample to 1e-10 relative (FP64), like every FP64 form; the
mixed-precision form to 1e-5 of the FP64 result.
"""

import ctypes

import numpy
import pytest
import scipy.sparse

import imate_amd
from imate_amd import _lib
from oracle import oracle

pytestmark = pytest.mark.gpu


def _dense_handle(lib, A):
    D = numpy.ascontiguousarray(A)
    h = ctypes.c_void_p()
    _lib.check(lib.slq_op_create_dense(D.shape[0], D.shape[1], _lib.ptr(D),
                                       _lib.dtype_code(D.dtype),
                                       ctypes.byref(h)), 'slq_op_create_dense')
    return h


@pytest.mark.parametrize('shape', [(300, 200), (130, 70), (1000, 1000),
                                   (64, 4100)])
@pytest.mark.parametrize('dtype', [numpy.float64, numpy.float32])
def test_dense_spmm_matches_numpy(gpu_lib, shape, dtype):
    rng = numpy.random.default_rng(sum(shape))
    A = rng.standard_normal(shape).astype(dtype)
    h = _dense_handle(gpu_lib, A)
    try:
        X = rng.standard_normal((shape[1], 64)).astype(dtype)
        Y = numpy.empty((shape[0], 64), dtype=dtype)
        _lib.check(gpu_lib.slq_op_spmm(h, 0.0, 64, 0, _lib.ptr(X), _lib.ptr(Y)),
                   'slq_op_spmm')
        ref = A.astype(numpy.float64) @ X.astype(numpy.float64)
        tol = 1e-13 if dtype == numpy.float64 else 1e-5
        assert numpy.max(numpy.abs(Y - ref)) <= tol * numpy.abs(ref).max()
        U = rng.standard_normal((shape[0], 64)).astype(dtype)
        Z = numpy.empty((shape[1], 64), dtype=dtype)
        _lib.check(gpu_lib.slq_op_spmm(h, 0.0, 64, 1, _lib.ptr(U), _lib.ptr(Z)),
                   'slq_op_spmm')
        refz = A.T.astype(numpy.float64) @ U.astype(numpy.float64)
        assert numpy.max(numpy.abs(Z - refz)) <= tol * numpy.abs(refz).max()
    finally:
        gpu_lib.slq_op_destroy(h)


def test_dense_affine_shift_spmm(gpu_lib):
    rng = numpy.random.default_rng(3)
    A = rng.standard_normal((257, 257))
    h = _dense_handle(gpu_lib, A)
    ha = ctypes.c_void_p()
    _lib.check(gpu_lib.slq_op_create_affine(h, None, ctypes.byref(ha)),
               'slq_op_create_affine')
    try:
        X = rng.standard_normal((257, 64))
        Y = numpy.empty((257, 64))
        _lib.check(gpu_lib.slq_op_spmm(ha, 0.75, 64, 0, _lib.ptr(X), _lib.ptr(Y)),
                   'slq_op_spmm')
        ref = A @ X + 0.75 * X
        assert numpy.max(numpy.abs(Y - ref)) <= 1e-13 * numpy.abs(ref).max()
    finally:
        gpu_lib.slq_op_destroy(ha)
        gpu_lib.slq_op_destroy(h)


def _spd(n, seed):
    rng = numpy.random.default_rng(seed)
    R = rng.uniform(-1.0, 1.0, (n, n))
    return 0.5 * (R + R.T) + 1.5 * numpy.sqrt(n) * numpy.eye(n)


@pytest.mark.parametrize('func', ['log', 'inv'])
def test_dense_slq_matches_oracle(gpu_lib, func):
    A = _spd(700, 1)
    fn = imate_amd.logdet if func == 'log' else imate_amd.traceinv
    val, info = fn(A, method='slq', gpu=True, lanczos_degree=20,
                   min_num_samples=70, max_num_samples=70, seed=2,
                   return_info=True)
    ref = oracle.slq(scipy.sparse.csr_matrix(A), func=func, lanczos_degree=20,
                     min_num_samples=70, max_num_samples=70, seed=2)
    got = info['convergence']['samples']
    want = ref['samples'][:, 0]
    assert numpy.max(numpy.abs(got - want) / numpy.abs(want)) < 1e-10
    exact = numpy.linalg.slogdet(A)[1] if func == 'log' else \
        numpy.trace(numpy.linalg.inv(A))
    assert abs(val - exact) <= 3 * info['error']['absolute_error'] + 1e-12


def test_dense_affine_and_gram_match_oracle(gpu_lib):
    A = _spd(300, 4)
    t = [0.05, 2.0]
    val, info = imate_amd.logdet(imate_amd.AffineMatrixFunction(A),
                                 parameters=t, method='slq', gpu=True,
                                 lanczos_degree=15, min_num_samples=64,
                                 max_num_samples=64, seed=6, return_info=True)
    ref = oracle.slq(scipy.sparse.csr_matrix(A), func='log', parameters=t,
                     affine=True, lanczos_degree=15, min_num_samples=64,
                     max_num_samples=64, seed=6)
    got = numpy.asarray(info['convergence']['samples'])
    assert numpy.max(numpy.abs(got - ref['samples']) /
                     numpy.abs(ref['samples'])) < 1e-10
    rng = numpy.random.default_rng(5)
    G = rng.standard_normal((450, 200))
    val, info = imate_amd.logdet(G, gram=True, method='slq', gpu=True,
                                 lanczos_degree=20, min_num_samples=64,
                                 max_num_samples=64, seed=7, return_info=True)
    ref = oracle.slq(scipy.sparse.csr_matrix(G), func='log', gram=True,
                     lanczos_degree=20, min_num_samples=64,
                     max_num_samples=64, seed=7)
    got = info['convergence']['samples']
    assert numpy.max(numpy.abs(got - ref['samples'][:, 0]) /
                     numpy.abs(ref['samples'][:, 0])) < 1e-10


@pytest.mark.parametrize('ortho', [0, -1])
def test_dense_mixed_precision_close_to_fp64(gpu_lib, ortho):
    A = _spd(500, 8)
    kw = dict(method='slq', gpu=True, lanczos_degree=20, min_num_samples=64,
              max_num_samples=64, seed=3, orthogonalize=ortho)
    a = imate_amd.logdet(A, **kw)
    b = imate_amd.logdet(A, precision='mixed', **kw)
    assert abs(a - b) <= 1e-5 * abs(a)


def test_dense_and_csr_paths_agree(gpu_lib):
    A = _spd(400, 9)
    kw = dict(method='slq', gpu=True, lanczos_degree=20, min_num_samples=64,
              max_num_samples=64, seed=1)
    a = imate_amd.logdet(A, **kw)
    b = imate_amd.logdet(scipy.sparse.csr_matrix(A), **kw)
    assert abs(a - b) <= 1e-10 * abs(b)


def test_dense_set_values_and_rejections(gpu_lib):
    rng = numpy.random.default_rng(2)
    A = rng.standard_normal((129, 129))
    h = _dense_handle(gpu_lib, A)
    try:
        A2 = rng.standard_normal((129, 129))
        _lib.check(gpu_lib.slq_op_set_values(h, _lib.ptr(A2)), 'slq_op_set_values')
        X = rng.standard_normal((129, 64))
        Y = numpy.empty((129, 64))
        _lib.check(gpu_lib.slq_op_spmm(h, 0.0, 64, 0, _lib.ptr(X), _lib.ptr(Y)),
                   'slq_op_spmm')
        assert numpy.max(numpy.abs(Y - A2 @ X)) <= 1e-13 * numpy.abs(A2 @ X).max()
        ip = numpy.empty(130, dtype=numpy.int64)
        assert gpu_lib.slq_op_get_csr(h, _lib.ptr(ip), None, None) == _lib.SLQ_E_ARG
        hb = ctypes.c_void_p()
        assert gpu_lib.slq_op_create_affine(h, h, ctypes.byref(hb)) == \
            _lib.SLQ_E_ARG
    finally:
        gpu_lib.slq_op_destroy(h)
