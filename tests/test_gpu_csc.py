"""GPU tests of CSC handles (imate's cuCSCMatrix counterpart): the CSR the
library builds from CSC arrays against scipy's conversion, products against
numpy, SLQ against the oracle (which runs on the CSR of the same matrix) on
the same seeded probes, and new values given in CSC order.

Tolerances: the converted CSR bit for bit; products to 1e-13 of the largest
entry (FP64); SLQ per sample to 1e-10 relative, like every FP64 form.
"""

import ctypes

import numpy
import pytest
import scipy.sparse

import imate_amd
from imate_amd import _lib
from imate_amd.sample_matrices import sparse_correlation_matrix
from oracle import oracle

pytestmark = pytest.mark.gpu


def _csc_handle(lib, M):
    ip = numpy.ascontiguousarray(M.indptr, dtype=numpy.int64)
    ix = numpy.ascontiguousarray(M.indices, dtype=numpy.int32)
    dv = numpy.ascontiguousarray(M.data)
    h = ctypes.c_void_p()
    _lib.check(lib.slq_op_create_csc(M.shape[0], M.shape[1], M.nnz,
                                     _lib.ptr(ip), _lib.ptr(ix), _lib.ptr(dv),
                                     _lib.dtype_code(dv.dtype),
                                     ctypes.byref(h)), 'slq_op_create_csc')
    return h


def _get_csr(lib, h, shape, nnz, dtype):
    ip = numpy.empty(shape[0] + 1, dtype=numpy.int64)
    ix = numpy.empty(nnz, dtype=numpy.int32)
    dv = numpy.empty(nnz, dtype=dtype)
    _lib.check(lib.slq_op_get_csr(h, _lib.ptr(ip), _lib.ptr(ix), _lib.ptr(dv)),
               'slq_op_get_csr')
    return ip, ix, dv


def _unsorted_csc(shape, seed):
    """Random CSC with the row indices of every column shuffled."""
    M = scipy.sparse.random(shape[0], shape[1], density=0.02,
                            random_state=seed, format='csc')
    rng = numpy.random.default_rng(seed)
    ix, dv = M.indices.copy(), M.data.copy()
    for c in range(shape[1]):
        a, b = M.indptr[c], M.indptr[c + 1]
        p = rng.permutation(b - a) + a
        ix[a:b], dv[a:b] = ix[p], dv[p]
    return scipy.sparse.csc_matrix((dv, ix, M.indptr.copy()), shape=shape)


@pytest.mark.parametrize('shape', [(500, 500), (700, 300), (64, 2000)])
def test_csc_conversion_and_products(gpu_lib, shape):
    M = _unsorted_csc(shape, sum(shape))
    h = _csc_handle(gpu_lib, M)
    try:
        ip, ix, dv = _get_csr(gpu_lib, h, shape, M.nnz, M.dtype)
        ref = M.tocsr()
        ref.sort_indices()
        assert numpy.array_equal(ip, ref.indptr)
        assert numpy.array_equal(ix, ref.indices)
        assert numpy.array_equal(dv, ref.data)
        rng = numpy.random.default_rng(1)
        X = rng.standard_normal((shape[1], 64))
        Y = numpy.empty((shape[0], 64))
        _lib.check(gpu_lib.slq_op_spmm(h, 0.0, 64, 0, _lib.ptr(X), _lib.ptr(Y)),
                   'slq_op_spmm')
        want = M @ X
        assert numpy.max(numpy.abs(Y - want)) <= 1e-13 * numpy.abs(want).max()
        U = rng.standard_normal((shape[0], 64))
        Z = numpy.empty((shape[1], 64))
        _lib.check(gpu_lib.slq_op_spmm(h, 0.0, 64, 1, _lib.ptr(U), _lib.ptr(Z)),
                   'slq_op_spmm')
        want = M.T @ U
        assert numpy.max(numpy.abs(Z - want)) <= 1e-13 * numpy.abs(want).max()
    finally:
        gpu_lib.slq_op_destroy(h)


def test_csc_set_values_in_csc_order(gpu_lib):
    M = sparse_correlation_matrix(3000, nnz_per_row=25, seed=3).tocsc()
    h = _csc_handle(gpu_lib, M)
    try:
        new = numpy.random.default_rng(4).standard_normal(M.nnz)
        _lib.check(gpu_lib.slq_op_set_values(h, _lib.ptr(new)),
                   'slq_op_set_values')
        M2 = scipy.sparse.csc_matrix((new, M.indices, M.indptr), shape=M.shape)
        X = numpy.random.default_rng(5).standard_normal((3000, 64))
        Y = numpy.empty((3000, 64))
        _lib.check(gpu_lib.slq_op_spmm(h, 0.0, 64, 0, _lib.ptr(X), _lib.ptr(Y)),
                   'slq_op_spmm')
        want = M2 @ X
        assert numpy.max(numpy.abs(Y - want)) <= 1e-13 * numpy.abs(want).max()
    finally:
        gpu_lib.slq_op_destroy(h)


@pytest.mark.parametrize('func', ['log', 'inv'])
def test_csc_slq_matches_oracle(func):
    A = sparse_correlation_matrix(4096, nnz_per_row=40, seed=0)
    fn = imate_amd.logdet if func == 'log' else imate_amd.traceinv
    val, info = fn(A.tocsc(), method='slq', gpu=True, lanczos_degree=20,
                   min_num_samples=64, max_num_samples=64, seed=2,
                   return_info=True)
    ref = oracle.slq(A, func=func, lanczos_degree=20, min_num_samples=64,
                     max_num_samples=64, seed=2)
    got = info['convergence']['samples']
    want = ref['samples'][:, 0]
    assert numpy.max(numpy.abs(got - want) / numpy.abs(want)) < 1e-10


def test_csc_gram_and_affine_match_oracle():
    G = scipy.sparse.random(600, 250, density=0.03, random_state=6,
                            format='csc')
    G = (G + scipy.sparse.eye(600, 250, format='csc') * 0.5).tocsc()
    val, info = imate_amd.logdet(G, gram=True, method='slq', gpu=True,
                                 lanczos_degree=20, min_num_samples=64,
                                 max_num_samples=64, seed=7, orthogonalize=-1,
                                 return_info=True)
    ref = oracle.slq(G.tocsr(), func='log', gram=True, lanczos_degree=20,
                     min_num_samples=64, max_num_samples=64, seed=7,
                     orthogonalize=-1)
    got = info['convergence']['samples']
    assert numpy.max(numpy.abs(got - ref['samples'][:, 0]) /
                     numpy.abs(ref['samples'][:, 0])) < 1e-10
    A = sparse_correlation_matrix(2048, nnz_per_row=30, seed=1)
    B = scipy.sparse.diags(numpy.linspace(0.5, 1.5, 2048)).tocsc()
    t = [0.05, 1.0]
    val, info = imate_amd.logdet(imate_amd.AffineMatrixFunction(A.tocsc(), B),
                                 parameters=t, method='slq', gpu=True,
                                 lanczos_degree=16, min_num_samples=64,
                                 max_num_samples=64, seed=8, return_info=True)
    ref = oracle.slq(A, func='log', parameters=t, B=B.tocsr(), affine=True,
                     lanczos_degree=16, min_num_samples=64,
                     max_num_samples=64, seed=8)
    got = numpy.asarray(info['convergence']['samples'])
    assert numpy.max(numpy.abs(got - ref['samples']) /
                     numpy.abs(ref['samples'])) < 1e-10
