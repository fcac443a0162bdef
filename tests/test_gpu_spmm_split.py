"""The matrix-core SpMM with two or four wavefronts per tile
(slq_op_set_launch spmm_split = 2, 4): per-sample SLQ against the oracle at
1e-10 (FP64) and 1e-5 (mixed precision), Lanczos and Golub-Kahn (whose
A^T u runs on the transposed tiles), tile counts that leave teams of the
last workgroup without a tile, and the products against numpy."""

import ctypes

import numpy
import pytest
import scipy.sparse

import imate_amd
from imate_amd import _lib
from imate_amd.sample_matrices import sparse_correlation_matrix
from oracle import oracle

pytestmark = pytest.mark.gpu


def _matrix(A, split):
    M = imate_amd.Matrix(A)
    h = M.get_linear_operator(gpu=True).handle
    _lib.check(_lib.load().slq_op_set_launch(h, 0, 0, split, 0),
               'slq_op_set_launch')
    return M


@pytest.mark.parametrize('split', [2, 4])
@pytest.mark.parametrize('n', [4096, 4099, 47])
def test_split_lanczos_vs_oracle(split, n):
    A = sparse_correlation_matrix(n, nnz_per_row=min(40, n - 1), seed=1)
    M = _matrix(A, split)
    for func, fn in (('log', imate_amd.logdet), ('inv', imate_amd.traceinv)):
        val, info = fn(M, method='slq', gpu=True, lanczos_degree=20,
                       min_num_samples=70, max_num_samples=70, seed=2,
                       return_info=True)
        ref = oracle.slq(A, func=func, lanczos_degree=20, min_num_samples=70,
                         max_num_samples=70, seed=2)
        got = info['convergence']['samples']
        want = ref['samples'][:, 0]
        assert numpy.max(numpy.abs(got - want) / numpy.abs(want)) < 1e-10


@pytest.mark.parametrize('split', [2, 4])
def test_split_reorth_mixed_and_affine(split):
    A = sparse_correlation_matrix(3000, nnz_per_row=30, seed=3)
    M = _matrix(A, split)
    kw = dict(method='slq', gpu=True, lanczos_degree=25, min_num_samples=64,
              max_num_samples=64, seed=4, return_info=True)
    val, info = imate_amd.logdet(M, orthogonalize=-1, precision='mixed', **kw)
    ref = oracle.slq(A, func='log', lanczos_degree=25, min_num_samples=64,
                     max_num_samples=64, seed=4, orthogonalize=-1)
    rel = numpy.abs(info['convergence']['samples'] - ref['samples'][:, 0]) / \
        numpy.abs(ref['samples'][:, 0])
    assert rel.max() < 1e-5
    F = imate_amd.AffineMatrixFunction(A)
    h = F.get_linear_operator(gpu=True).handle
    _lib.check(_lib.load().slq_op_set_launch(h, 0, 0, split, 0),
               'slq_op_set_launch')
    ts = [1e-3, 0.5]
    vals, info = imate_amd.logdet(F, parameters=ts, **kw)
    ref = oracle.slq(A, func='log', parameters=ts, affine=True,
                     lanczos_degree=25, min_num_samples=64,
                     max_num_samples=64, seed=4)
    rel = numpy.abs(info['convergence']['samples'] - ref['samples']) / \
        numpy.abs(ref['samples'])
    assert rel.max() < 1e-10


@pytest.mark.parametrize('split', [2, 4])
def test_split_gram_vs_oracle(split):
    # a banded rectangular matrix: dense enough for 16-row tiles in both
    # directions (A v and A^T u)
    rng = numpy.random.default_rng(5)
    r, c = 2000, 1200
    rows, cols, vals = [], [], []
    for i in range(r):
        j0 = int(i * c / r)
        for j in range(max(0, j0 - 6), min(c, j0 + 7)):
            rows.append(i)
            cols.append(j)
            vals.append(rng.standard_normal() + (3.0 if j == j0 else 0.0))
    A = scipy.sparse.csr_matrix((vals, (rows, cols)), shape=(r, c))
    M = _matrix(A, split)
    val, info = imate_amd.logdet(M, gram=True, method='slq', gpu=True,
                                 lanczos_degree=15, min_num_samples=64,
                                 max_num_samples=64, seed=6, return_info=True)
    ref = oracle.slq(A, func='log', gram=True, lanczos_degree=15,
                     min_num_samples=64, max_num_samples=64, seed=6)
    got = info['convergence']['samples']
    want = ref['samples'][:, 0]
    assert numpy.max(numpy.abs(got - want) / numpy.abs(want)) < 1e-10


@pytest.mark.parametrize('split', [1, 2, 4])
def test_split_products_vs_numpy(gpu_lib, split):
    A = sparse_correlation_matrix(5003, nnz_per_row=35, seed=7)
    M = _matrix(A, split)
    h = M.gpu_Aop.handle
    X = numpy.random.default_rng(8).standard_normal((5003, 64))
    Y = numpy.empty((5003, 64))
    _lib.check(gpu_lib.slq_op_spmm(h, 0.0, 64, 0, _lib.ptr(X), _lib.ptr(Y)),
               'slq_op_spmm')
    ref = A @ X
    assert numpy.max(numpy.abs(Y - ref)) <= 1e-13 * numpy.abs(ref).max()
    got = ctypes.c_int32(0)
    _lib.check(gpu_lib.slq_op_get_launch(h, None, None, ctypes.byref(got),
                                         None), 'slq_op_get_launch')
    assert got.value == split
    assert gpu_lib.slq_op_set_launch(h, 0, 0, 3, 0) == _lib.SLQ_E_ARG
    assert gpu_lib.slq_op_set_launch(h, 0, 0, 1, 2) == _lib.SLQ_E_ARG


def test_csr_operand_form_vs_oracle_and_tiles():
    """tile_mode 1 (the CSR rows on a handle that has 16-row tiles): the
    oracle's samples at 1e-10, and different bits from the tile path (so
    the run did take the other path)."""
    A = sparse_correlation_matrix(4096, nnz_per_row=40, seed=9)
    kw = dict(method='slq', gpu=True, lanczos_degree=20, min_num_samples=64,
              max_num_samples=64, seed=3, return_info=True)
    M = imate_amd.Matrix(A)
    h = M.get_linear_operator(gpu=True).handle
    _, tiles = imate_amd.logdet(M, **kw)
    _lib.check(_lib.load().slq_op_set_launch(h, 0, 0, 0, 1),
               'slq_op_set_launch')
    _, csr = imate_amd.logdet(M, **kw)
    ref = oracle.slq(A, func='log', lanczos_degree=20, min_num_samples=64,
                     max_num_samples=64, seed=3)
    want = ref['samples'][:, 0]
    for info in (tiles, csr):
        got = info['convergence']['samples']
        assert numpy.max(numpy.abs(got - want) / numpy.abs(want)) < 1e-10
    assert not numpy.array_equal(tiles['convergence']['samples'],
                                 csr['convergence']['samples'])
