"""Host-side tests of the drop-in (CPU only): the C ABI library loads and
exports every entry point that include/slq_hip.h declares, the Python side
binds exactly those, and argument validation behaves like imate's
(imate/_trace_estimator/trace_estimator_utilities.py:108-305) before any
device work starts."""

import ctypes
import os
import re

import numpy
import pytest
import scipy.sparse

import imate_amd
from imate_amd import _lib
from imate_amd.sample_matrices import (toeplitz, sparse_correlation_matrix,
                                       morton_order, hilbert_order,
                                       correlation_scale)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, 'include', 'slq_hip.h')) as f:
        src = f.read()
    src = re.sub(r'/\*[\s\S]*?\*/', '', src)
    return set(re.findall(r'\bint\s+(slq_\w+)\s*\(', src))


def test_header_declares_the_expected_entry_points():
    names = _declared()
    assert {'slq_estimate', 'slq_op_create_csr', 'slq_op_create_affine',
            'slq_op_destroy', 'slq_rademacher', 'slq_op_spmm',
            'slq_recurrences', 'slq_quadrature', 'slq_last_error',
            'slq_device_properties'} <= names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_lib.SIGNATURES) == _declared()


def test_error_path_without_device_is_an_exception():
    # device queries must fail loudly (never fall back) when no GPU is seen
    try:
        props = _lib.device_properties()
    except RuntimeError as e:
        assert 'device' in str(e).lower()
    else:
        assert props['num_devices'] >= 1


def test_invalid_csr_rejected_by_the_c_abi():
    lib = _lib.load()
    indptr = numpy.array([0, 1, 2], dtype=numpy.int64)
    indices = numpy.array([0, 5], dtype=numpy.int32)   # column 5 of 2
    data = numpy.ones(2)
    h = ctypes.c_void_p()
    r = lib.slq_op_create_csr(2, 2, 2, _lib.ptr(indptr), _lib.ptr(indices),
                              _lib.ptr(data), _lib.SLQ_F64, ctypes.byref(h))
    assert r in (_lib.SLQ_E_ARG, _lib.SLQ_E_NODEVICE)


def test_invalid_csc_rejected_by_the_c_abi():
    lib = _lib.load()
    indptr = numpy.array([0, 1, 2], dtype=numpy.int64)
    data = numpy.ones(2)
    h = ctypes.c_void_p()
    for rows in ([0, 3], [-1, 0]):       # row 3 of 3 rows; a negative row
        indices = numpy.array(rows, dtype=numpy.int32)
        r = lib.slq_op_create_csc(3, 2, 2, _lib.ptr(indptr), _lib.ptr(indices),
                                  _lib.ptr(data), _lib.SLQ_F64,
                                  ctypes.byref(h))
        assert r == _lib.SLQ_E_ARG
    bad_ptr = numpy.array([0, 2, 1, 2], dtype=numpy.int64)   # decreasing
    indices = numpy.array([0, 1], dtype=numpy.int32)
    r = lib.slq_op_create_csc(3, 3, 2, _lib.ptr(bad_ptr), _lib.ptr(indices),
                              _lib.ptr(data), _lib.SLQ_F64, ctypes.byref(h))
    assert r == _lib.SLQ_E_ARG


A_SMALL = toeplitz(2.0, 1.0, 20, gram=True)


@pytest.mark.parametrize('kw, exc', [
    (dict(min_num_samples=20, max_num_samples=10), ValueError),
    (dict(min_num_samples=0), ValueError),
    (dict(max_num_samples=1.5), TypeError),
    (dict(lanczos_degree=0), ValueError),
    (dict(lanczos_degree=10, orthogonalize=11), ValueError),
    (dict(confidence_level=1.5), ValueError),
    (dict(confidence_level=0.95, outlier_significance_level=0.1), ValueError),
    (dict(error_atol=-1.0), ValueError),
    (dict(error_rtol=-1.0), ValueError),
    (dict(lanczos_tol=-1.0), ValueError),
    (dict(seed=1.5), TypeError),
    (dict(num_gpu_devices=-1), ValueError),
    (dict(verbose=1), TypeError),
    (dict(plot=None), TypeError),
    (dict(gpu=None), TypeError),
    (dict(precision='half'), ValueError),
    (dict(block_size=65), ValueError),
])
def test_argument_validation(kw, exc):
    kw = dict(kw)
    kw.setdefault('gpu', True)
    with pytest.raises(exc):
        imate_amd.logdet(A_SMALL, method='slq', **kw)


def test_return_info_must_be_bool():
    with pytest.raises(TypeError):
        imate_amd.logdet(A_SMALL, method='slq', return_info=1, gpu=True)


def test_non_square_without_gram_rejected():
    A = scipy.sparse.random(5, 4, density=0.5, format='csr')
    with pytest.raises(ValueError):
        imate_amd.logdet(A, method='slq', gpu=True)


def test_other_methods_are_out_of_scope():
    with pytest.raises(NotImplementedError):
        imate_amd.logdet(A_SMALL, method='cholesky')
    with pytest.raises(ValueError):
        imate_amd.traceinv(A_SMALL, method='nonsense')


def test_affine_needs_parameters():
    Aop = imate_amd.AffineMatrixFunction(A_SMALL)
    with pytest.raises(ValueError):
        imate_amd.logdet(Aop, method='slq', gpu=True)


def test_float128_rejected_on_gpu():
    if not hasattr(numpy, 'float128'):
        pytest.skip('no float128 on this platform')
    A = A_SMALL.astype(numpy.float128)
    with pytest.raises(TypeError):
        imate_amd.logdet(A, method='slq', gpu=True)


def test_toeplitz_matches_definition():
    A = toeplitz(3.0, 0.5, 6).toarray()
    expect = numpy.diag(numpy.full(6, 3.0)) + numpy.diag(numpy.full(5, 0.5), 1)
    assert numpy.array_equal(A, expect)
    G = toeplitz(3.0, 0.5, 6, gram=True).toarray()
    assert numpy.allclose(G, expect.T @ expect)


def test_correlation_matrix_is_spd_and_sparse():
    A = sparse_correlation_matrix(2000, nnz_per_row=30, seed=4)
    assert abs(A.nnz / 2000 - 30) < 4
    assert (A - A.T).nnz == 0
    w = numpy.linalg.eigvalsh(A.toarray())
    assert w.min() > 0
    # the hard-thresholded kernel with dense neighbourhoods is not SPD, the
    # tapered one is
    B = sparse_correlation_matrix(2000, scale=0.02, seed=4)
    assert numpy.linalg.eigvalsh(B.toarray()).min() > 0
    assert correlation_scale(4096, 40) > 0


def test_hilbert_order_walks_the_grid_by_edge_neighbours():
    """On the cell centres of a 32 x 32 grid the curve visits every cell
    once, each step to an edge neighbour, from (0, 0) to (31, 0); a wrong
    reflection or swap of the quadrant frames breaks the adjacency."""
    g = numpy.array([(i, j) for i in range(32) for j in range(32)], float)
    p = hilbert_order((g + 0.5) / 32, bits=5)
    assert numpy.array_equal(numpy.sort(p), numpy.arange(1024))
    q = g[p]
    assert numpy.all(numpy.abs(numpy.diff(q, axis=0)).sum(axis=1) == 1)
    assert tuple(q[0]) == (0, 0) and tuple(q[-1]) == (31, 0)


def test_curve_orders_give_one_matrix_up_to_permutation():
    """The Hilbert and Morton orderings of the same seeded points are one
    matrix under a symmetric permutation: the same spectrum."""
    A = sparse_correlation_matrix(600, nnz_per_row=20, seed=2)
    B = sparse_correlation_matrix(600, nnz_per_row=20, seed=2, order='morton')
    assert A.nnz == B.nnz
    ea = numpy.linalg.eigvalsh(A.toarray())
    eb = numpy.linalg.eigvalsh(B.toarray())
    assert numpy.allclose(ea, eb, rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        sparse_correlation_matrix(600, nnz_per_row=20, order='rows')


def test_morton_order_is_a_permutation_with_locality():
    pts = numpy.random.default_rng(0).random((4096, 2))
    p = morton_order(pts)
    assert numpy.array_equal(numpy.sort(p), numpy.arange(4096))
    q = pts[p]
    step = numpy.sqrt(((q[1:] - q[:-1])**2).sum(axis=1))
    assert numpy.median(step) < 0.05
