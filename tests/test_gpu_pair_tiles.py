"""GPU tests of the tiled A + tB operator (the general affine operator on the
matrix-core SpMM): the union-pattern tiles with A + tB gathered per t,
against numpy products and against the oracle's SLQ on the same seeded
probes; against the two-stream CSR path (SLQ_PAIR_TILES=0) it runs beside.

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
from imate_amd.sample_matrices import sparse_correlation_matrix
from oracle import oracle

pytestmark = pytest.mark.gpu


def _csr_handle(lib, M):
    M = M.tocsr()
    ip = numpy.ascontiguousarray(M.indptr, dtype=numpy.int64)
    ix = numpy.ascontiguousarray(M.indices, dtype=numpy.int32)
    dv = numpy.ascontiguousarray(M.data)
    h = ctypes.c_void_p()
    _lib.check(lib.slq_op_create_csr(M.shape[0], M.shape[1], M.nnz,
                                     _lib.ptr(ip), _lib.ptr(ix), _lib.ptr(dv),
                                     _lib.dtype_code(dv.dtype),
                                     ctypes.byref(h)), 'slq_op_create_csr')
    return h


def _spd_other_pattern(n, nnz_per_row, seed, rseed):
    """SPD, on a pattern that holds the correlation matrix of the same size,
    seed and ``nnz_per_row`` (same points) and 60 % more entries, plus about
    one random symmetric pair per row far from the diagonal; the union with
    that matrix still fills more than a quarter of its 16-row tile slabs."""
    C = sparse_correlation_matrix(n, nnz_per_row=int(nnz_per_row * 1.6),
                                  seed=seed)
    R = scipy.sparse.random(n, n, density=0.5 / n, random_state=rseed,
                            format='csr')
    S = 0.1 * (R + R.T)
    d = numpy.asarray(abs(S).sum(axis=1)).ravel()
    return (C + S + scipy.sparse.diags(d)).tocsr()


def _spmm(lib, h, t, X):
    Y = numpy.empty((X.shape[0], X.shape[1]), dtype=X.dtype)
    _lib.check(lib.slq_op_spmm(h, t, X.shape[1], 0, _lib.ptr(X), _lib.ptr(Y)),
               'slq_op_spmm')
    return Y


@pytest.mark.parametrize('which_b', ['diag', 'other'])
@pytest.mark.parametrize('dtype', [numpy.float64, numpy.float32])
def test_pair_spmm_matches_numpy(gpu_lib, monkeypatch, which_b, dtype):
    n = 3000
    A = sparse_correlation_matrix(n, nnz_per_row=30, seed=4).astype(dtype)
    B = (scipy.sparse.diags(numpy.linspace(0.5, 1.5, n)).tocsr()
         if which_b == 'diag' else _spd_other_pattern(n, 30, 4, 5)).astype(dtype)
    ha, hb = _csr_handle(gpu_lib, A), _csr_handle(gpu_lib, B)
    hab = ctypes.c_void_p()
    _lib.check(gpu_lib.slq_op_create_affine(ha, hb, ctypes.byref(hab)),
               'slq_op_create_affine')
    try:
        rng = numpy.random.default_rng(1)
        X = rng.standard_normal((n, 64)).astype(dtype)
        tol = 1e-13 if dtype == numpy.float64 else 1e-5
        for t in (0.0, 0.37, -2.5, 0.37):
            ref = (A.astype(numpy.float64) + t * B.astype(numpy.float64)) @ \
                X.astype(numpy.float64)
            got = _spmm(gpu_lib, hab, t, X)
            assert numpy.max(numpy.abs(got - ref)) <= tol * numpy.abs(ref).max()
        monkeypatch.setenv('SLQ_PAIR_TILES', '0')
        csr = _spmm(gpu_lib, hab, 0.37, X)
        monkeypatch.delenv('SLQ_PAIR_TILES')
        pair = _spmm(gpu_lib, hab, 0.37, X)
        ref = (A.astype(numpy.float64) + 0.37 * B.astype(numpy.float64)) @ \
            X.astype(numpy.float64)
        assert numpy.max(numpy.abs(csr - ref)) <= tol * numpy.abs(ref).max()
        assert numpy.max(numpy.abs(pair - ref)) <= tol * numpy.abs(ref).max()
        # the transposed product runs on the union tiles of A^T and B^T
        Z = numpy.empty_like(X)
        _lib.check(gpu_lib.slq_op_spmm(hab, 0.37, 64, 1, _lib.ptr(X),
                                       _lib.ptr(Z)), 'slq_op_spmm')
        ref = (A.astype(numpy.float64) + 0.37 * B.astype(numpy.float64)).T @ \
            X.astype(numpy.float64)
        assert numpy.max(numpy.abs(Z - ref)) <= tol * numpy.abs(ref).max()
    finally:
        gpu_lib.slq_op_destroy(hab)
        gpu_lib.slq_op_destroy(hb)
        gpu_lib.slq_op_destroy(ha)


def test_pair_spmm_sees_new_values(gpu_lib):
    """Values set on A or B after the affine operator was made are used
    (the pair values are gathered from the CSR values at every call)."""
    n = 2000
    A = sparse_correlation_matrix(n, nnz_per_row=20, seed=6)
    B = _spd_other_pattern(n, 20, 6, 7)
    ha, hb = _csr_handle(gpu_lib, A), _csr_handle(gpu_lib, B)
    hab = ctypes.c_void_p()
    _lib.check(gpu_lib.slq_op_create_affine(ha, hb, ctypes.byref(hab)),
               'slq_op_create_affine')
    try:
        X = numpy.random.default_rng(2).standard_normal((n, 64))
        _spmm(gpu_lib, hab, 0.5, X)
        A2, B2 = A.copy(), B.copy()
        A2.data = A2.data * 1.7
        B2.data = B2.data - 0.25
        _lib.check(gpu_lib.slq_op_set_values(ha, _lib.ptr(
            numpy.ascontiguousarray(A2.data))), 'slq_op_set_values')
        _lib.check(gpu_lib.slq_op_set_values(hb, _lib.ptr(
            numpy.ascontiguousarray(B2.data))), 'slq_op_set_values')
        ref = (A2 + 0.5 * B2) @ X
        got = _spmm(gpu_lib, hab, 0.5, X)
        assert numpy.max(numpy.abs(got - ref)) <= 1e-13 * numpy.abs(ref).max()
    finally:
        gpu_lib.slq_op_destroy(hab)
        gpu_lib.slq_op_destroy(hb)
        gpu_lib.slq_op_destroy(ha)


def test_pair_spmm_repeated_columns_fall_back(gpu_lib):
    """A CSR with a repeated column in a row keeps the two-stream CSR path;
    the product is still A + tB with the repeats summed."""
    n = 1500
    A = sparse_correlation_matrix(n, nnz_per_row=20, seed=8)
    ip, ix, dv = A.indptr.copy(), A.indices.copy(), A.data.copy()
    r = 10
    k0 = ip[r]
    ix[k0 + 1] = ix[k0]           # row 10 names one column twice
    Adup = scipy.sparse.csr_matrix((dv, ix, ip), shape=A.shape)
    B = _spd_other_pattern(n, 20, 8, 9)
    ha, hb = _csr_handle(gpu_lib, Adup), _csr_handle(gpu_lib, B)
    hab = ctypes.c_void_p()
    _lib.check(gpu_lib.slq_op_create_affine(ha, hb, ctypes.byref(hab)),
               'slq_op_create_affine')
    try:
        X = numpy.random.default_rng(3).standard_normal((n, 64))
        dense_a = numpy.zeros(A.shape)
        for row in range(n):
            for k in range(ip[row], ip[row + 1]):
                dense_a[row, ix[k]] += dv[k]
        ref = (dense_a + 1.25 * B.toarray()) @ X
        got = _spmm(gpu_lib, hab, 1.25, X)
        assert numpy.max(numpy.abs(got - ref)) <= 1e-13 * numpy.abs(ref).max()
    finally:
        gpu_lib.slq_op_destroy(hab)
        gpu_lib.slq_op_destroy(hb)
        gpu_lib.slq_op_destroy(ha)


_C = None


def corr():
    global _C
    if _C is None:
        _C = sparse_correlation_matrix(4096, nnz_per_row=40, seed=0)
    return _C


@pytest.mark.parametrize('ortho', [0, -1])
@pytest.mark.parametrize('func', ['log', 'inv'])
def test_pair_slq_matches_oracle(func, ortho):
    A = corr()
    B = _spd_other_pattern(A.shape[0], 40, 0, 11)
    t = [0.02, 0.6, 3.0]
    fn = imate_amd.logdet if func == 'log' else imate_amd.traceinv
    val, info = fn(imate_amd.AffineMatrixFunction(A, B), parameters=t,
                   method='slq', gpu=True, lanczos_degree=18,
                   min_num_samples=40, max_num_samples=40, seed=4,
                   orthogonalize=ortho, return_info=True)
    ref = oracle.slq(A, func=func, parameters=t, B=B, affine=True,
                     lanczos_degree=18, min_num_samples=40,
                     max_num_samples=40, seed=4, orthogonalize=ortho)
    got = numpy.asarray(info['convergence']['samples'])
    want = ref['samples']
    assert numpy.max(numpy.abs(got - want) / numpy.abs(want)) < 1e-10
    assert numpy.allclose(val, ref['trace'], rtol=1e-10, atol=0)


def test_pair_and_csr_paths_agree(monkeypatch):
    A = corr()
    B = _spd_other_pattern(A.shape[0], 40, 0, 12)
    kw = dict(parameters=[0.1, 1.0], method='slq', gpu=True,
              lanczos_degree=20, min_num_samples=64, max_num_samples=64,
              seed=5, return_info=True)
    Aop = imate_amd.AffineMatrixFunction(A, B)
    _, pair = imate_amd.logdet(Aop, **kw)
    monkeypatch.setenv('SLQ_PAIR_TILES', '0')
    _, csr = imate_amd.logdet(Aop, **kw)
    monkeypatch.delenv('SLQ_PAIR_TILES')
    p = numpy.asarray(pair['convergence']['samples'])
    c = numpy.asarray(csr['convergence']['samples'])
    # different summation orders: equal to 1e-10, not bit for bit (which
    # shows that the two runs took different paths)
    assert numpy.max(numpy.abs(p - c) / numpy.abs(c)) < 1e-10
    assert not numpy.array_equal(p, c)


def test_pair_mixed_precision_close_to_fp64():
    A = corr()
    B = _spd_other_pattern(A.shape[0], 40, 0, 13)
    kw = dict(parameters=[0.05, 2.0], method='slq', gpu=True,
              lanczos_degree=20, min_num_samples=64, max_num_samples=64,
              seed=6)
    Aop = imate_amd.AffineMatrixFunction(A, B)
    a = numpy.asarray(imate_amd.logdet(Aop, **kw))
    b = numpy.asarray(imate_amd.logdet(Aop, precision='mixed', **kw))
    assert numpy.all(numpy.abs(a - b) <= 1e-5 * numpy.abs(a))


def test_gram_general_b_sparse_union_keeps_csr_streams():
    """Golub-Kahn on A + tB where the union is too sparse for tiles (a
    bidiagonal A): the two CSR streams and their transposes; per sample
    against the oracle."""
    from imate_amd.sample_matrices import toeplitz
    A = toeplitz(2.0, 1.0, 60, gram=False)
    B = scipy.sparse.diags(numpy.linspace(0.5, 1.5, 60)).tocsr()
    t = [0.3, 1.2]
    val, info = imate_amd.logdet(imate_amd.AffineMatrixFunction(A, B),
                                 parameters=t, gram=True, method='slq',
                                 gpu=True, lanczos_degree=20,
                                 min_num_samples=16, max_num_samples=16,
                                 seed=3, return_info=True)
    ref = oracle.slq(A, func='log', gram=True, parameters=t, B=B,
                     affine=True, lanczos_degree=20, min_num_samples=16,
                     max_num_samples=16, seed=3)
    got = numpy.asarray(info['convergence']['samples'])
    assert numpy.max(numpy.abs(got - ref['samples']) /
                     numpy.abs(ref['samples'])) < 1e-10


def test_gram_general_b_on_union_tiles(monkeypatch):
    """Golub-Kahn on A + tB: u = (A + tB) v on the union tiles of A and B,
    v = (A + tB)^T u on those of A^T and B^T; per sample against the oracle,
    and against the two CSR streams (equal to 1e-10, not bit for bit)."""
    A = sparse_correlation_matrix(2048, nnz_per_row=30, seed=3)
    B = _spd_other_pattern(2048, 30, 3, 14)
    t = [0.1, 1.5]
    kw = dict(parameters=t, gram=True, method='slq', gpu=True,
              lanczos_degree=16, min_num_samples=32, max_num_samples=32,
              seed=9, return_info=True)
    Aop = imate_amd.AffineMatrixFunction(A, B)
    _, info = imate_amd.logdet(Aop, **kw)
    ref = oracle.slq(A, func='log', gram=True, parameters=t, B=B, affine=True,
                     lanczos_degree=16, min_num_samples=32,
                     max_num_samples=32, seed=9)
    got = numpy.asarray(info['convergence']['samples'])
    assert numpy.max(numpy.abs(got - ref['samples']) /
                     numpy.abs(ref['samples'])) < 1e-10
    monkeypatch.setenv('SLQ_PAIR_TILES', '0')
    _, csr = imate_amd.logdet(Aop, **kw)
    monkeypatch.delenv('SLQ_PAIR_TILES')
    c = numpy.asarray(csr['convergence']['samples'])
    assert numpy.max(numpy.abs(got - c) / numpy.abs(c)) < 1e-10
    assert not numpy.array_equal(got, c)


@pytest.mark.parametrize('gram,precision', [(False, None), (False, 'mixed'),
                                            (True, None)])
def test_t_lanes_equal_sequential(monkeypatch, gram, precision):
    """The t of A + tB run side by side (eight lanes, so eleven t take two
    groups) give the samples of the one-after-the-other run bit for bit,
    also when some t converge before others (early stop per inquiry)."""
    A = corr()
    B = _spd_other_pattern(A.shape[0], 40, 0, 14)
    ts = list(numpy.logspace(-2, 1, 11))
    kw = dict(parameters=ts, method='slq', gpu=True, lanczos_degree=16,
              min_num_samples=20, max_num_samples=150, error_rtol=2e-3,
              seed=8, gram=gram, return_info=True)
    if precision:
        kw['precision'] = precision
    Aop = imate_amd.AffineMatrixFunction(A, B)
    monkeypatch.setenv('SLQ_CONCURRENT_T', '0')
    v0, i0 = imate_amd.logdet(Aop, **kw)
    monkeypatch.delenv('SLQ_CONCURRENT_T')
    v1, i1 = imate_amd.logdet(Aop, **kw)
    s0 = numpy.asarray(i0['convergence']['samples'])
    s1 = numpy.asarray(i1['convergence']['samples'])
    assert numpy.array_equal(s0, s1, equal_nan=True)
    assert numpy.array_equal(numpy.asarray(v0), numpy.asarray(v1))
    used = numpy.asarray(i1['convergence']['num_samples_used'])
    assert used.min() < used.max()   # some t stopped earlier than others
    if not gram and precision is None:
        ref = oracle.slq(A, func='log', parameters=ts, B=B, affine=True,
                         lanczos_degree=16, min_num_samples=20,
                         max_num_samples=150, error_rtol=2e-3, seed=8)
        want = ref['samples']
        ok = ~numpy.isnan(want)
        assert numpy.array_equal(ok, ~numpy.isnan(s1))
        assert numpy.max(numpy.abs(s1[ok] - want[ok]) /
                         numpy.abs(want[ok])) < 1e-10
