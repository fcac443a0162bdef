"""The device-generated sparse Gaussian matrix and the chunked transpose
(the 2^20 x 2^18 Golub-Kahn workload), at sizes the oracle finishes."""

import numpy
import pytest
import scipy.sparse

import imate_amd
from imate_amd.sample_matrices import sparse_gaussian_device
from oracle import oracle

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def _mix(z):
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9 & M64
    z = (z ^ (z >> 27)) * 0x94d049bb133111eb & M64
    return z ^ (z >> 31)


def _columns(nrows, ncols, k, seed):
    """Host restatement of the generator's column hash (exact integers)."""
    cols = numpy.empty((nrows, k), dtype=numpy.int64)
    for r in range(nrows):
        for s in range(k):
            c0, c1 = s * ncols // k, (s + 1) * ncols // k
            key = (seed * 0x9e3779b97f4a7c15 + r * 0xd1b54a32d192ed03 + s) & M64
            cols[r, s] = c0 + _mix(key) % (c1 - c0)
    return cols


def test_generator_pattern_and_values(gpu_lib):
    nrows, ncols, density, seed = 400, 300, 0.1, 11
    D = sparse_gaussian_device(nrows, ncols, density, seed=seed)
    k = int(round(density * ncols))
    A = D.to_csr()
    assert A.shape == (nrows, ncols) and A.nnz == nrows * k
    assert numpy.array_equal(A.indptr, numpy.arange(nrows + 1) * k)
    assert numpy.array_equal(A.indices.reshape(nrows, k), _columns(nrows, ncols, k, seed))
    assert numpy.all(numpy.diff(A.indices.reshape(nrows, k), axis=1) > 0)
    big = sparse_gaussian_device(2000, 1000, 0.1, seed=1).to_csr().data
    assert abs(big.mean()) < 0.02 and abs(big.var() - 1.0) < 0.03
    again = sparse_gaussian_device(nrows, ncols, density, seed=seed).to_csr()
    assert (abs(again - A)).max() == 0
    other = sparse_gaussian_device(nrows, ncols, density, seed=seed + 1).to_csr()
    assert not numpy.array_equal(other.indices, A.indices)


@pytest.mark.parametrize('chunk', [None, '5000'])
def test_gram_on_device_matrix_matches_oracle(gpu_lib, monkeypatch, chunk):
    """Golub-Kahn on the device matrix (transpose built in one sort, or in
    column chunks of 5000 entries) against the oracle on its CSR copy."""
    if chunk:
        monkeypatch.setenv('SLQ_TRANSPOSE_CHUNK', chunk)
    D = sparse_gaussian_device(1500, 600, 0.05, seed=3)
    A = D.to_csr()
    val, info = imate_amd.logdet(D, gram=True, method='slq', gpu=True,
                                 lanczos_degree=20, min_num_samples=64,
                                 max_num_samples=64, seed=4, return_info=True)
    ref = oracle.slq(A, func='log', gram=True, lanczos_degree=20,
                     min_num_samples=64, max_num_samples=64, seed=4)
    got = info['convergence']['samples']
    want = ref['samples'][:, 0]
    assert numpy.max(numpy.abs(got - want) / numpy.abs(want)) < 1e-10
    # the chunked and the one-sort transposes are the same matrix
    assert abs(val - ref['trace'][0]) <= 1e-10 * abs(ref['trace'][0])


def test_chunked_transpose_spmm(gpu_lib, monkeypatch):
    monkeypatch.setenv('SLQ_TRANSPOSE_CHUNK', '777')
    from imate_amd import _lib
    D = sparse_gaussian_device(900, 350, 0.08, seed=8)
    A = D.to_csr()
    rng = numpy.random.default_rng(0)
    U = rng.standard_normal((900, 64))
    Y = numpy.empty((350, 64))
    _lib.check(gpu_lib.slq_op_spmm(D.gpu_Aop.handle, 0.0, 64, 1, _lib.ptr(U),
                                   _lib.ptr(Y)), 'slq_op_spmm')
    ref = A.T @ U
    assert numpy.max(numpy.abs(Y - ref)) <= 1e-13 * numpy.abs(ref).max()


def test_blocked_spmm_matches_scipy(gpu_lib):
    """Rows of ~280 non-zeros over 70,000 columns (a 36 MB operand): the
    column-blocked CSR kernel for A X, the row kernel for A^T U; and the
    same matrix from host arrays, sorted (blocked) and with rows reversed
    (unsorted: the row kernel)."""
    import ctypes
    from imate_amd import _lib
    D = sparse_gaussian_device(2048, 70000, 0.004, seed=11)
    A = D.to_csr()
    rng = numpy.random.default_rng(1)
    X = rng.standard_normal((70000, 64))
    Y = numpy.empty((2048, 64))
    _lib.check(gpu_lib.slq_op_spmm(D.gpu_Aop.handle, 0.0, 64, 0, _lib.ptr(X),
                                   _lib.ptr(Y)), 'slq_op_spmm')
    ref = A @ X
    assert numpy.max(numpy.abs(Y - ref)) <= 1e-13 * numpy.abs(ref).max()
    U = rng.standard_normal((2048, 64))
    Z = numpy.empty((70000, 64))
    _lib.check(gpu_lib.slq_op_spmm(D.gpu_Aop.handle, 0.0, 64, 1, _lib.ptr(U),
                                   _lib.ptr(Z)), 'slq_op_spmm')
    refz = A.T @ U
    assert numpy.max(numpy.abs(Z - refz)) <= 1e-13 * numpy.abs(refz).max()
    indptr = A.indptr.astype(numpy.int64)
    rev_ix = A.indices.astype(numpy.int32).copy()
    rev_v = A.data.copy()
    for r in range(A.shape[0]):
        a, b = indptr[r], indptr[r + 1]
        rev_ix[a:b] = rev_ix[a:b][::-1].copy()
        rev_v[a:b] = rev_v[a:b][::-1].copy()
    for ix, v in ((A.indices.astype(numpy.int32), A.data), (rev_ix, rev_v)):
        h = ctypes.c_void_p()
        _lib.check(gpu_lib.slq_op_create_csr(2048, 70000, A.nnz, _lib.ptr(indptr),
                                             _lib.ptr(ix), _lib.ptr(v),
                                             _lib.SLQ_F64, ctypes.byref(h)),
                   'slq_op_create_csr')
        try:
            Y2 = numpy.empty((2048, 64))
            _lib.check(gpu_lib.slq_op_spmm(h, 0.0, 64, 0, _lib.ptr(X),
                                           _lib.ptr(Y2)), 'slq_op_spmm')
        finally:
            gpu_lib.slq_op_destroy(h)
        assert numpy.max(numpy.abs(Y2 - ref)) <= 1e-13 * numpy.abs(ref).max()


def test_gram_blocked_transpose_matches_oracle(gpu_lib):
    """Tall A (70,000 x 2048, ~8 non-zeros per row): Golub-Kahn's A^T u
    gathers a 36 MB operand through the column-blocked kernel."""
    D = sparse_gaussian_device(70000, 2048, 0.004, seed=12)
    A = D.to_csr()
    val, info = imate_amd.logdet(D, gram=True, method='slq', gpu=True,
                                 lanczos_degree=20, min_num_samples=64,
                                 max_num_samples=64, seed=5, return_info=True)
    ref = oracle.slq(A, func='log', gram=True, lanczos_degree=20,
                     min_num_samples=64, max_num_samples=64, seed=5)
    got = info['convergence']['samples']
    want = ref['samples'][:, 0]
    assert numpy.max(numpy.abs(got - want) / numpy.abs(want)) < 1e-10


@pytest.mark.parametrize('nrows', [2051, 20001])
def test_blocked_spmm_ragged_rows(gpu_lib, nrows):
    """The column-blocked kernel with a row count that is not a multiple of
    its eight-row groups, and (20001 rows) more groups than one round of
    the grid, so the last round has empty groups that still pass every
    block barrier."""
    from imate_amd import _lib
    D = sparse_gaussian_device(nrows, 70000, 0.004, seed=13)
    A = D.to_csr()
    X = numpy.random.default_rng(2).standard_normal((70000, 64))
    Y = numpy.empty((nrows, 64))
    _lib.check(gpu_lib.slq_op_spmm(D.gpu_Aop.handle, 0.0, 64, 0, _lib.ptr(X),
                                   _lib.ptr(Y)), 'slq_op_spmm')
    ref = A @ X
    assert numpy.max(numpy.abs(Y - ref)) <= 1e-13 * numpy.abs(ref).max()


def test_unsorted_csr_transpose_with_chunking_set(gpu_lib, monkeypatch):
    """A CSR given through the C ABI with every row's columns reversed
    (unsorted) while the chunked transpose is requested: the chunked path
    needs sorted rows, so the one-sort path must build A^T, and the
    transposed product and Golub-Kahn must equal those of the sorted
    matrix."""
    import ctypes
    from imate_amd import _lib
    monkeypatch.setenv('SLQ_TRANSPOSE_CHUNK', '777')
    # singular values near 2: the recurrences move by ~1e-15 under a
    # relative perturbation of 1e-15 of the values (checked on the oracle),
    # so a different summation order stays inside the 1e-10 bar
    G = scipy.sparse.random(700, 300, density=0.06, random_state=21,
                            format='csr',
                            data_rvs=numpy.random.default_rng(1).standard_normal)
    A = (2.0 * scipy.sparse.eye(700, 300, format='csr') + 0.3 * G).tocsr()
    A.sort_indices()
    indptr = A.indptr.astype(numpy.int64)
    ix = A.indices.astype(numpy.int32).copy()
    v = A.data.copy()
    for r in range(A.shape[0]):
        a, b = indptr[r], indptr[r + 1]
        ix[a:b] = ix[a:b][::-1].copy()
        v[a:b] = v[a:b][::-1].copy()
    h = ctypes.c_void_p()
    _lib.check(gpu_lib.slq_op_create_csr(700, 300, A.nnz, _lib.ptr(indptr),
                                         _lib.ptr(ix), _lib.ptr(v),
                                         _lib.SLQ_F64, ctypes.byref(h)),
               'slq_op_create_csr')
    try:
        U = numpy.random.default_rng(3).standard_normal((700, 64))
        Z = numpy.empty((300, 64))
        _lib.check(gpu_lib.slq_op_spmm(h, 0.0, 64, 1, _lib.ptr(U),
                                       _lib.ptr(Z)), 'slq_op_spmm')
        refz = A.T @ U
        assert numpy.max(numpy.abs(Z - refz)) <= 1e-13 * numpy.abs(refz).max()
        m, s = 12, 40
        alpha = numpy.zeros((s, m))
        beta = numpy.zeros((s, m))
        sizes = numpy.zeros(s, dtype=numpy.int32)
        _lib.check(gpu_lib.slq_recurrences(h, 0.0, 1, 6, 0, s, m,
                                           float(numpy.finfo(float).eps), 0,
                                           _lib.PREC_NATIVE, _lib.ptr(alpha),
                                           _lib.ptr(beta), _lib.ptr(sizes)),
                   'slq_recurrences')
    finally:
        gpu_lib.slq_op_destroy(h)
    op = oracle.Operator(A)
    a2, b2, sz2 = oracle.recurrences(op, 6, s, m,
                                     float(numpy.finfo(float).eps), 0,
                                     gram=True)
    assert numpy.array_equal(sizes, sz2)
    for k in range(s):
        n_k = sizes[k]
        assert numpy.allclose(alpha[k, :n_k], a2[k, :n_k], rtol=1e-10, atol=0)
        assert numpy.allclose(beta[k, :n_k], b2[k, :n_k], rtol=1e-10, atol=0)
