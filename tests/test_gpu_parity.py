"""GPU parity: the HIP path against the CPU oracle on the same seeded probes."""

import numpy
import pytest
import scipy.sparse

from imate_amd import _lib
from oracle import oracle

pytestmark = pytest.mark.gpu


def _corr(n, k, seed, dtype='float64'):
    from imate_amd.sample_matrices import sparse_correlation_matrix
    return sparse_correlation_matrix(n, nnz_per_row=k, seed=seed, dtype=dtype)


@pytest.mark.parametrize('n', [1, 5, 63, 64, 65, 200, 4096])
@pytest.mark.parametrize('seed', [0, 1, 12345])
def test_rademacher_bit_exact(gpu_lib, n, seed):
    count = 7
    first = 3
    out = numpy.empty((count, n), dtype=numpy.float64)
    W = n // 64 + 1
    words = numpy.empty((count, W), dtype=numpy.uint64)
    _lib.check(gpu_lib.slq_rademacher(seed, n, first, count, _lib.SLQ_F64,
                                      _lib.ptr(out), _lib.ptr(words)),
               'slq_rademacher')
    ref = oracle.rademacher(seed, n, first + count)[first:]
    assert numpy.array_equal(out, ref)
    s = oracle.thread_state(seed, 0)
    allw = oracle.xoshiro_words(s, W * (first + count))
    assert numpy.array_equal(words.ravel(), allw[W * first:])


@pytest.mark.parametrize('n', [300000, 1 << 20])
def test_rademacher_bit_exact_many_chunks(gpu_lib, n):
    """Several 64-chunk groups per probe: every lane of the jump doubling and
    the group offsets are exercised."""
    count, first, seed = 5, 70, 7
    W = n // 64 + 1
    out = numpy.empty((count, n), dtype=numpy.float64)
    words = numpy.empty((count, W), dtype=numpy.uint64)
    _lib.check(gpu_lib.slq_rademacher(seed, n, first, count, _lib.SLQ_F64,
                                      _lib.ptr(out), _lib.ptr(words)),
               'slq_rademacher')
    s = oracle.thread_state(seed, 0)
    allw = oracle.xoshiro_words(s, W * (first + count))
    assert numpy.array_equal(words.ravel(), allw[W * first:])
    bits = numpy.unpackbits(words.view(numpy.uint8), axis=1,
                            bitorder='little')[:, :n]
    assert numpy.array_equal(out, numpy.where(bits == 1, 1.0, -1.0))


def _op(A):
    from imate_amd.linear_operator import DeviceOperator
    return DeviceOperator.from_matrix(A)


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
@pytest.mark.parametrize('transpose', [0, 1])
def test_spmm_matches_scipy(gpu_lib, dtype, transpose):
    rng = numpy.random.default_rng(0)
    A = scipy.sparse.random(300, 170, density=0.05, random_state=1,
                            format='csr', dtype=dtype)
    dev = _op(A)
    b = 64
    inl = A.shape[0] if transpose else A.shape[1]
    X = rng.standard_normal((inl, b)).astype(dtype)
    Y = numpy.empty((A.shape[1] if transpose else A.shape[0], b), dtype=dtype)
    _lib.check(gpu_lib.slq_op_spmm(dev.handle, 0.0, b, transpose,
                                   _lib.ptr(X), _lib.ptr(Y)), 'slq_op_spmm')
    ref = (A.T @ X) if transpose else (A @ X)
    tol = 1e-14 if dtype == 'float64' else 2e-6
    assert numpy.max(numpy.abs(Y - ref)) <= tol * max(1.0, numpy.abs(ref).max())


def test_recurrences_match_oracle(gpu_lib):
    A = _corr(4096, 40, 0)
    dev = _op(A)
    m, count, seed = 20, 16, 3
    a = numpy.zeros((count, m))
    b = numpy.zeros((count, m))
    sz = numpy.zeros(count, dtype=numpy.int32)
    _lib.check(gpu_lib.slq_recurrences(dev.handle, 0.0, 0, seed, 0, count, m,
                                       numpy.finfo(float).eps, 0, 0,
                                       _lib.ptr(a), _lib.ptr(b), _lib.ptr(sz)),
               'slq_recurrences')
    op = oracle.Operator(A)
    ra, rb, rs = oracle.recurrences(op, seed, count, m,
                                    numpy.finfo(float).eps, 0)
    assert numpy.array_equal(sz, rs)
    assert numpy.max(numpy.abs(a - ra)) < 1e-10
    assert numpy.max(numpy.abs(b - rb)) < 1e-10


def test_logdet_per_sample_parity_c1():
    import imate_amd
    A = _corr(4096, 40, 0)
    for seed in (0, 1):
        ld, info = imate_amd.logdet(A, method='slq', gpu=True,
                                    lanczos_degree=20, min_num_samples=100,
                                    max_num_samples=100, seed=seed,
                                    return_info=True)
        ref = oracle.slq(A, func='log', lanczos_degree=20,
                         min_num_samples=100, max_num_samples=100, seed=seed)
        got = info['convergence']['samples']
        want = ref['samples'][:, 0]
        rel = numpy.abs(got - want) / numpy.abs(want)
        assert rel.max() < 1e-10, rel.max()
        assert abs(ld - ref['trace'][0]) <= 1e-10 * abs(ref['trace'][0])


@pytest.mark.parametrize('tile_rows', ['0', '8', '16'])
@pytest.mark.parametrize('gram', [0, 1])
@pytest.mark.parametrize('ortho', [0, -1])
def test_recurrences_every_spmm_layout(gpu_lib, monkeypatch, tile_rows, gram,
                                       ortho):
    """CSR, 8-row tiles (scalar-loaded slabs) and 16-row tiles (matrix
    cores) give the oracle's recurrences, for Lanczos and Golub-Kahn, with
    and without re-orthogonalisation."""
    monkeypatch.setenv('SLQ_TILE_ROWS', tile_rows)
    A = _corr(3000, 40, 2)
    dev = _op(A)
    m, count, seed = 16, 64, 9
    a = numpy.zeros((count, m))
    b = numpy.zeros((count, m))
    sz = numpy.zeros(count, dtype=numpy.int32)
    _lib.check(gpu_lib.slq_recurrences(dev.handle, 0.0, gram, seed, 0, count,
                                       m, numpy.finfo(float).eps, ortho, 0,
                                       _lib.ptr(a), _lib.ptr(b), _lib.ptr(sz)),
               'slq_recurrences')
    op = oracle.Operator(A)
    ra, rb, rs = oracle.recurrences(op, seed, count, m,
                                    numpy.finfo(float).eps, ortho, gram=gram)
    assert numpy.array_equal(sz, rs)
    scale = max(1.0, numpy.abs(ra).max())
    assert numpy.max(numpy.abs(a - ra)) < 1e-10 * scale
    assert numpy.max(numpy.abs(b - rb)) < 1e-10 * scale


@pytest.mark.parametrize('tile_rows', ['0', '8', '16'])
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
@pytest.mark.parametrize('transpose', [0, 1])
@pytest.mark.parametrize('t', [0.0, 0.7])
def test_spmm_layouts_match_scipy(gpu_lib, monkeypatch, tile_rows, dtype,
                                  transpose, t):
    """The block SpMM through every layout (CSR, 8-row tiles, 16-row tiles on
    the matrix cores), plain and affine (A + tI), against scipy."""
    monkeypatch.setenv('SLQ_TILE_ROWS', tile_rows)
    A = _corr(2500, 40, 5, dtype=dtype)
    if t:
        from imate_amd.linear_operator import DeviceOperator
        dev = DeviceOperator.affine(_op(A))
    else:
        dev = _op(A)
    rng = numpy.random.default_rng(1)
    X = rng.standard_normal((2500, 64)).astype(dtype)
    Y = numpy.empty((2500, 64), dtype=dtype)
    _lib.check(gpu_lib.slq_op_spmm(dev.handle, t, 64, transpose, _lib.ptr(X),
                                   _lib.ptr(Y)), 'slq_op_spmm')
    A64 = A.astype(numpy.float64)
    ref = (A64.T if transpose else A64) @ X.astype(numpy.float64) + \
        t * X.astype(numpy.float64)
    tol = 1e-13 if dtype == 'float64' else 3e-6
    assert numpy.max(numpy.abs(Y - ref)) <= tol * numpy.abs(ref).max()
