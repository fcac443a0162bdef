"""The CPU path (gpu=False, ``slq_estimate_host``) against the oracle.

The host path is product code of its own (probe blocks of 8, rows split over
threads, CGS2 re-orthogonalisation, the device's eigensolve and quadrature
code); the oracle restates imate's one-probe-per-thread CPU loop. Bars:
* FP64, per sample: 1e-10 relative (the two differ only in rounding: block
  sums in double vs the reference's long-double row sums, CGS2 vs MGS);
* FP32: 2e-5 relative to the oracle's FP32 and FP64 samples (the oracle's
  own FP32 run is within 3.4e-6 of its FP64 run on these matrices);
* float128: 1e-12 relative to the oracle's FP64 samples.
The result must not depend on the thread count (bit for bit), and dense,
CSC and CSR copies of one matrix give the same samples.
"""

import math

import numpy
import pytest
import scipy.sparse

import imate_amd
from imate_amd._trace_estimator import trace_estimator
from imate_amd.sample_matrices import sparse_correlation_matrix, sparse_gaussian
from oracle import oracle


def _corr(n=1500, k=30, seed=1, dtype='float64'):
    return sparse_correlation_matrix(n, nnz_per_row=k, seed=seed, dtype=dtype)


def _host(A, func='log', gram=False, p=1.0, parameters=None, m=20, s=24,
          smax=None, ortho=0, seed=3, threads=4, tol=None, rtol=1e-2):
    return trace_estimator(A, parameters, func, gram, p, s,
                           s if smax is None else smax, None, rtol, 0.95,
                           0.001, m, tol, ortho, seed, threads, 0, False,
                           False, False)


def _rel(got, want):
    got = numpy.asarray(got, dtype=numpy.float64).reshape(want.shape)
    scale = numpy.maximum(1.0, numpy.abs(want))
    ok = ~numpy.isnan(want)
    assert numpy.array_equal(numpy.isnan(got), ~ok)
    return float(numpy.max(numpy.abs(got[ok] - want[ok]) / scale[ok]))


FUNCS = ['log', 'inv', ('exp', [-0.5]), ('linfrac', [2.0, 1.0, 1.0, 3.0]),
         ('indicator', [0.5, 2.0]), ('gaussian', [1.0, 0.5])]


@pytest.mark.parametrize('func', FUNCS)
@pytest.mark.parametrize('ortho', [0, 4, -1])
def test_lanczos_matches_oracle(func, ortho):
    A = _corr()
    val, info = _host(A, func, ortho=ortho)
    ref = oracle.slq(A, func=func, lanczos_degree=20, min_num_samples=24,
                     max_num_samples=24, seed=3, orthogonalize=ortho)
    assert _rel(info['convergence']['samples'], ref['samples']) < 1e-10
    assert abs(val - ref['trace'][0]) <= 1e-10 * max(1.0, abs(ref['trace'][0]))


@pytest.mark.parametrize('p', [2.0, 0.5])
def test_exponent_matches_oracle(p):
    A = _corr(900)
    val, info = _host(A, 'identity', p=p, m=15)
    ref = oracle.slq(A, func='identity', exponent=p, lanczos_degree=15,
                     min_num_samples=24, max_num_samples=24, seed=3)
    assert _rel(info['convergence']['samples'], ref['samples']) < 1e-10


@pytest.mark.parametrize('ortho', [0, 3, -1])
def test_golub_kahn_non_square_matches_oracle(ortho):
    A = sparse_gaussian(700, 300, 0.05, seed=4) * 0.2
    val, info = _host(A, 'log', gram=True, m=15, ortho=ortho, seed=5)
    ref = oracle.slq(A, func='log', gram=True, lanczos_degree=15,
                     min_num_samples=24, max_num_samples=24, seed=5,
                     orthogonalize=ortho)
    assert _rel(info['convergence']['samples'], ref['samples']) < 1e-10


@pytest.mark.parametrize('gram', [False, True])
def test_affine_identity_ritz_shift_matches_oracle(gram):
    A = _corr(1200)
    ts = [0.0, 0.25, 1.0, 3.0]
    op = imate_amd.AffineMatrixFunction(A)
    val, info = _host(op, 'log', gram=gram, parameters=ts, m=20, seed=6)
    ref = oracle.slq(A, func='log', gram=gram, parameters=ts, affine=True,
                     lanczos_degree=20, min_num_samples=24,
                     max_num_samples=24, seed=6)
    assert _rel(info['convergence']['samples'], ref['samples']) < 1e-10
    assert numpy.allclose(val, ref['trace'], rtol=1e-10)


def test_affine_general_b_matches_oracle():
    A = _corr(1000)
    B = scipy.sparse.diags(numpy.linspace(0.5, 2.0, 1000)).tocsr()
    ts = [0.1, 1.0, 2.5]
    op = imate_amd.AffineMatrixFunction(A, B)
    val, info = _host(op, 'inv', parameters=ts, m=20, seed=7)
    ref = oracle.slq(A, func='inv', parameters=ts, affine=True, B=B,
                     lanczos_degree=20, min_num_samples=24,
                     max_num_samples=24, seed=7)
    assert _rel(info['convergence']['samples'], ref['samples']) < 1e-10


def test_convergence_stop_and_outputs_match_oracle():
    """min < max: the stopping sample, samples used, outliers, error and the
    NaN rows after the stop are those of the reference's loop."""
    A = _corr(1000)
    val, info = _host(A, 'log', s=10, smax=200, rtol=5e-3, seed=9)
    ref = oracle.slq(A, func='log', lanczos_degree=20, min_num_samples=10,
                     max_num_samples=200, seed=9, error_rtol=5e-3)
    conv = info['convergence']
    assert conv['num_samples_used'] == ref['num_samples_used'][0]
    assert conv['num_outliers'] == ref['num_outliers'][0]
    assert conv['converged'] == bool(ref['converged'][0])
    assert _rel(conv['samples'], ref['samples']) < 1e-10
    assert abs(info['error']['absolute_error'] - ref['error'][0]) <= \
        1e-8 * ref['error'][0]
    assert abs(val - ref['trace'][0]) <= 1e-10 * abs(ref['trace'][0])


def test_breakdown_three_eigenvalues():
    # Krylov dimension 3: every probe stops after 3 steps and the 3-node
    # Gauss rule is exact
    d = numpy.repeat([1.0, 2.0, 5.0], 40)
    A = scipy.sparse.diags(d).tocsr()
    val, info = _host(A, 'log', m=20, tol=1e-10, s=9, seed=2)
    exact = 40 * (math.log(1.0) + math.log(2.0) + math.log(5.0))
    assert numpy.allclose(info['convergence']['samples'], exact, rtol=1e-12)
    ref = oracle.slq(A, func='log', lanczos_degree=20, lanczos_tol=1e-10,
                     min_num_samples=9, max_num_samples=9, seed=2)
    assert _rel(info['convergence']['samples'], ref['samples']) < 1e-10


def test_float32_matches_oracle_fp32_and_fp64():
    A = _corr(2000, dtype='float32')
    val, info = _host(A, 'log', m=20, seed=4)
    assert info['convergence']['samples'].dtype == numpy.float32
    r32 = oracle.slq(A, func='log', lanczos_degree=20, min_num_samples=24,
                     max_num_samples=24, seed=4)
    r64 = oracle.slq(A.astype(numpy.float64), func='log', lanczos_degree=20,
                     min_num_samples=24, max_num_samples=24, seed=4)
    got = info['convergence']['samples']
    assert _rel(got, r32['samples'].astype(numpy.float64)) < 2e-5
    assert _rel(got, r64['samples']) < 2e-5


def test_float128_matches_oracle_fp64():
    A = _corr(1000).astype(numpy.longdouble)
    val, info = _host(A, 'inv', m=20, seed=4)
    assert info['matrix']['data_type'] == b'float128'
    ref = oracle.slq(A.astype(numpy.float64), func='inv', lanczos_degree=20,
                     min_num_samples=24, max_num_samples=24, seed=4)
    assert _rel(info['convergence']['samples'], ref['samples']) < 1e-12


def test_thread_count_does_not_change_the_result():
    A = _corr(9000, k=20)   # five row chunks
    out = [_host(A, 'log', m=12, s=16, ortho=-1, threads=t)[1]
           ['convergence']['samples'] for t in (1, 3, 8)]
    assert numpy.array_equal(out[0], out[1])
    assert numpy.array_equal(out[0], out[2])


def test_dense_csc_and_csr_agree():
    A = _corr(600)
    ref = _host(A, 'log', m=15)[1]['convergence']['samples']
    for M in (A.toarray(), A.tocsc(), numpy.asfortranarray(A.toarray())):
        got = _host(M, 'log', m=15)[1]['convergence']['samples']
        assert _rel(got, ref) < 1e-13
    G = sparse_gaussian(300, 120, 0.1, seed=1)
    ref = _host(G, 'log', gram=True, m=10)[1]['convergence']['samples']
    for M in (G.toarray(), G.tocsc()):
        got = _host(M, 'log', gram=True, m=10)[1]['convergence']['samples']
        assert _rel(got, ref) < 1e-13


def test_public_api_and_info_schema():
    A = _corr(500)
    v1 = imate_amd.logdet(A, method='slq', gpu=False, seed=1)
    v2, info = imate_amd.logdet(A, method='slq', gpu=False, seed=1,
                                return_info=True)
    assert v1 == v2
    assert info['device']['num_gpu_multiprocessors'] == 0
    assert info['device']['num_cpu_threads'] >= 1
    assert info['solver']['method'] == 'slq'
    s = imate_amd.schatten(A, p=2, method='slq', gpu=False, seed=1)
    assert s > 0
    M = imate_amd.Matrix(A)
    assert imate_amd.traceinv(M, method='slq', gpu=False, seed=2) == \
        imate_amd.traceinv(A, method='slq', gpu=False, seed=2)


def test_set_values_refreshes_host_copy():
    A = _corr(400)
    M = imate_amd.Matrix(A)
    imate_amd.logdet(M, method='slq', gpu=False, seed=1)
    M.set_values(A.data * 2.0)
    got = imate_amd.logdet(M, method='slq', gpu=False, seed=1)
    want = imate_amd.logdet(A * 2.0, method='slq', gpu=False, seed=1)
    assert got == want


def test_argument_errors():
    A = _corr(300)
    with pytest.raises(ValueError):
        imate_amd.logdet(A, method='slq', gpu=False, precision='mixed')
    with pytest.raises(ValueError):
        imate_amd.logdet(imate_amd.AffineMatrixFunction(A), method='slq',
                         gpu=False)
    with pytest.raises(ValueError):
        imate_amd.logdet(scipy.sparse.random(5, 4, density=0.5,
                                             format='csr'),
                         method='slq', gpu=False)
    with pytest.raises(TypeError):
        imate_amd.logdet(A.astype(numpy.int64), method='slq', gpu=False)


def test_invalid_host_csr_and_csc_rejected():
    import ctypes
    from imate_amd import _lib
    lib = _lib.load()
    data = numpy.ones(2)
    h = ctypes.c_void_p()
    cases = [([0, 1, 2], [0, 5]),      # column 5 of 2
             ([0, 1, 2], [-1, 0]),     # negative index
             ([0, 2, 1], [0, 1]),      # decreasing indptr
             ([1, 1, 2], [0, 1])]      # indptr not starting at 0
    for ptr, idx in cases:
        indptr = numpy.array(ptr, dtype=numpy.int64)
        indices = numpy.array(idx, dtype=numpy.int32)
        for fn in (lib.slq_host_op_create_csr, lib.slq_host_op_create_csc):
            r = fn(2, 2, 2, _lib.ptr(indptr), _lib.ptr(indices),
                   _lib.ptr(data), _lib.SLQ_F64, ctypes.byref(h))
            assert r == _lib.SLQ_E_ARG, (ptr, idx)
    assert 'slq_host_op_create' in _lib.last_error()
