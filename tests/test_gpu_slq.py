"""GPU tests of the drop-in through the C ABI: every SLQ form against the CPU
oracle on the same seeded probes (per-sample values first, then the
convergence / outlier outputs), plus closed forms and size-independent
properties at larger sizes.

Tolerances: 1e-10 relative per sample for the FP64 forms (the parity bar;
GPU sums run in a different order than the reference's long-double sums);
re-orthogonalised runs use classical Gram-Schmidt twice on the GPU against
the reference's modified Gram-Schmidt, also held to 1e-10; FP32 matrices
are held to 1e-4 and the mixed-precision form to 1e-5 of the FP64 result.
"""

import math

import numpy
import pytest
import scipy.sparse

import imate_amd
from imate_amd.sample_matrices import (sparse_correlation_matrix, toeplitz,
                                       toeplitz_logdet, toeplitz_traceinv)
from oracle import oracle

pytestmark = pytest.mark.gpu

FUNCS = {'log': imate_amd.logdet, 'inv': imate_amd.traceinv,
         'identity': imate_amd.trace}


def _both(A, func='log', gram=False, p=1.0, m=20, s=40, seed=3, ortho=0,
          params=None, B=None, min_s=None, rtol=1e-2, atol=None,
          outlier=0.001, **extra):
    min_s = s if min_s is None else min_s
    if params is not None:
        Aop = imate_amd.AffineMatrixFunction(A, B)
    else:
        Aop = A
    val, info = FUNCS[func](Aop, gram=gram, p=p, method='slq', gpu=True,
                            lanczos_degree=m, min_num_samples=min_s,
                            max_num_samples=s, seed=seed,
                            orthogonalize=ortho, parameters=params,
                            error_rtol=rtol, error_atol=atol,
                            outlier_significance_level=outlier,
                            return_info=True, **extra)
    ref = oracle.slq(A, func=func, gram=gram, exponent=p, lanczos_degree=m,
                     min_num_samples=min_s, max_num_samples=s, seed=seed,
                     orthogonalize=ortho, parameters=params, B=B,
                     affine=params is not None, error_rtol=rtol,
                     error_atol=0.0 if atol is None else atol,
                     outlier_significance_level=outlier)
    return val, info, ref


def _assert_samples(info, ref, tol):
    got = numpy.atleast_2d(numpy.asarray(info['convergence']['samples']).T).T
    want = ref['samples']
    assert got.shape == want.shape
    assert numpy.array_equal(numpy.isnan(got), numpy.isnan(want))
    ok = ~numpy.isnan(want)
    rel = numpy.abs(got[ok] - want[ok]) / numpy.maximum(numpy.abs(want[ok]),
                                                         1e-300)
    assert rel.max() < tol, rel.max()


C1 = None


def c1():
    global C1
    if C1 is None:
        C1 = sparse_correlation_matrix(4096, nnz_per_row=40, seed=0)
    return C1


@pytest.mark.parametrize('func', ['log', 'inv', 'identity'])
@pytest.mark.parametrize('seed', [0, 2])
def test_functions_per_sample(func, seed):
    val, info, ref = _both(c1(), func=func, m=20, s=100, seed=seed)
    _assert_samples(info, ref, 1e-10)
    assert abs(val - ref['trace'][0]) <= 1e-10 * abs(ref['trace'][0])


@pytest.mark.parametrize('p', [2.0, 0.5])
def test_exponent(p):
    val, info, ref = _both(c1(), func='identity', p=p, m=20, s=32)
    _assert_samples(info, ref, 1e-10)


@pytest.mark.parametrize('ortho', [-1, 5, 20])
def test_reorthogonalization(ortho):
    val, info, ref = _both(c1(), m=20, s=32, ortho=ortho)
    _assert_samples(info, ref, 1e-10)


def test_affine_identity_shift_many_t():
    t = [0.0, 1e-3, 0.05, 0.7, 3.0]
    val, info, ref = _both(c1(), m=20, s=40, params=t)
    _assert_samples(info, ref, 1e-10)
    assert numpy.allclose(val, ref['trace'], rtol=1e-10)
    assert list(info['convergence']['num_samples_used']) == \
        list(ref['num_samples_used'])


def test_affine_general_b():
    A = c1()
    B = scipy.sparse.diags(numpy.linspace(0.5, 1.5, A.shape[0])).tocsr()
    t = [0.1, 0.4]
    val, info, ref = _both(A, m=16, s=24, params=t, B=B)
    _assert_samples(info, ref, 1e-10)


@pytest.mark.parametrize('ortho', [0, -1, 4])
def test_gram_golub_kahn_toeplitz(ortho):
    A = toeplitz(2.0, 1.0, 50, gram=False)
    val, info, ref = _both(A, gram=True, m=30, s=50, ortho=ortho)
    _assert_samples(info, ref, 1e-10)


def _rect():
    A = scipy.sparse.random(400, 150, density=0.05, random_state=2,
                            format='csr')
    return (A + scipy.sparse.eye(400, 150) * 0.5).tocsr()


def test_gram_golub_kahn_rectangular_reorthogonalized():
    val, info, ref = _both(_rect(), gram=True, m=20, s=24, ortho=-1)
    _assert_samples(info, ref, 1e-10)


def test_gram_golub_kahn_rectangular_rounding_bound():
    """Without re-orthogonalisation this recurrence loses orthogonality
    within 20 steps and amplifies rounding: a one-ulp perturbation of the
    matrix moves the oracle's own per-sample values by ~1e-9. The GPU must
    stay within a small multiple of that intrinsic sensitivity."""
    A = _rect()
    val, info, ref = _both(A, gram=True, m=20, s=24)
    B = A.copy()
    B.data = B.data * (1 + 2e-16 * numpy.random.default_rng(0).standard_normal(B.nnz))
    pert = oracle.slq(B, func='log', gram=True, lanczos_degree=20,
                      min_num_samples=24, max_num_samples=24, seed=3)
    sens = numpy.max(numpy.abs(pert['samples'] - ref['samples']) /
                     numpy.abs(ref['samples']))
    _assert_samples(info, ref, max(1e-10, 10 * sens))


def test_gram_affine_shift():
    A = toeplitz(2.0, 1.0, 60, gram=False)
    val, info, ref = _both(A, gram=True, m=20, s=16, params=[0.5])
    _assert_samples(info, ref, 1e-10)


def test_convergence_stop_and_outliers_match():
    # positive log-determinant so that the relative tolerance can fire; the
    # oracle stops at sample 100 (in the second 64-probe block)
    A = (c1() * 8.0).tocsr()
    val, info, ref = _both(A, m=15, s=200, min_s=10, rtol=2e-3, outlier=0.01)
    assert info['convergence']['num_samples_used'] == \
        ref['num_samples_used'][0]
    assert info['convergence']['num_samples_used'] < 200
    assert info['convergence']['converged'] == bool(ref['converged'][0])
    assert info['convergence']['num_outliers'] == ref['num_outliers'][0]
    _assert_samples(info, ref, 1e-10)
    assert abs(info['error']['absolute_error'] - ref['error'][0]) <= \
        1e-8 * ref['error'][0]


def test_samples_beyond_block_boundaries():
    # 150 samples = two full 64-probe blocks and a partial one
    val, info, ref = _both(c1(), m=10, s=150)
    _assert_samples(info, ref, 1e-10)
    assert list(info['convergence']['samples_processed_order']) == \
        list(range(150))


@pytest.mark.parametrize('n', [1, 2, 63, 64, 65, 130])
def test_small_and_ragged_sizes(n):
    A = sparse_correlation_matrix(max(n, 2), nnz_per_row=5, seed=9)
    if n == 1:
        A = scipy.sparse.csr_matrix(numpy.array([[2.5]]))
    val, info, ref = _both(A, m=min(8, n), s=10)
    _assert_samples(info, ref, 1e-10)


def test_identity_breakdown():
    n = 500
    A = scipy.sparse.identity(n, format='csr') * 2.0
    val, info, ref = _both(A, m=10, s=8)
    assert numpy.allclose(info['convergence']['samples'], n * math.log(2.0),
                          rtol=1e-14)
    _assert_samples(info, ref, 1e-12)


def test_three_eigenvalue_breakdown_exact():
    d = numpy.repeat([1.0, 2.0, 5.0], 40)
    A = scipy.sparse.diags(d).tocsr()
    val, info, ref = _both(A, m=20, s=6, lanczos_tol=1e-10)
    exact = 40 * (math.log(1.0) + math.log(2.0) + math.log(5.0))
    assert numpy.allclose(info['convergence']['samples'], exact, rtol=1e-12)


def test_toeplitz_closed_forms_gpu():
    A = toeplitz(2.0, 1.0, 100, gram=True)
    ld, info = imate_amd.logdet(A, method='slq', gpu=True, lanczos_degree=30,
                                min_num_samples=64, max_num_samples=256,
                                orthogonalize=-1, seed=1, return_info=True)
    exact = toeplitz_logdet(2.0, 1.0, 100, gram=True)
    assert abs(ld - exact) < 3 * info['error']['absolute_error'] + 1e-9
    ti, info = imate_amd.traceinv(A, method='slq', gpu=True,
                                  lanczos_degree=30, min_num_samples=64,
                                  max_num_samples=256, orthogonalize=-1,
                                  seed=2, return_info=True)
    exact = toeplitz_traceinv(2.0, 1.0, 100, gram=True)
    assert abs(ti - exact) < 3 * info['error']['absolute_error'] + 1e-9


def test_schatten():
    A = c1()
    kw = dict(method='slq', gpu=True, lanczos_degree=20, min_num_samples=32,
              max_num_samples=32, seed=4)
    n = A.shape[0]
    s0 = imate_amd.schatten(A, p=0, **kw)
    ld = imate_amd.logdet(A, **kw)
    assert abs(s0 - math.exp(ld / n)) < 1e-12 * s0
    s2 = imate_amd.schatten(A, p=2, **kw)
    tr2 = imate_amd.trace(A, p=2, **kw)
    assert abs(s2 - math.sqrt(tr2 / n)) < 1e-12 * s2
    sm1 = imate_amd.schatten(A, p=-1, **kw)
    ti = imate_amd.traceinv(A, p=1, **kw)
    assert abs(sm1 - (ti / n)**(-1.0)) < 1e-12 * sm1


def test_float32_matrix():
    A = c1().astype(numpy.float32)
    val, info, ref = _both(A, m=20, s=32)
    _assert_samples(info, ref, 1e-4)


def test_mixed_precision_close_to_fp64():
    A = c1()
    kw = dict(method='slq', gpu=True, lanczos_degree=20, min_num_samples=64,
              max_num_samples=64, seed=5, return_info=True)
    v64, i64 = imate_amd.logdet(A, **kw)
    v32, i32 = imate_amd.logdet(A, precision='mixed', **kw)
    rel = numpy.abs(i32['convergence']['samples'] -
                    i64['convergence']['samples']) / \
        numpy.abs(i64['convergence']['samples'])
    assert rel.max() < 1e-5


def test_matrix_object_reuse_and_determinism():
    Aop = imate_amd.Matrix(c1())
    kw = dict(method='slq', gpu=True, lanczos_degree=20, min_num_samples=64,
              max_num_samples=64, seed=6, return_info=True)
    v1, i1 = imate_amd.logdet(Aop, **kw)
    v2, i2 = imate_amd.logdet(Aop, **kw)
    assert numpy.array_equal(i1['convergence']['samples'],
                             i2['convergence']['samples'])
    assert v1 == v2


def test_info_schema_and_verbose(capsys):
    val, info = imate_amd.logdet(c1(), method='slq', gpu=True,
                                 lanczos_degree=10, min_num_samples=8,
                                 max_num_samples=8, seed=1, verbose=True,
                                 return_info=True)
    out = capsys.readouterr().out
    assert 'results' in out and 'lanczos degree' in out
    assert set(info) == {'matrix', 'error', 'convergence', 'time', 'device',
                         'solver'}
    assert info['solver']['method'] == 'slq'
    assert info['matrix']['data_type'] == b'float64'
    assert info['device']['num_gpu_devices'] == 1
    assert info['time']['alg_wall_time'] > 0


def test_cholesky_accuracy_c1():
    A = c1()
    L = numpy.linalg.cholesky(A.toarray())
    exact = 2.0 * numpy.sum(numpy.log(numpy.diag(L)))
    ld, info = imate_amd.logdet(A, method='slq', gpu=True, lanczos_degree=20,
                                min_num_samples=256, max_num_samples=256,
                                seed=0, return_info=True)
    assert abs(ld - exact) <= 0.01 * abs(exact)
    assert abs(ld - exact) <= 3 * info['error']['absolute_error']


@pytest.mark.parametrize('m', [50, 70])
def test_full_reorthogonalization_long_basis(m):
    """Full re-orthogonalisation past 32 stored vectors: steps with K <= 64
    take the fused subtraction + product pass (64-vector form above 32), and
    m = 70 also runs the two-pass form beyond 64."""
    val, info, ref = _both(c1(), m=m, s=24, ortho=-1)
    _assert_samples(info, ref, 1e-10)


def test_gram_full_reorthogonalization_long_basis():
    A = _rect()
    val, info, ref = _both(A, gram=True, m=45, s=24, ortho=-1)
    _assert_samples(info, ref, 1e-10)


def test_fused_gram_schmidt_matches_two_passes(monkeypatch):
    kw = dict(method='slq', gpu=True, lanczos_degree=60, min_num_samples=64,
              max_num_samples=64, seed=6, orthogonalize=-1, return_info=True)
    _, fused = imate_amd.logdet(c1(), **kw)
    monkeypatch.setenv('SLQ_FUSED_GS', '0')
    _, plain = imate_amd.logdet(c1(), **kw)
    monkeypatch.delenv('SLQ_FUSED_GS')
    a = numpy.asarray(fused['convergence']['samples'])
    b = numpy.asarray(plain['convergence']['samples'])
    assert numpy.max(numpy.abs(a - b) / numpy.abs(b)) < 1e-10
    _, mixed = imate_amd.logdet(c1(), precision='mixed', **kw)
    c = numpy.asarray(mixed['convergence']['samples'])
    assert abs(c.mean() - b.mean()) <= 1e-5 * abs(b.mean())
