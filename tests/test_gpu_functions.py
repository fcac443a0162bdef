"""trexp / trlinfrac / eigencount / density on the GPU against the oracle on
the same seeded probes (per-sample, FP64: 1e-10 relative)."""

import numpy
import pytest
import scipy.sparse

import imate_amd
from oracle import oracle

pytestmark = pytest.mark.gpu


def _corr(n=2000, k=30, seed=1, dtype='float64'):
    from imate_amd.sample_matrices import sparse_correlation_matrix
    return sparse_correlation_matrix(n, nnz_per_row=k, seed=seed, dtype=dtype)


CASES = [
    (imate_amd.trexp, dict(coeff=0.4), ('exp', [0.4])),
    (imate_amd.trexp, dict(coeff=-1.5), ('exp', [-1.5])),
    (imate_amd.trlinfrac, dict(coeff=[2.0, 1.0, 1.0, 3.0]),
     ('linfrac', [2.0, 1.0, 1.0, 3.0])),
    (imate_amd.eigencount, dict(interval=[0.5, 2.0]), ('indicator', [0.5, 2.0])),
    (imate_amd.density, dict(mu=1.0, sigma=0.5), ('gaussian', [1.0, 0.5])),
]


def _compare(got, ref):
    want = ref['samples'][:got.shape[0], :]
    scale = numpy.maximum(1.0, numpy.abs(want))
    assert numpy.max(numpy.abs(got - want) / scale) < 1e-10


@pytest.mark.parametrize('fn,kw,spec', CASES)
@pytest.mark.parametrize('p', [1.0, 2.0])
def test_functions_match_oracle(fn, kw, spec, p):
    A = _corr()
    val, info = fn(A, p=p, method='slq', gpu=True, lanczos_degree=20,
                   min_num_samples=70, max_num_samples=70, seed=3,
                   return_info=True, **kw)
    ref = oracle.slq(A, func=spec, exponent=p, lanczos_degree=20,
                     min_num_samples=70, max_num_samples=70, seed=3)
    got = numpy.asarray(info['convergence']['samples']).reshape(70, -1)
    _compare(got, ref)
    assert abs(val - ref['trace'][0]) <= 1e-10 * max(1.0, abs(ref['trace'][0]))


@pytest.mark.parametrize('fn,kw,spec', CASES[::2])
def test_functions_gram(fn, kw, spec):
    from imate_amd.sample_matrices import sparse_gaussian
    A = sparse_gaussian(600, 300, 0.05, seed=4) * 0.2
    val, info = fn(A, gram=True, method='slq', gpu=True, lanczos_degree=15,
                   min_num_samples=40, max_num_samples=40, seed=5,
                   orthogonalize=-1, return_info=True, **kw)
    ref = oracle.slq(A, func=spec, gram=True, lanczos_degree=15,
                     min_num_samples=40, max_num_samples=40, seed=5,
                     orthogonalize=-1)
    got = numpy.asarray(info['convergence']['samples']).reshape(40, -1)
    _compare(got, ref)


def test_trexp_affine_ritz_shift():
    A = _corr(1500)
    ts = [0.0, 0.1, 1.0]
    op = imate_amd.AffineMatrixFunction(A)
    val, info = imate_amd.trexp(op, coeff=-0.5, parameters=ts, method='slq',
                                gpu=True, lanczos_degree=20,
                                min_num_samples=64, max_num_samples=64, seed=6,
                                return_info=True)
    ref = oracle.slq(A, func=('exp', [-0.5]), parameters=ts, affine=True,
                     lanczos_degree=20, min_num_samples=64,
                     max_num_samples=64, seed=6)
    got = numpy.asarray(info['convergence']['samples']).reshape(64, -1)
    _compare(got, ref)


def test_eigencount_identity_exact():
    I = scipy.sparse.identity(300, format='csr') * 2.0
    c = imate_amd.eigencount(I, interval=[1.5, 2.5], method='slq', gpu=True,
                             min_num_samples=8, max_num_samples=8)
    assert c == 300.0
