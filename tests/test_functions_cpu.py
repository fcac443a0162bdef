"""The matrix functions of trexp / trlinfrac / eigencount / density (CPU):
the oracle's quadrature pinned by Lanczos exactness, and the argument checks
of the drop-in wrappers."""

import numpy
import pytest
import scipy.sparse

import imate_amd
from oracle import oracle

FUNCS = [
    ('log', None),
    ('inv', None),
    ('identity', None),
    (('exp', [0.3]), lambda x: numpy.exp(0.3 * x)),
    (('linfrac', [2.0, 1.0, 1.0, 3.0]), lambda x: (2 * x + 1) / (x + 3)),
    (('indicator', [2.5, 6.5]), lambda x: ((x >= 2.5) & (x <= 6.5)) * 1.0),
    (('gaussian', [4.0, 1.5]),
     lambda x: numpy.exp(-0.5 * ((x - 4.0) / 1.5) ** 2) / (1.5 * numpy.sqrt(2 * numpy.pi))),
]
SIMPLE = {'log': numpy.log, 'inv': lambda x: 1.0 / x, 'identity': lambda x: x}


def _matrix(n=40, seed=0):
    # symmetric with known, well separated eigenvalues in [1, 9]
    rng = numpy.random.default_rng(seed)
    lam = numpy.linspace(1.0, 9.0, n) + 0.01 * rng.random(n)
    Q, _ = numpy.linalg.qr(rng.standard_normal((n, n)))
    return (Q * lam) @ Q.T, lam, Q


@pytest.mark.parametrize('func,f', FUNCS)
@pytest.mark.parametrize('p', [1.0, 2.0])
def test_quadrature_exact_with_full_krylov(func, f, p):
    """A Lanczos run to completion makes Gauss quadrature exact for every f:
    each sample equals n v^T f(A^p) v / v^T v."""
    A, lam, Q = _matrix()
    n = A.shape[0]
    if f is None:
        f = SIMPLE[func]
    Mf = (Q * f(lam ** p)) @ Q.T
    S = scipy.sparse.csr_matrix(A)
    ref = oracle.slq(S, func=func, exponent=p, lanczos_degree=n,
                     lanczos_tol=1e-14, orthogonalize=-1, min_num_samples=6,
                     max_num_samples=6, seed=4)
    V = oracle.rademacher(4, n, 6)
    want = numpy.array([n * v @ Mf @ v / (v @ v) for v in V])
    got = ref['samples'][:, 0]
    scale = max(1.0, numpy.abs(want).max())
    assert numpy.abs(got - want).max() <= 1e-9 * scale


def test_quadrature_exact_gram():
    """Golub-Kahn to completion: samples of f(A^T A) for a rectangular A."""
    rng = numpy.random.default_rng(3)
    A = rng.standard_normal((50, 30))
    s = numpy.linalg.svd(A, compute_uv=False)
    _, _, Vt = numpy.linalg.svd(A, full_matrices=False)
    g = lambda x: numpy.exp(-0.5 * ((x - 40.0) / 9.0) ** 2) / (9.0 * numpy.sqrt(2 * numpy.pi))
    Mf = (Vt.T * g(s ** 2)) @ Vt
    ref = oracle.slq(scipy.sparse.csr_matrix(A), func=('gaussian', [40.0, 9.0]),
                     gram=True, lanczos_degree=30, lanczos_tol=1e-14,
                     orthogonalize=-1, min_num_samples=4, max_num_samples=4,
                     seed=2)
    V = oracle.rademacher(2, 30, 4)
    want = numpy.array([30 * v @ Mf @ v / (v @ v) for v in V])
    assert numpy.abs(ref['samples'][:, 0] - want).max() <= 1e-9 * max(1.0, abs(want).max())


def test_indicator_counts_eigenvalues():
    """eigencount of the identity: all n eigenvalues (1) in [0.5, 1.5], none
    in [2, 3] (the run stops after one step with theta = 1)."""
    I = scipy.sparse.identity(64, format='csr')
    a = oracle.slq(I, func=('indicator', [0.5, 1.5]), min_num_samples=3,
                   max_num_samples=3)
    b = oracle.slq(I, func=('indicator', [2.0, 3.0]), min_num_samples=3,
                   max_num_samples=3)
    assert numpy.all(a['samples'][:, 0] == 64.0)
    assert numpy.all(b['samples'][:, 0] == 0.0)


def test_exp_of_scaled_identity():
    c = 0.7
    I = scipy.sparse.identity(30, format='csr') * 2.0
    r = oracle.slq(I, func=('exp', [c]), min_num_samples=2, max_num_samples=2)
    assert numpy.allclose(r['samples'][:, 0], 30 * numpy.exp(2.0 * c), rtol=1e-15)


@pytest.mark.parametrize('call,exc', [
    (lambda A: imate_amd.trlinfrac(A, coeff=[1, 2, 2, 4], method='slq'), ValueError),
    (lambda A: imate_amd.trlinfrac(A, coeff=[1, 2, 3], method='slq'), ValueError),
    (lambda A: imate_amd.trlinfrac(A, coeff=1.0, method='slq'), TypeError),
    (lambda A: imate_amd.eigencount(A, interval=[0, 1, 2], method='slq'), ValueError),
    (lambda A: imate_amd.eigencount(A, interval=3.0, method='slq'), TypeError),
    (lambda A: imate_amd.density(A, sigma=0.0, method='slq'), ValueError),
    (lambda A: imate_amd.density(A, sigma='a', method='slq'), TypeError),
    (lambda A: imate_amd.density(A, mu='a', method='slq'), TypeError),
    (lambda A: imate_amd.density(A, mu=[1.0, 2.0], method='slq'), TypeError),
    (lambda A: imate_amd.trexp(A, return_info=1, method='slq'), TypeError),
    (lambda A: imate_amd.trexp(A), NotImplementedError),
    (lambda A: imate_amd.density(A, method='eigenvalue'), NotImplementedError),
    (lambda A: imate_amd.eigencount(A, method='cholesky'), ValueError),
])
def test_argument_checks(call, exc):
    A = scipy.sparse.identity(10, format='csr')
    with pytest.raises(exc):
        call(A)
