"""GPU tests over the grid of imate's own SLQ tests (tests/test_logdet.py:
142-191, tests/test_traceinv.py:260-330, tests/test_trace.py:126-170): the
50 x 50 bidiagonal Toeplitz
matrix a = 2, b = 1 (its Gramian when gram is False), float32 and float64,
sparse and dense, gram True and False, exponent p in {0, 1, 2}; SLQ with
orthogonalize=-1, lanczos_degree=30, 100 to 200 samples and error_rtol=1e-2
as there (a fixed seed instead of seed=-1, so that the oracle can run on the
same probes).

Checks: per sample against the oracle (1e-10 relative for float64; for
float32 1e-4, or four times the change of the oracle's own samples under a
one-ulp perturbation of the matrix where that is larger: the float32
Golub-Kahn run of traceinv with p = 2 moves by 5e-4 under it; absolute
where the samples are zero), the number of samples used
(the convergence decision), and the estimate against the exact value the
reference tests use (the Toeplitz closed forms of toeplitz.py:355-540 for
p = 1, dense eigenvalues otherwise) within four times its 95 % error.
"""

import numpy
import pytest

import imate_amd
from imate_amd.sample_matrices import (toeplitz, toeplitz_logdet,
                                       toeplitz_trace, toeplitz_traceinv)
from oracle import oracle

pytestmark = pytest.mark.gpu

A_, B_, SIZE = 2.0, 1.0, 50
OPTS = dict(min_num_samples=100, max_num_samples=200, orthogonalize=-1,
            lanczos_degree=30, error_rtol=1e-2, seed=5)
GRID = [(dtype, gram, sparse, p)
        for dtype in ('float32', 'float64')
        for gram in (True, False)
        for sparse in (True, False)
        for p in (0, 1, 2)]


def _exact(func, K, gram, p):
    D = K.toarray().astype(numpy.float64)
    M = D.T @ D if gram else D
    lam = numpy.linalg.eigvalsh(M)
    if func == 'log':
        if p == 1:
            ld = toeplitz_logdet(A_, B_, SIZE, gram)
            return ld if gram else 2.0 * ld
        return p * float(numpy.sum(numpy.log(lam)))
    if func == 'identity':
        if p == 1 and not gram:
            return toeplitz_trace(A_, B_, SIZE, True)
        return float(numpy.sum(lam**float(p)))
    if p == 1 and not gram:
        return toeplitz_traceinv(A_, B_, SIZE, True)
    return float(numpy.sum(lam**(-float(p))))


def _sensitivity(K, func, gram, p, want):
    """Largest relative change of the oracle's samples when every stored
    value of K moves by one ulp (signs seeded)."""
    K2 = K.tocsr().copy()
    eps = numpy.finfo(K2.dtype).eps
    sign = numpy.random.default_rng(0).choice([-1.0, 1.0], K2.nnz)
    K2.data = (K2.data * (1.0 + eps * sign)).astype(K2.dtype)
    pert = oracle.slq(K2, func=func, gram=gram, exponent=p, **OPTS)
    b = numpy.asarray(pert['samples'][:, 0], dtype=numpy.float64)
    ok = ~numpy.isnan(want) & ~numpy.isnan(b) & (want != 0)
    if not numpy.any(ok):
        return 0.0
    return float(numpy.max(numpy.abs(b[ok] - want[ok]) / numpy.abs(want[ok])))


FUNCS = {'log': imate_amd.logdet, 'inv': imate_amd.traceinv,
         'identity': imate_amd.trace}


@pytest.mark.parametrize('func', ['log', 'inv', 'identity'])
@pytest.mark.parametrize('dtype,gram,sparse,p', GRID)
def test_reference_grid(func, dtype, gram, sparse, p):
    K = toeplitz(A_, B_, SIZE, gram=(not gram), dtype=dtype)
    Kin = K if sparse else K.toarray()
    fn = FUNCS[func]
    val, info = fn(Kin, method='slq', gpu=True, gram=gram, p=p,
                   return_info=True, **OPTS)
    ref = oracle.slq(K.tocsr(), func=func, gram=gram, exponent=p, **OPTS)
    got = numpy.asarray(info['convergence']['samples'], dtype=numpy.float64)
    want = numpy.asarray(ref['samples'][:, 0], dtype=numpy.float64)
    assert numpy.array_equal(numpy.isnan(got), numpy.isnan(want))
    ok = ~numpy.isnan(want)
    tol = 1e-10 if dtype == 'float64' else max(1e-4, 4.0 * _sensitivity(
        K, func, gram, p, want))
    diff = numpy.abs(got[ok] - want[ok])
    assert numpy.all(diff <= tol * numpy.maximum(numpy.abs(want[ok]), 1.0))
    assert info['convergence']['num_samples_used'] == ref['num_samples_used'][0]
    exact = _exact(func, K, gram, p)
    assert abs(val - exact) <= 4.0 * info['error']['absolute_error'] + \
        1e-5 * max(abs(exact), 1.0)
