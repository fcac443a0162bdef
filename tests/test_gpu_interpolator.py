"""GPU tests of the interpolators: the points are estimated through the C ABI
(one call per batch of points) and compared with the oracle's SLQ on the same
seeded probes; the fitted interpolants are compared with the same fits of the
oracle's values, and with per-point ``schatten(A + t_i B)`` calls (imate's own
call pattern).

Tolerances: the points, like every FP64 SLQ output, to 1e-10 relative; the
interpolants to 1e-8 (the host fits' linear solves amplify the 1e-12-level
differences of the points by their condition numbers); per-point calls on
explicit A + t_i B to 1e-8 (a different rounding of the same operator).
"""

import numpy
import pytest
import scipy.linalg
import scipy.sparse

import imate_amd
from imate_amd.interpolator import (InterpolateSchatten, InterpolateLogdet,
                                    InterpolateTrace, _interpolant_base)
from imate_amd.sample_matrices import sparse_correlation_matrix
from oracle import oracle

pytestmark = pytest.mark.gpu

M, S, SEED = 20, 24, 5
OPTIONS = dict(method='slq', gpu=True, lanczos_degree=M, min_num_samples=S,
               max_num_samples=S, seed=SEED)
TQ = numpy.geomspace(2e-3, 20.0, 13)

_A = None


def matrix():
    global _A
    if _A is None:
        _A = sparse_correlation_matrix(2048, nnz_per_row=30, seed=1)
    return _A


def diag_b(n):
    return scipy.sparse.diags(numpy.linspace(0.5, 1.5, n)).tocsr()


def _oracle_schatten(A, p, t=None, B=None):
    if p == 0:
        func, e = 'log', 1.0
    elif p > 0:
        func, e = 'identity', float(p)
    else:
        func, e = 'inv', float(-p)
    r = oracle.slq(A, func=func, exponent=e, parameters=t, B=B,
                   affine=t is not None, lanczos_degree=M, min_num_samples=S,
                   max_num_samples=S, seed=SEED)
    tr = numpy.asarray(r['trace'], dtype=float)
    n = A.shape[0]
    return numpy.exp(tr / n) if p == 0 else (tr / n)**(1.0 / p)


@pytest.fixture
def oracle_values(monkeypatch):
    """Make new interpolators take their points from the oracle."""
    def schatten_at(self, t):
        t = numpy.atleast_1d(numpy.asarray(t, dtype=float))
        B = None if self.B_is_identity else self.B
        return _oracle_schatten(self.A, self.p, t, B)

    def compute_schatten(self, Mx, p):
        return float(_oracle_schatten(Mx, p)[0])

    def use():
        monkeypatch.setattr(_interpolant_base.InterpolantBase, '_schatten_at',
                            schatten_at)
        monkeypatch.setattr(_interpolant_base.InterpolantBase,
                            '_compute_schatten', compute_schatten)
    return use


CASES = [('imbf', {}, [1e-3, 1e-2, 1e-1, 1.0, 10.0]),
         ('imbf', dict(basis_func_type='non-ortho'), [1e-2, 1e-1, 1.0, 10.0]),
         ('rbf', dict(func_type=2), [1e-3, 1e-2, 1e-1, 1.0, 10.0]),
         ('rpf', {}, [1e-2, 1e-1, 1.0, 10.0]),
         ('crf', dict(func_type=1), [1e-2, 1e-1, 1.0, 10.0]),
         ('spl', dict(func_type=1), [1e-2, 1e-1, 1.0, 10.0]),
         ('mbf', {}, [0.5])]


@pytest.mark.parametrize('kind,kw,ti', CASES)
@pytest.mark.parametrize('p', [0, -1, 2])
def test_points_and_fit_match_oracle_identity_b(oracle_values, kind, kw, ti,
                                                p):
    A = matrix()
    gpu = InterpolateSchatten(A, p=p, kind=kind, ti=ti, options=OPTIONS, **kw)
    oracle_values()
    ref = InterpolateSchatten(A, p=p, kind=kind, ti=ti, options=OPTIONS, **kw)
    g, r = gpu.interpolator, ref.interpolator
    assert abs(g.schatten_A - r.schatten_A) <= 1e-10 * abs(r.schatten_A)
    assert numpy.allclose(g.schatten_i, r.schatten_i, rtol=1e-10, atol=0)
    assert numpy.allclose(gpu.interpolate(TQ), ref.interpolate(TQ),
                          rtol=1e-8, atol=0)


@pytest.mark.parametrize('kind,kw,ti', CASES[:4])
def test_points_and_fit_match_oracle_general_b(oracle_values, kind, kw, ti):
    A = matrix()
    B = diag_b(A.shape[0])
    gpu = InterpolateLogdet(A, B=B, kind=kind, ti=ti, options=OPTIONS, **kw)
    oracle_values()
    ref = InterpolateLogdet(A, B=B, kind=kind, ti=ti, options=OPTIONS, **kw)
    g, r = gpu.interpolator, ref.interpolator
    assert abs(g.schatten_B - r.schatten_B) <= 1e-10 * abs(r.schatten_B)
    assert numpy.allclose(g.schatten_i, r.schatten_i, rtol=1e-10, atol=0)
    assert numpy.allclose(gpu.interpolate(TQ), ref.interpolate(TQ),
                          rtol=1e-8, atol=0)


@pytest.mark.parametrize('with_b', [False, True])
def test_points_match_per_point_calls(with_b):
    """imate forms A + t_i B and calls schatten once per point."""
    A = matrix()
    n = A.shape[0]
    B = diag_b(n) if with_b else None
    ti = [1e-2, 0.3, 4.0]
    L = InterpolateLogdet(A, B=B, kind='imbf', ti=ti, options=OPTIONS)
    Bm = scipy.sparse.eye(n, format='csr') if B is None else B
    for tk, got in zip(ti, L.interpolator.schatten_i):
        want = imate_amd.schatten((A + tk * Bm).tocsr(), p=0, **OPTIONS)
        assert abs(got - want) <= 1e-8 * abs(want)
    want0 = imate_amd.schatten(A, p=0, **OPTIONS)
    assert abs(L.interpolator.schatten_A - want0) <= 1e-8 * abs(want0)


def test_ext_and_eval_one_call_per_batch():
    A = matrix()
    E = InterpolateTrace(A, p=-1, kind='ext', options=OPTIONS)
    t = numpy.array([0.01, 0.1, 1.0])
    got = E.interpolate(t)
    want = numpy.array([imate_amd.traceinv(imate_amd.AffineMatrixFunction(A),
                                           parameters=tk, **OPTIONS)
                        for tk in t])
    assert numpy.allclose(got, want, rtol=1e-10, atol=0)
    I = InterpolateTrace(A, p=-1, kind='imbf', ti=[1e-2, 1e-1, 1.0, 10.0],
                         options=OPTIONS)
    assert numpy.allclose(I.eval(t), got, rtol=1e-10, atol=0)


def test_accuracy_against_dense_eigenvalues():
    """Enough probes for the Monte-Carlo error to sit below the fit error:
    the interpolated logdet / n follows the exact curve."""
    A = sparse_correlation_matrix(1024, nnz_per_row=30, seed=2)
    lam = scipy.linalg.eigvalsh(A.toarray())
    opts = dict(OPTIONS, min_num_samples=256, max_num_samples=256)
    L = InterpolateLogdet(A, kind='imbf', ti=[1e-3, 1e-2, 1e-1, 1.0, 10.0],
                          options=opts)
    t = numpy.geomspace(1e-3, 10.0, 9)
    exact = numpy.array([numpy.sum(numpy.log(lam + tk)) for tk in t])
    assert numpy.max(numpy.abs(L.interpolate(t) - exact)) / 1024 < 2e-2
