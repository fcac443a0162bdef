"""CPU tests of the interpolators' host side: every kind is fitted to exact
Schatten norms (dense eigenvalues of small matrices, computed here as the
reference values) injected in place of the device estimates, and pinned by what the
mathematics fixes: interpolation through the points, closed forms a kind
reproduces exactly (MBF for p = 2; RPF for p = -1 with q + 1 distinct
eigenvalues), orthonormality of the root-monomial bases, the bounds, and the
library routines the kinds restate (scipy's Gaussian Rbf). The device path
itself is covered by test_gpu_interpolator.py.
"""

import numpy
import numpy.polynomial.chebyshev as cheb
import pytest
import scipy.integrate
import scipy.interpolate
import scipy.linalg
import scipy.sparse

import imate_amd
from imate_amd.interpolator import _interpolant_base, _methods
from imate_amd.interpolator import (InterpolateSchatten, InterpolateLogdet,
                                    InterpolateTrace)


def _exact_schatten(M, p):
    lam = scipy.linalg.eigvalsh(M)
    if p == 0:
        return float(numpy.exp(numpy.mean(numpy.log(lam))))
    return float(numpy.mean(lam**p)**(1.0 / p))


@pytest.fixture
def exact(monkeypatch):
    """Replace the device estimates by dense eigenvalue values."""
    calls = []

    def schatten_at(self, t):
        t = numpy.atleast_1d(numpy.asarray(t, dtype=float))
        calls.append(t.size)
        A = self.A.toarray() if scipy.sparse.issparse(self.A) else self.A
        if self.B_is_identity:
            B = numpy.eye(self.n)
        else:
            B = self.B.toarray() if scipy.sparse.issparse(self.B) else self.B
        return numpy.array([_exact_schatten(A + tk * B, self.p) for tk in t])

    def compute_schatten(self, M, p):
        M = M.toarray() if scipy.sparse.issparse(M) else M
        return _exact_schatten(M, p)

    monkeypatch.setattr(_interpolant_base.InterpolantBase, '_schatten_at',
                        schatten_at)
    monkeypatch.setattr(_interpolant_base.InterpolantBase,
                        '_compute_schatten', compute_schatten)
    return calls


def _spd(n, seed, cond=50.0):
    rng = numpy.random.default_rng(seed)
    Q, _ = numpy.linalg.qr(rng.standard_normal((n, n)))
    lam = numpy.geomspace(1.0 / cond, 1.0, n)
    return (Q * lam) @ Q.T


def _exact_curve(A, B, p, t):
    B = numpy.eye(A.shape[0]) if B is None else B
    return numpy.array([_exact_schatten(A + tk * B, p) for tk in t])


TI = [1e-3, 1e-2, 1e-1, 1.0, 10.0]
TI4 = [1e-2, 1e-1, 1.0, 10.0]


# ------------------------------------------------------- orthonormal sets

@pytest.mark.parametrize('first', [1, 2])
def test_root_monomial_sets_are_orthonormal(first):
    C = _methods.orthonormal_root_monomials(first)
    K = C.shape[0]

    def psi(j, t):
        k = numpy.arange(K) + first
        return numpy.sum(C[j] * t**(1.0 / k))

    # <f, g> = int_0^1 f g dt / t = int_0^inf f(e^-u) g(e^-u) du
    G = numpy.empty((K, K))
    for i in range(K):
        for j in range(i + 1):
            G[i, j] = G[j, i] = scipy.integrate.quad(
                lambda u: psi(i, numpy.exp(-u)) * psi(j, numpy.exp(-u)),
                0.0, numpy.inf, limit=400, epsabs=1e-11, epsrel=1e-10)[0]
    assert numpy.allclose(G, numpy.eye(K), atol=1e-8)
    assert numpy.all(numpy.diag(C) > 0)
    assert numpy.allclose(numpy.triu(C, 1), 0.0)


def test_root_monomial_first_functions():
    """psi_0 = t^(1/first) / |t^(1/first)|: |t^(1/a)|^2 = a / 2."""
    for first in (1, 2):
        C = _methods.orthonormal_root_monomials(first)
        assert abs(C[0, 0] - numpy.sqrt(2.0 / first)) < 1e-15


# ------------------------------------------------------- fits through points

KINDS = [('imbf', dict(basis_func_type='non-ortho'), TI),
         ('imbf', dict(basis_func_type='ortho'), TI),
         ('imbf', dict(basis_func_type='ortho2'), TI),
         ('rbf', dict(func_type=1), TI),
         ('rbf', dict(func_type=2), TI),
         ('rpf', {}, TI4),
         ('crf', dict(func_type=1), TI4),
         ('crf', dict(func_type=2, scale=0.1), TI4),
         ('spl', dict(func_type=1), TI4),
         ('spl', dict(func_type=2), TI4)]


@pytest.mark.parametrize('kind,kw,ti', KINDS)
@pytest.mark.parametrize('p', [0, -1, 2])
@pytest.mark.parametrize('with_b', [False, True])
def test_kind_passes_through_points(exact, kind, kw, ti, p, with_b):
    A = _spd(30, 0)
    B = _spd(30, 1, cond=4.0) if with_b else None
    S = InterpolateSchatten(A, B=B, p=p, kind=kind, ti=ti, **kw)
    want = _exact_curve(A, B, p, ti)
    got = S.interpolate(numpy.array(ti))
    assert numpy.allclose(got, want, rtol=1e-9, atol=0), (got, want)
    # one estimator call for t = 0 and the points, one for |B|_p
    assert exact[0] == len(ti) + 1


@pytest.mark.parametrize('kind,kw,ti', KINDS)
def test_kind_accuracy_between_points(exact, kind, kw, ti):
    """Sanity of each fit between its points: the error of logdet / n (the
    relative error of |A + tI|_0) over [t_1, t_q]. The 'ortho' basis (whose
    asymptote is t / scale_t) and the func_type 2 spline are the loosest
    kinds, as in imate's own comparison of the kinds."""
    A = _spd(40, 2)
    S = InterpolateLogdet(A, kind=kind, ti=ti, **kw)
    t = numpy.geomspace(ti[0], ti[-1], 9)
    want = 40 * numpy.log(_exact_curve(A, None, 0, t))
    got = S.interpolate(t)
    loose = kw.get('basis_func_type') == 'ortho' or \
        (kind == 'spl' and kw.get('func_type') == 2)
    assert numpy.max(numpy.abs(got - want)) / 40 < (0.1 if loose else 2e-2)


def test_mbf_exact_for_p2_identity(exact):
    A = _spd(25, 3)
    for ti in ([], [0.3]):
        S = InterpolateSchatten(A, p=2, kind='mbf', ti=ti)
        t = numpy.geomspace(1e-4, 1e3, 15)
        assert numpy.allclose(S.interpolate(t), _exact_curve(A, None, 2, t),
                              rtol=1e-12)
    S = InterpolateSchatten(A, p=2, kind='mbf')
    assert abs(S.interpolator.t1 - 1.0 / S.interpolator.tau0) < 1e-15


def test_rpf_exact_for_three_eigenvalues(exact):
    """tau(t) = 1 / mean(1 / (lambda + t)) for eigenvalues {1, 2, 5} is a
    ratio of a monic cubic and a monic quadratic: q = 2 recovers it."""
    A = numpy.diag(numpy.repeat([1.0, 2.0, 5.0], 4))
    S = InterpolateSchatten(A, p=-1, kind='rpf', ti=TI4)
    t = numpy.geomspace(1e-4, 1e3, 20)
    assert numpy.allclose(S.interpolate(t), _exact_curve(A, None, -1, t),
                          rtol=1e-10)
    num, den = S.interpolator.numerator, S.interpolator.denominator
    assert numpy.allclose(num, numpy.poly([-1.0, -2.0, -5.0]), rtol=1e-9)
    assert len(den) == 3


def test_rpf_positive_pole_raises(exact):
    A = _spd(10, 4)
    m = InterpolateSchatten(A, p=-1, kind='rpf', ti=TI4).interpolator
    # data of (t^2 + t + 1) / (t - 2): the fit recovers its pole at t = 2
    t = numpy.array([0.1, 10.0])
    f = (t**2 + t + 1.0) / (t - 2.0)
    with pytest.raises(ValueError):
        m.fit_rational_poly(t, f, -0.5)


def test_crf_scale_minimises_curvature(exact):
    A = _spd(30, 5)
    S = InterpolateLogdet(A, kind='crf', ti=TI4)
    m = S.interpolator
    yi = m._data()
    s = S.get_scale()
    f0 = m.curvature_integral(s, m.t_i, yi)
    for r in (0.8, 0.95, 1.05, 1.25):
        assert f0 <= m.curvature_integral(s * r, m.t_i, yi) + 1e-12


def test_crf_integer_points(exact):
    A = _spd(20, 6)
    S = InterpolateSchatten(A, p=0, kind='crf', ti=5, scale=1.0)
    t = S.interpolator.t_i
    x = (t - 1.0) / (t + 1.0)
    assert numpy.allclose(cheb.chebval(x, [0] * 5 + [1]), 0.0, atol=1e-12)
    assert numpy.all(numpy.diff(t) > 0)
    with pytest.raises(ValueError):
        InterpolateSchatten(A, p=0, kind='crf', ti=2.5)


def test_gaussian_rbf_matches_scipy():
    rng = numpy.random.default_rng(7)
    x = numpy.sort(rng.uniform(-4, 3, 9))
    y = rng.standard_normal(9)
    eps = float(numpy.mean(numpy.diff(x)))
    w = _methods.gaussian_rbf_fit(x, y, eps)
    ref = scipy.interpolate.Rbf(x, y, function='gaussian', epsilon=eps)
    xq = numpy.linspace(-4, 3, 31)
    assert numpy.allclose(_methods.gaussian_rbf_eval(xq, x, w, eps), ref(xq),
                          rtol=1e-8, atol=1e-10)


# ------------------------------------------------------- front ends, bounds

def test_logdet_and_trace_front_ends(exact):
    A = _spd(30, 8)
    n = 30
    t = numpy.array([0.05, 0.5, 5.0])
    L = InterpolateLogdet(A, kind='imbf', ti=TI)
    S = InterpolateSchatten(A, p=0, kind='imbf', ti=TI)
    assert numpy.allclose(L.interpolate(t), n * numpy.log(S.interpolate(t)),
                          rtol=1e-14)
    assert numpy.allclose(L.eval(t), n * numpy.log(_exact_curve(A, None, 0, t)),
                          rtol=1e-12)
    T = InterpolateTrace(A, p=-1, kind='imbf', ti=TI)
    S1 = InterpolateSchatten(A, p=-1, kind='imbf', ti=TI)
    assert numpy.allclose(T.interpolate(t), n / S1.interpolate(t), rtol=1e-14)
    assert isinstance(L(0.5), float)
    assert L(numpy.array([0.5])).shape == (1,)


@pytest.mark.parametrize('with_b', [False, True])
def test_bounds(exact, with_b):
    A = _spd(30, 9)
    B = _spd(30, 10, cond=3.0) if with_b else None
    Bd = numpy.eye(30) if B is None else B
    t = numpy.geomspace(1e-3, 1e2, 12)
    L = InterpolateLogdet(A, B=B, kind='ext')
    exact_ld = numpy.array([numpy.linalg.slogdet(A + tk * Bd)[1] for tk in t])
    assert numpy.all(L.lower_bound(t) <= exact_ld + 1e-9)
    T = InterpolateTrace(A, B=B, p=-1, kind='ext')
    exact_ti = numpy.array([numpy.trace(numpy.linalg.inv(A + tk * Bd))
                            for tk in t])
    assert numpy.all(T.upper_bound(t) >= exact_ti * (1 - 1e-12))
    assert numpy.all(T.lower_bound(t) <= exact_ti * (1 + 1e-12))
    S2 = InterpolateSchatten(A, B=B, p=2, kind='ext')
    assert numpy.all(S2.bound(t) >= _exact_curve(A, B, 2, t) * (1 - 1e-12))


def test_ext_is_exact(exact):
    A = _spd(15, 11)
    S = InterpolateSchatten(A, p=1.5, kind='ext')
    t = numpy.array([0.0, 0.1, 2.0])
    assert numpy.allclose(S.interpolate(t), _exact_curve(A, None, 1.5, t),
                          rtol=1e-13)
    assert exact == [3]      # one call for the three points


def test_argument_checks(exact):
    A = _spd(10, 12)
    with pytest.raises(ValueError):
        InterpolateSchatten(A, kind='bogus', ti=TI)
    with pytest.raises(NotImplementedError):
        InterpolateSchatten(A, kind='eig')
    with pytest.raises(ValueError):
        InterpolateSchatten(A, kind='eig', B=A)
    for kind in ('imbf', 'rbf', 'rpf', 'spl', 'crf'):
        with pytest.raises(ValueError):
            InterpolateSchatten(A, kind=kind, ti=[])
    with pytest.raises(ValueError):
        InterpolateSchatten(A, p=-1, kind='rpf', ti=[0.1, 1.0, 10.0])
    with pytest.raises(ValueError):
        InterpolateSchatten(A, kind='imbf', ti=numpy.geomspace(1e-3, 1, 10))
    with pytest.raises(ValueError):
        InterpolateSchatten(A, kind='imbf', ti=numpy.geomspace(1e-3, 1, 9),
                            basis_func_type='ortho')
    with pytest.raises(ValueError):
        InterpolateSchatten(A, kind='imbf', ti=TI, basis_func_type='x')
    with pytest.raises(TypeError):
        InterpolateSchatten(A, kind='mbf', ti=[0.1, 0.2])
    with pytest.raises(ValueError):
        InterpolateTrace(A, p=0, ti=TI)
    with pytest.raises(ValueError):
        InterpolateSchatten(A, ti=TI, options={'parameters': [1.0]})
    with pytest.raises(NotImplementedError):
        InterpolateSchatten(A, kind='imbf', ti=TI).get_scale()
    with pytest.raises(ValueError):
        InterpolateSchatten(A, p=2, kind='imbf', ti=TI).upper_bound(1.0)
    with pytest.raises(NotImplementedError):
        InterpolateSchatten(A, kind='imbf', ti=TI).plot(numpy.array(TI))


def test_device_path_needs_slq_method():
    """Without injected values the points go to ``schatten``; the default
    method is not part of this package and raises before any device use."""
    A = _spd(10, 13)
    with pytest.raises(NotImplementedError):
        InterpolateLogdet(A, kind='imbf', ti=TI)
    assert imate_amd.InterpolateLogdet is InterpolateLogdet
