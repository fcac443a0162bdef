"""The interpolation kinds of imate's interpolators.

Each kind fits tau(t) = |A + tB|_p / |B|_p from its values tau_i at the
points t_i (computed on the device by ``InterpolantBase``) and tau0 = tau(0);
the fits are small dense problems solved on the host. Each class cites the
reference method it follows; all evaluate elementwise over numpy arrays of t.
"""

import math
from fractions import Fraction
from numbers import Number

import numpy
import numpy.polynomial.chebyshev as cheb
import scipy.interpolate
import scipy.integrate
import scipy.optimize

from ._interpolant_base import InterpolantBase

__all__ = ['ExactMethod', 'EigenvaluesMethod', 'MonomialBasisFunctionsMethod',
           'InverseMonomialBasisFunctionsMethod', 'RadialBasisFunctionsMethod',
           'RationalPolynomialFunctionsMethod',
           'ChebyshevRationalFunctionsMethod', 'SplineMethod']


def _points(ti):
    """``ti`` as a 1-D float array; raises on None or an empty list."""
    if ti is None:
        raise ValueError('"ti" should be a list or array.')
    if isinstance(ti, Number):
        ti = [ti]
    ti = numpy.asarray(ti, dtype=float).ravel()
    if ti.size == 0:
        raise ValueError('"ti" should be a list or array.')
    return ti


def _scalar_or_array(t, value):
    return float(value) if numpy.isscalar(t) else value


# ------------------------------------------------------------------ EXT, EIG

class ExactMethod(InterpolantBase):
    """No interpolation: every query is an estimator call
    (imate/interpolator/_exact_method.py:252-270)."""

    def __init__(self, A, B=None, p=2, options={}, verbose=False):
        super(ExactMethod, self).__init__(A, B=B, p=p, ti=None,
                                          options=options, verbose=verbose)
        self.q = 0

    def interpolate(self, t):
        return self.eval(t)


class EigenvaluesMethod(InterpolantBase):
    """imate's 'eig' kind computes every eigenvalue of A on the CPU
    (imate/interpolator/_eigenvalues_method.py:317-380): a dense eigenvalue
    method, not the stochastic Lanczos quadrature this package runs."""

    def __init__(self, A, B=None, p=2, options={}, verbose=False,
                 non_zero_ratio=0.9, tol=1e-3):
        if B is not None:
            raise ValueError('In "EIG" method, "B" should be "None".')
        raise NotImplementedError(
            'kind "eig" (all eigenvalues on the CPU) is not part of '
            'imate_amd; use kind "ext" with method="slq" for exact values.')


# ------------------------------------------------------------------ MBF

class MonomialBasisFunctionsMethod(InterpolantBase):
    """tau(t)^2 = t^2 + c t + tau0^2 with c from one point t1 (default
    1 / tau0) (imate/interpolator/_monomial_basis_functions_method.py:
    369-453). Exact for p = 2 and B = I, where tau(t)^2 = mean((lambda +
    t)^2)."""

    def __init__(self, A, B=None, p=2, options={}, verbose=False, ti=[]):
        if ti is None or (not isinstance(ti, Number) and len(ti) == 0):
            # t1 = 1 / tau0 needs tau0 first: two estimator calls
            super(MonomialBasisFunctionsMethod, self).__init__(
                A, B=B, p=p, ti=[], options=options, verbose=verbose)
            self.t1 = 1.0 / self.tau0
            schatten_1 = self.eval(self.t1)
            self.t_i = numpy.array([self.t1])
            self.q = 1
            self.schatten_i = numpy.array([schatten_1])
            self.tau_i = self.schatten_i / self.schatten_B
        else:
            t = _points(ti)
            if t.size != 1:
                raise TypeError('"ti" for the "mbf" method should be a single '
                                'number, or an array or a list of a single '
                                'number.')
            # t = 0 and t1 in one call
            super(MonomialBasisFunctionsMethod, self).__init__(
                A, B=B, p=p, ti=t, options=options, verbose=verbose)
            self.t1 = float(t[0])
        self.tau1 = float(self.tau_i[0])
        self.scale_t = self.find_scale_ti()

    def interpolate(self, t):
        ta = numpy.asarray(t, dtype=float)
        c = (self.tau1**2 - self.tau0**2 - self.t1**2) / self.t1
        tau = numpy.sqrt(ta**2 + c * ta + self.tau0**2)
        return _scalar_or_array(t, tau * self.schatten_B)


# ------------------------------------------------------------------ IMBF

_ORTHO_MAX = 9      # basis functions of the orthogonal sets


def orthonormal_root_monomials(first, count=_ORTHO_MAX):
    """Coefficients C[j, i] of psi_j(t) = sum_i C[j, i] t^(1/(first + i)),
    i <= j: Gram-Schmidt of t^(1/first), t^(1/(first+1)), ... under
    <f, g> = int_0^1 f(t) g(t) dt / t, whose Gram matrix 1 / (1/a + 1/b) is
    rational, so the process runs in exact fractions; psi_j has a positive
    coefficient on its newest monomial. first = 1 gives imate's 'ortho' set
    and first = 2 its 'ortho2' set (the tables of
    _inverse_monomial_basis_functions_method.py:631-723, alpha[j] a[j, i])."""
    e = [Fraction(1, first + i) for i in range(count)]
    gram = [[1 / (e[i] + e[k]) for k in range(count)] for i in range(count)]

    def inner(u, v):
        return sum(u[i] * v[k] * gram[i][k] for i in range(count) if u[i]
                   for k in range(count) if v[k])

    C = numpy.zeros((count, count))
    done = []
    for j in range(count):
        v = [Fraction(0)] * count
        v[j] = Fraction(1)
        for (b, bb) in done:
            c = inner(v, b) / bb
            v = [v[i] - c * b[i] for i in range(count)]
        vv = inner(v, v)
        done.append((v, vv))
        norm = math.sqrt(float(vv))
        for i in range(j + 1):
            C[j, i] = float(v[i]) / norm
    return C


_ORTHO_SETS = {}


def _ortho_set(first):
    if first not in _ORTHO_SETS:
        _ORTHO_SETS[first] = orthonormal_root_monomials(first)
    return _ORTHO_SETS[first]


class InverseMonomialBasisFunctionsMethod(InterpolantBase):
    """tau(t) = tau0 + sum_j w_j phi_j(t) [+ t] over root monomials
    t^(1/k) (imate/interpolator/_inverse_monomial_basis_functions_method.py:
    421-783):

    * 'non-ortho': phi_j = t^(1/(j+2)), plus t; w solves the q x q system.
    * 'ortho': orthonormal set from t, t^(1/2), ..., of t / scale_t, q + 1
      functions with the extra condition that the coefficient of t is one.
    * 'ortho2' (default): orthonormal set from t^(1/2), t^(1/3), ..., of
      t / scale_t, plus t.
    """

    def __init__(self, A, B=None, p=2, options={}, verbose=False, ti=[],
                 basis_func_type='ortho2'):
        ti = _points(ti)
        if basis_func_type not in ('non-ortho', 'ortho', 'ortho2'):
            raise ValueError('Method is invalid.')
        limit = _ORTHO_MAX - 1 if basis_func_type == 'ortho' else _ORTHO_MAX
        if basis_func_type != 'non-ortho' and ti.size > limit:
            raise ValueError('Cannot regress order higher than %d. Decrease '
                             'the number of interpolation points.' % limit)
        super(InverseMonomialBasisFunctionsMethod, self).__init__(
            A, B=B, p=p, ti=ti, options=options, verbose=verbose)
        self.basis_func_type = basis_func_type
        self.w = None
        self.initialize_interpolator()

    def basis(self, t):
        """[len(t), K] values of the basis functions at t (t already
        divided by scale_t for the orthogonal sets)."""
        t = numpy.atleast_1d(numpy.asarray(t, dtype=float))
        if self.basis_func_type == 'non-ortho':
            k = numpy.arange(self.q) + 2.0
            return t[:, None]**(1.0 / k[None, :])
        first = 1 if self.basis_func_type == 'ortho' else 2
        K = self.q + 1 if self.basis_func_type == 'ortho' else self.q
        C = _ortho_set(first)[:K, :K]
        k = numpy.arange(K) + float(first)
        monomials = t[:, None]**(1.0 / k[None, :])
        return monomials @ C.T

    def initialize_interpolator(self):
        if self.basis_func_type == 'non-ortho':
            C = self.basis(self.t_i)
            b = self.tau_i - self.tau0 - self.t_i
        elif self.basis_func_type == 'ortho':
            q = self.q
            C = numpy.zeros((q + 1, q + 1))
            C[:q] = self.basis(self.t_i / self.scale_t)
            # the coefficient of the monomial t (first of the set) is one
            C[q] = _ortho_set(1)[:q + 1, 0]
            b = numpy.r_[self.tau_i - self.tau0, 1.0]
        else:
            C = self.basis(self.t_i / self.scale_t)
            b = self.tau_i - self.tau0 - self.t_i
        if self.verbose:
            print('Condition number: %f' % numpy.linalg.cond(C))
        self.w = numpy.linalg.solve(C, b)

    def interpolate(self, t):
        ta = numpy.asarray(t, dtype=float)
        flat = numpy.atleast_1d(ta).ravel()
        if self.basis_func_type == 'non-ortho':
            tau = self.tau0 + self.basis(flat) @ self.w + flat
        elif self.basis_func_type == 'ortho':
            tau = self.tau0 + self.basis(flat / self.scale_t) @ self.w
        else:
            tau = self.tau0 + self.basis(flat / self.scale_t) @ self.w + flat
        tau = tau.reshape(ta.shape)
        return _scalar_or_array(t, tau * self.schatten_B)


# ------------------------------------------------------------------ RBF

def gaussian_rbf_fit(x, y, epsilon):
    """Weights of the Gaussian radial basis interpolant
    s(x) = sum_j w_j exp(-((x - x_j) / epsilon)^2) through (x_j, y_j) (the
    'gaussian' function of scipy.interpolate.Rbf, no smoothing)."""
    phi = numpy.exp(-((x[:, None] - x[None, :]) / epsilon)**2)
    return numpy.linalg.solve(phi, y)


def gaussian_rbf_eval(xq, x, w, epsilon):
    xq = numpy.atleast_1d(xq)
    return numpy.exp(-((xq[:, None] - x[None, :]) / epsilon)**2) @ w


class RadialBasisFunctionsMethod(InterpolantBase):
    """Interpolation in x = log10(t) (imate/interpolator/
    _radial_basis_functions_method.py:377-525):

    * func_type 1: y = tau - tau0 - t by a cubic spline with zero slope at
      the left end and zero curvature at the right end; tau = y + tau0 + t.
    * func_type 2: y = tau / (tau0 + t) - 1 by Gaussian radial basis
      functions of width mean(diff(x)), padded with three zeros beyond each
      end; tau = (y + 1)(tau0 + t).

    Outside [low, high] (by default log10 t in [-4.5, 3.5], widened past the
    points by one spacing) y is zero.
    """

    def __init__(self, A, B=None, p=0, options={}, verbose=False, ti=[],
                 func_type=1):
        ti = _points(ti)
        if func_type not in (1, 2):
            raise ValueError('Invalid function type.')
        super(RadialBasisFunctionsMethod, self).__init__(
            A, B=B, p=p, ti=ti, options=options, verbose=verbose)
        self.func_type = func_type
        self.RBF = None
        self.low_log_threshold = None
        self.high_log_threshold = None
        self.initialize_interpolator()

    def initialize_interpolator(self):
        xi = numpy.log10(self.t_i)
        dxi = float(numpy.mean(numpy.diff(xi))) if xi.size > 1 else 1.0
        if self.func_type == 1:
            yi = self.tau_i - self.tau0 - self.t_i
        else:
            yi = self.tau_i / (self.tau0 + self.t_i) - 1.0
        low, high = -4.5, 3.5
        if low >= numpy.min(xi):
            low = float(numpy.min(xi)) - dxi
        if high <= numpy.max(xi):
            high = float(numpy.max(xi)) + dxi
        self.low_log_threshold = low
        self.high_log_threshold = high
        if self.func_type == 1:
            self.RBF = scipy.interpolate.CubicSpline(
                xi, yi, bc_type=((1, 0.0), (2, 0.0)), extrapolate=True)
        else:
            pad = 3
            xi = numpy.r_[numpy.linspace(low - dxi, low, pad), xi,
                          numpy.linspace(high, high + dxi, pad)]
            yi = numpy.r_[numpy.zeros(pad), yi, numpy.zeros(pad)]
            w = gaussian_rbf_fit(xi, yi, dxi)
            self.RBF = (lambda x, xi=xi, w=w, e=dxi:
                        gaussian_rbf_eval(x, xi, w, e))

    def interpolate(self, t):
        ta = numpy.asarray(t, dtype=float)
        flat = numpy.atleast_1d(ta).ravel()
        x = numpy.log10(flat)
        inside = (x >= self.low_log_threshold) & (x <= self.high_log_threshold)
        y = numpy.zeros_like(flat)
        if numpy.any(inside):
            y[inside] = self.RBF(x[inside])
        if self.func_type == 1:
            tau = y + self.tau0 + flat
        else:
            tau = (y + 1.0) * (self.tau0 + flat)
        return _scalar_or_array(t, tau.reshape(ta.shape) * self.schatten_B)


# ------------------------------------------------------------------ RPF

class RationalPolynomialFunctionsMethod(InterpolantBase):
    """tau(t) = (t^(q+1) + b_q t^q + ... + b_0) / (t^q + a_(q-1) t^(q-1) +
    ... + a_0) with a_0 = b_0 / tau0 (so tau(0) = tau0), the 2q unknowns from
    2q points (imate/interpolator/_rational_polynomial_functions_method.py:
    361-515). Exact for p = -1, B = I and a matrix with q + 1 distinct
    eigenvalues of equal multiplicity, where tau is such a ratio."""

    def __init__(self, A, B=None, p=0, options={}, verbose=False, ti=[]):
        ti = _points(ti)
        if ti.size % 2 != 0:
            raise ValueError('In rational polynomial interpolation method, '
                             'the number of interpolation points should be '
                             'even.')
        super(RationalPolynomialFunctionsMethod, self).__init__(
            A, B=B, p=p, ti=ti, options=options, verbose=verbose)
        self.numerator = None
        self.denominator = None
        self.initialize_interpolator()

    def fit_rational_poly(self, t_i, tau_i, tau0):
        """(numerator, denominator) coefficient lists, highest power first."""
        q = t_i.size // 2
        # unknowns x = [b_0 .. b_q, a_1 .. a_(q-1)]; at each point
        # tau_i (t^q + b_0/tau0 + sum a_j t^j) = t^(q+1) + sum b_j t^j
        powers = t_i[:, None]**numpy.arange(q + 1)[None, :]
        C = numpy.zeros((2 * q, 2 * q))
        C[:, 0] = 1.0 - tau_i / tau0
        C[:, 1:q + 1] = powers[:, 1:]
        C[:, q + 1:] = -tau_i[:, None] * powers[:, 1:q]
        rhs = tau_i * t_i**q - t_i**(q + 1)
        if self.verbose:
            print('Condition number: %0.2e' % numpy.linalg.cond(C))
        x = numpy.linalg.solve(C, rhs)
        b = x[:q + 1]
        a = numpy.r_[b[0] / tau0, x[q + 1:]]
        numerator = numpy.r_[1.0, b[::-1]]
        denominator = numpy.r_[1.0, a[::-1]]
        poles = numpy.roots(denominator)
        if numpy.any(numpy.real(poles) > 0):
            raise ValueError('rational_polynomial has positive poles: %s'
                             % poles)
        return numerator, denominator

    def initialize_interpolator(self):
        self.numerator, self.denominator = self.fit_rational_poly(
            self.t_i, self.tau_i, self.tau0)

    def interpolate(self, t):
        ta = numpy.asarray(t, dtype=float)
        tau = numpy.polyval(self.numerator, ta) / \
            numpy.polyval(self.denominator, ta)
        return _scalar_or_array(t, tau * self.schatten_B)


# ------------------------------------------------------------------ CRF

def chebyshev_rational_roots(degree):
    """Points t = (1 + x) / (1 - x) at the roots x of T_degree, ascending."""
    k = numpy.arange(1, degree + 1)
    x = numpy.sort(numpy.cos((2.0 * k - 1.0) * numpy.pi / (2.0 * degree)))
    return (1.0 + x) / (1.0 - x)


def _crf_series(coeff):
    """Chebyshev series of y(x) = sum_j coeff_j (1 - T_(j+1)(x)) / 2."""
    c = numpy.zeros(coeff.size + 1)
    c[0] = 0.5 * numpy.sum(coeff)
    c[1:] = -0.5 * coeff
    return c


class ChebyshevRationalFunctionsMethod(InterpolantBase):
    """y(x) = sum_j c_j (1 - T_(j+1)(x)) / 2 over x = (t/s - 1)/(t/s + 1)
    (imate/interpolator/_chebyshev_rational_functions_method.py:435-706):

    * func_type 1: y = tau / (tau0 + t) - 1, with y(-1) = 0 added.
    * func_type 2: y = (tau - tau0) / t - 1.

    ``ti`` may be an integer: that many Chebyshev-rational points. With
    ``scale=None`` the scale s minimises the curvature integral
    int_{-1}^{1} y''^2 / (1 + y'^2)^(5/2) dx (Nelder-Mead from s = 1).
    """

    def __init__(self, A, B=None, p=0, options={}, verbose=False, ti=[],
                 scale=None, func_type=1):
        if ti is None or (not numpy.isscalar(ti) and len(ti) == 0):
            raise ValueError('"ti" should be an integer, list, or array.')
        if numpy.isscalar(ti):
            if not isinstance(ti, (int, numpy.integer)):
                raise ValueError('"ti" as a scalar should be an integer.')
            ti = chebyshev_rational_roots(int(ti))
        if func_type not in (1, 2):
            raise ValueError('"type" should be 1 or 2.')
        super(ChebyshevRationalFunctionsMethod, self).__init__(
            A, B=B, p=p, ti=ti, options=options, verbose=verbose)
        self.func_type = func_type
        self.scale = scale
        self.coeff = None
        self.initialize_interpolator()

    def _data(self):
        if self.func_type == 1:
            return self.tau_i / (self.tau0 + self.t_i) - 1.0
        return (self.tau_i - self.tau0) / self.t_i - 1.0

    def linear_system(self, scale, t_i, yi):
        """Coefficients c interpolating (x_i, y_i) for scale s."""
        scale = abs(float(scale))
        ts = t_i / scale
        xi = (ts - 1.0) / (ts + 1.0)
        if self.func_type == 1:
            xi = numpy.r_[-1.0, xi]
            yi = numpy.r_[0.0, yi]
        K = xi.size
        T = numpy.empty((K, K))
        for j in range(K):
            e = numpy.zeros(j + 2)
            e[-1] = 1.0
            T[:, j] = 0.5 * (1.0 - cheb.chebval(xi, e))
        return numpy.linalg.solve(T, yi)

    def curvature_integral(self, scale, t_i, yi):
        scale = float(numpy.ravel(scale)[0])
        if scale < 1e-5:
            return numpy.inf
        c = _crf_series(self.linear_system(scale, t_i, yi))
        d1 = cheb.chebder(c)
        d2 = cheb.chebder(c, 2)

        def integrand(x):
            return cheb.chebval(x, d2)**2 / \
                (1.0 + cheb.chebval(x, d1)**2)**2.5

        return scipy.integrate.quad(integrand, -1.0, 1.0)[0]

    def initialize_interpolator(self):
        yi = self._data()
        if self.scale is None:
            res = scipy.optimize.minimize(self.curvature_integral, 1.0,
                                          args=(self.t_i, yi),
                                          method='Nelder-Mead')
            self.scale = abs(float(res.x[0]))
        self.coeff = self.linear_system(self.scale, self.t_i, yi)
        if self.verbose:
            print('tau0: %f' % self.tau0)
            print('scale: %f' % self.scale)

    def interpolate(self, t):
        ta = numpy.asarray(t, dtype=float)
        ts = ta / abs(float(self.scale))
        x = (ts - 1.0) / (ts + 1.0)
        y = cheb.chebval(x, _crf_series(self.coeff))
        if self.func_type == 1:
            tau = (y + 1.0) * (self.tau0 + ta)
        else:
            tau = (y + 1.0) * ta + self.tau0
        return _scalar_or_array(t, tau * self.schatten_B)


# ------------------------------------------------------------------ SPL

class SplineMethod(InterpolantBase):
    """Cubic spline of y over x = (t - 1)/(t + 1) in [-1, 1]
    (imate/interpolator/_spline_method.py:366-450), y as in the 'crf' kind;
    func_type 1 pins y(-1) = y(1) = 0, func_type 2 pins y(1) = 0."""

    def __init__(self, A, B=None, p=0, options={}, verbose=False, ti=[],
                 func_type=1):
        ti = _points(ti)
        if func_type not in (1, 2):
            raise ValueError('"type" should be 1 or 2.')
        super(SplineMethod, self).__init__(A, B=B, p=p, ti=ti,
                                           options=options, verbose=verbose)
        self.func_type = func_type
        self.interp_obj = None
        self.initialize_interpolator()

    def initialize_interpolator(self):
        if self.func_type == 1:
            yi = self.tau_i / (self.tau0 + self.t_i) - 1.0
        else:
            yi = (self.tau_i - self.tau0) / self.t_i - 1.0
        xi = (self.t_i - 1.0) / (self.t_i + 1.0)
        if self.func_type == 1:
            xi = numpy.r_[-1.0, xi, 1.0]
            yi = numpy.r_[0.0, yi, 0.0]
        else:
            xi = numpy.r_[xi, 1.0]
            yi = numpy.r_[yi, 0.0]
        self.interp_obj = scipy.interpolate.CubicSpline(xi, yi)

    def interpolate(self, t):
        ta = numpy.asarray(t, dtype=float)
        x = (ta - 1.0) / (ta + 1.0)
        y = self.interp_obj(x)
        if self.func_type == 1:
            tau = (y + 1.0) * (self.tau0 + ta)
        else:
            tau = (y + 1.0) * ta + self.tau0
        return _scalar_or_array(t, tau * self.schatten_B)
