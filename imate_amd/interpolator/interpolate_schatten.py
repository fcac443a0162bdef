"""``InterpolateSchatten``: t -> |A + tB|_p from a few estimator calls
(imate/interpolator/interpolate_schatten.py:42-1064, the call surface without
plotting)."""

from numbers import Number

import numpy

from ._methods import (ExactMethod, EigenvaluesMethod,
                       MonomialBasisFunctionsMethod,
                       InverseMonomialBasisFunctionsMethod,
                       RadialBasisFunctionsMethod,
                       RationalPolynomialFunctionsMethod,
                       ChebyshevRationalFunctionsMethod, SplineMethod)

__all__ = ['InterpolateSchatten']

_KINDS = {
    'ext': ExactMethod,
    'eig': EigenvaluesMethod,
    'mbf': MonomialBasisFunctionsMethod,
    'imbf': InverseMonomialBasisFunctionsMethod,
    'rbf': RadialBasisFunctionsMethod,
    'crf': ChebyshevRationalFunctionsMethod,
    'spl': SplineMethod,
    'rpf': RationalPolynomialFunctionsMethod,
}


class InterpolateSchatten(object):
    """Interpolation of the normalised Schatten norm of A + tB over t.

    ``options`` are passed to ``schatten`` at the interpolation points: use
    ``{'method': 'slq', 'gpu': True, ...}`` (the other methods of imate are not
    part of this package and raise). The points of a kind are estimated in one
    call; with B = None the Lanczos recurrences run once per probe for all of
    them. ``interpolate``, ``eval``, ``bound`` and ``upper_bound`` take a number
    or an array of t.
    """

    def __init__(self, A, B=None, p=2, options={}, verbose=False, kind='imbf',
                 ti=[], **kwargs):
        if not isinstance(kind, str) or kind.lower() not in _KINDS:
            raise ValueError("'kind' is invalid. Select one of 'ext', 'eig', "
                             "'mbf', 'imbf', 'rbf', 'crf', 'spl', or 'rpf'.")
        self.kind = kind
        self.verbose = verbose
        self.n = A.shape[0]
        if ti is not None and not numpy.isscalar(ti):
            self.q = len(ti)
        else:
            self.q = 0
        self.p = p
        method = _KINDS[kind.lower()]
        if kind.lower() in ('ext', 'eig'):
            self.interpolator = method(A, B, p=self.p, options=options,
                                       verbose=verbose, **kwargs)
        else:
            self.interpolator = method(A, B, p=self.p, options=options,
                                       verbose=verbose, ti=ti, **kwargs)

    def get_scale(self):
        """The scale of the 'crf' kind."""
        if self.kind.lower() != 'crf':
            raise NotImplementedError('This function can be called only if '
                                      '"kind" is set to "crf"')
        scale = self.interpolator.scale
        if not isinstance(scale, Number):
            scale = scale[0]
        return scale

    def __call__(self, t):
        return self.interpolate(t)

    @staticmethod
    def _values(t, value):
        if isinstance(t, Number):
            return float(value)
        return numpy.asarray(value, dtype=float).reshape(numpy.shape(t))

    def eval(self, t):
        """Exact |A + tB|_p (estimator calls, no interpolation)."""
        return self._values(t, self.interpolator.eval(
            t if isinstance(t, Number) else numpy.asarray(t, dtype=float)))

    def _compare_with_exact_solution(self, t, schatten):
        if self.kind.lower() == 'ext':
            return schatten, numpy.zeros(numpy.shape(t))
        exact = self.eval(t)
        return exact, 1.0 - schatten / exact

    def interpolate(self, t):
        return self._values(t, self.interpolator.interpolate(
            t if isinstance(t, Number) else numpy.asarray(t, dtype=float)))

    def bound(self, t):
        return self._values(t, self.interpolator.bound(t))

    def upper_bound(self, t):
        return self._values(t, self.interpolator.upper_bound(t))

    def plot(self, t, normalize=True, compare=False):
        raise NotImplementedError('plotting is not part of imate_amd.')
