"""``InterpolateLogdet``: t -> log det(A + tB) through the p = 0 Schatten norm,
logdet = n log |A + tB|_0 (imate/interpolator/interpolate_logdet.py:30-811,
without plotting)."""

import numpy

from .interpolate_schatten import InterpolateSchatten

__all__ = ['InterpolateLogdet']


class InterpolateLogdet(InterpolateSchatten):

    def __init__(self, A, B=None, options={}, verbose=False, kind='imbf',
                 ti=[], **kwargs):
        super(InterpolateLogdet, self).__init__(
            A, B=B, p=0, ti=ti, kind=kind, verbose=verbose, options=options,
            **kwargs)

    def _schatten_to_logdet(self, schatten):
        return self.n * numpy.log(schatten)

    def _logdet_to_schatten(self, logdet_):
        return numpy.exp(logdet_ / self.n)

    def eval(self, t):
        return self._schatten_to_logdet(
            super(InterpolateLogdet, self).eval(t))

    def _compare_with_exact_solution(self, t, logdet_):
        if self.kind.lower() == 'ext':
            return logdet_, numpy.zeros(numpy.shape(t))
        exact = self.eval(t)
        return exact, (logdet_ - exact) / exact

    def interpolate(self, t):
        return self._schatten_to_logdet(
            super(InterpolateLogdet, self).interpolate(t))

    def lower_bound(self, t):
        """n log(|A|_0 + t |B|_0) (Brunn-Minkowski)."""
        return self._schatten_to_logdet(
            super(InterpolateLogdet, self).bound(t))
