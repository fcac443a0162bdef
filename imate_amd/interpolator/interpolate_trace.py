"""``InterpolateTrace``: t -> trace((A + tB)^p) through the Schatten norm,
trace = n |A + tB|_p^p, p != 0 (imate/interpolator/interpolate_trace.py:
30-970, without plotting)."""

import numpy

from .interpolate_schatten import InterpolateSchatten

__all__ = ['InterpolateTrace']


class InterpolateTrace(InterpolateSchatten):

    def __init__(self, A, B=None, p=-1, options={}, verbose=False,
                 kind='imbf', ti=[], **kwargs):
        if p == 0:
            raise ValueError('"p" cannot be zero in "InterpolateTrace". '
                             '"p=0" use "InterpolateLogdet" class.')
        super(InterpolateTrace, self).__init__(
            A, B=B, p=p, ti=ti, kind=kind, verbose=verbose, options=options,
            **kwargs)

    def _schatten_to_tracep(self, schatten):
        return schatten**self.p * self.n

    def _tracep_to_schatten(self, tracep):
        return (tracep / self.n)**(1.0 / self.p)

    def eval(self, t):
        return self._schatten_to_tracep(
            super(InterpolateTrace, self).eval(t))

    def _compare_with_exact_solution(self, t, tracep):
        if self.kind.lower() == 'ext':
            return tracep, numpy.zeros(numpy.shape(t))
        exact = self.eval(t)
        return exact, (tracep - exact) / exact

    def interpolate(self, t):
        return self._schatten_to_tracep(
            super(InterpolateTrace, self).interpolate(t))

    def upper_bound(self, t):
        """For p = -1: n / (|A|_-1 + t |B|_-1), from the norm's lower
        bound."""
        return self._schatten_to_tracep(
            super(InterpolateTrace, self).bound(t))

    def lower_bound(self, t):
        """For p = -1: n^2 / tr(A + tB) (arithmetic-harmonic means)."""
        return self._schatten_to_tracep(
            super(InterpolateTrace, self).upper_bound(t))
