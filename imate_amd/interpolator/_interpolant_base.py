"""Base of the interpolation kinds: the Schatten norm of A + tB at the
interpolation points, which every kind fits.

Follows imate/interpolator/_interpolant_base.py:67-269. The values come from
this package's SLQ path: A and B go to the device once, as one
``AffineMatrixFunction``, and each batch of points is one estimator call with
``parameters`` set to the points. With B = I one Lanczos run per probe serves
every point through a shift of its Ritz values (imate's known-eigenvalue
relation of A + tI); with a general B each point runs its own recurrence with
the fused A x + t B x SpMM. imate forms A + t_i B and calls ``schatten`` once
per point instead; with the same seed both see the same probes, and when the
sample count is fixed (``min_num_samples == max_num_samples``) the values
agree to rounding. With an adaptive sample count a batch stops when every
point has converged, so all points of a batch use the same number of samples.
"""

from numbers import Number

import numpy
import scipy.sparse

from ..linear_operator import AffineMatrixFunction, Matrix
from ..schatten import schatten


class InterpolantBase(object):
    """Data of an interpolation of t -> |A + tB|_p.

    Attributes follow imate's: ``schatten_A``, ``schatten_B``, ``tau0`` (=
    schatten_A / schatten_B), ``t_i``, ``q``, ``schatten_i``, ``tau_i`` (=
    schatten_i / schatten_B) and ``scale_t`` (max t_i when it exceeds 1).
    """

    def __init__(self, A, B=None, p=2, ti=None, options={}, verbose=False):
        if 'parameters' in options:
            raise ValueError('"options" cannot set "parameters": the '
                             'interpolation points are the parameters.')
        self.verbose = verbose
        self.A = A
        self.n = A.shape[0]
        self.use_sparse = scipy.sparse.issparse(A)
        self.B_is_identity = B is None
        self.B = B
        self.p = p
        self.options = dict(options)
        self._affine = None

        self.schatten_A = None
        self.schatten_B = None
        self.tau0 = None
        self.t_i = None
        self.q = None
        self.schatten_i = None
        self.tau_i = None
        self.scale_t = None

        if ti is None:
            return
        if isinstance(ti, Number):
            ti = [ti]
        ti = numpy.asarray(ti, dtype=float).ravel()
        self._compute_schatten_of_input_matrices(ti)
        if ti.size > 0:
            self.scale_t = self.find_scale_ti()

    # ------------------------------------------------------------ device

    def _operator(self):
        """A + tB on the device (built on first use, kept for every call)."""
        if self._affine is None:
            B = None if self.B_is_identity else self.B
            self._affine = AffineMatrixFunction(self.A, B)
        return self._affine

    def _schatten_at(self, t):
        """|A + t_k B|_p for every t_k of the array ``t`` (one call)."""
        t = numpy.atleast_1d(numpy.asarray(t, dtype=float))
        if t.size == 0:
            return numpy.zeros(0)
        values = schatten(self._operator(), p=self.p, parameters=t,
                          **self.options)
        return numpy.atleast_1d(numpy.asarray(values, dtype=float)).copy()

    def _compute_schatten(self, A, p):
        """|A|_p of a plain matrix (imate's _compute_schatten)."""
        return float(schatten(Matrix(A), p=p, **self.options))

    # ------------------------------------------------------ points, tau0

    def _compute_schatten_of_input_matrices(self, ti=None):
        """schatten_A, schatten_B, tau0 and, when ``ti`` is given, the values
        at the points: schatten_A is the point t = 0 of the same batch."""
        if self.B_is_identity:
            self.schatten_B = 1.0
        else:
            self.schatten_B = self._compute_schatten(self.B, self.p)
        if ti is None or ti.size == 0:
            self.schatten_A = float(self._schatten_at([0.0])[0])
        else:
            if self.verbose:
                print('Evaluate function at interpolant points ...', end='')
            values = self._schatten_at(numpy.r_[0.0, ti])
            if self.verbose:
                print(' Done.')
            self.schatten_A = float(values[0])
            self.t_i = ti
            self.q = ti.size
            self.schatten_i = values[1:]
            self.tau_i = self.schatten_i / self.schatten_B
        self.tau0 = self.schatten_A / self.schatten_B

    def find_scale_ti(self):
        """max(t_i) when it exceeds 1, else 1: the orthogonal bases are
        defined on [0, 1]."""
        scale_t = 1.0
        if self.t_i is not None and self.t_i.size > 0:
            if numpy.max(self.t_i) > 1.0:
                scale_t = float(numpy.max(self.t_i))
        return scale_t

    # ------------------------------------------------------ exact, bounds

    def eval(self, t):
        """|A + tB|_p without interpolation (one estimator call for all t)."""
        values = self._schatten_at(t)
        return float(values[0]) if numpy.isscalar(t) else values

    def _compute(self, t):
        return self.eval(t)

    def upper_bound(self, t):
        """For p = -1, the norm's upper bound (tr(A + tB) / n) from the
        arithmetic-harmonic mean inequality tr((A + tB)^-1) >= n^2 /
        tr(A + tB)."""
        if self.p != -1:
            raise ValueError('Lower bound is only available for "p=-1".')
        trace_A = float(self.A.diagonal().sum())
        if self.B_is_identity:
            trace_B = float(self.n)
        else:
            trace_B = float(self.B.diagonal().sum())
        trace_ub = self.n**2 / (trace_A + numpy.asarray(t) * trace_B)
        return (trace_ub / self.n)**(1.0 / self.p)

    def bound(self, t):
        """(tau0 + t) schatten_B: a lower bound of the norm for p <= 0
        (det^(1/n) and (tr X^p / n)^(1/p), p < 0, are concave and
        homogeneous, hence superadditive) and an upper bound for p >= 1
        (Minkowski)."""
        if self.tau0 is None:
            self._compute_schatten_of_input_matrices()
        return (self.tau0 + numpy.asarray(t)) * self.schatten_B
