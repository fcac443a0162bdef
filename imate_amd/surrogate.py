"""Learned surrogate of logdet(K(l) + t I) over a matrix family.

For a family of matrices K(l) with one sparsity pattern (a compactly
supported correlation kernel over its length scale l, see
``sample_matrices.CorrelationFamily``) a sweep over (l, t) asks for many
values of

    y(l, t) = logdet(K(l) + t I),

each of which costs an SLQ run. The surrogate is a 2 -> 128 -> 128 -> 1 MLP
with SiLU activations, trained on the GPU (hand-written FP32 forward,
backward and Adam kernels of ``libslq_hip.so``) on SLQ estimates that the GPU
path itself generates, and evaluated by a BF16 matrix-core inference kernel.

Inputs and target. The inputs are log l and log t mapped affinely to
[-1, 1] over the trained box. The network learns

    g(l, t) = (1 + t) (y / n - log(1 + t)),

which is O(1) over the box and tends to the jitter of the diagonal as t
grows (y = n log(1 + t) + n g / (1 + t)), standardised by the training
mean and deviation. Only n enters the transform; nothing of K(l) does.

SLQ is the fallback and the judge. ``query(..., judge='slq')`` runs SLQ for
every query and replaces each surrogate value that leaves the SLQ
confidence interval by the SLQ value (the evaluation mode). ``judge='trust'``
(the sweep mode) decides without running SLQ first: a query is answered by
the surrogate only where the surrogate's validated residuals nearby stay
inside their SLQ intervals; every other query is re-solved by SLQ. The
residual check uses a validation set of SLQ estimates that training never
saw (``calibrate``): for a query, the k nearest validation points in the
normalised input plane give the root mean square of

    rho_i = (surrogate_i - slq_i) / ci_i,

which is about 1/z = 0.51 (z = 1.96 at 95 %) for a surrogate that matches
the truth (the SLQ noise alone) and grows with the surrogate's own error;
the query is trusted when it is at most ``trust_rms`` (default 1.0: the
surrogate's local error is at most about the SLQ interval). Queries outside
the trained box are never trusted.

No counterpart in imate, whose surrogates over t are the interpolators of
imate/interpolator (not part of this package).
"""

import ctypes
import time

import numpy

from . import _lib
from .linear_operator import AffineMatrixFunction
from .logdet import logdet

__all__ = ['LogdetSurrogate', 'sample_queries']

HIDDEN = 128


def sample_queries(num_scales, num_t, scale_range, t_range, seed,
                   stratified=False):
    """Seeded queries: ``num_scales`` length scales log-uniform over
    ``scale_range``, each with ``num_t`` t values log-uniform over
    ``t_range``. Returns (scales[num_scales], t[num_scales, num_t]).

    ``stratified=True`` draws one value per equal slice of the log range
    (scales, and the t values of each scale), which covers the edges of the
    box as densely as its middle."""
    rng = numpy.random.default_rng(seed)
    a, b = numpy.log(scale_range[0]), numpy.log(scale_range[1])
    c, d = numpy.log(t_range[0]), numpy.log(t_range[1])
    if stratified:
        us = (rng.permutation(num_scales) + rng.random(num_scales)) / num_scales
        ut = (numpy.argsort(rng.random((num_scales, num_t)), axis=1) +
              rng.random((num_scales, num_t))) / num_t
    else:
        us = rng.random(num_scales)
        ut = rng.random((num_scales, num_t))
    ls = numpy.exp(a + (b - a) * us)
    ts = numpy.exp(c + (d - c) * ut)
    return ls, numpy.sort(ts, axis=1)


def init_params(seed=0):
    """He initialisation of the flat parameter vector (layout of the
    kernels: W1[128][2], b1, W2[128][128], b2, W3[128], b3), FP32."""
    rng = numpy.random.default_rng(seed)
    parts = [rng.standard_normal(2 * HIDDEN),          # W1 * sqrt(2 / 2)
             numpy.zeros(HIDDEN),
             rng.standard_normal(HIDDEN * HIDDEN) * numpy.sqrt(2.0 / HIDDEN),
             numpy.zeros(HIDDEN),
             rng.standard_normal(HIDDEN) * numpy.sqrt(1.0 / HIDDEN),
             numpy.zeros(1)]
    return numpy.concatenate(parts).astype(numpy.float32)


class LogdetSurrogate(object):
    """Surrogate of logdet(K(l) + t I) for a ``CorrelationFamily``.

    Parameters
    ----------
    family : CorrelationFamily
    scale_range, t_range : (low, high) of the trained box.
    lanczos_degree, num_samples, seed : SLQ settings of every estimate it
        generates or falls back to (``num_samples`` probes, no early stop).
    confidence_level : of the SLQ intervals that judge the surrogate.
    """

    def __init__(self, family, scale_range=(0.01, 0.1), t_range=(1e-4, 1e2),
                 lanczos_degree=30, num_samples=128, seed=0,
                 confidence_level=0.95, trust_neighbors=8, trust_rms=1.0):
        self.family = family
        self.n = family.n
        self.scale_range = tuple(float(x) for x in scale_range)
        self.t_range = tuple(float(x) for x in t_range)
        self.lanczos_degree = int(lanczos_degree)
        self.num_samples = int(num_samples)
        self.seed = int(seed)
        self.confidence_level = float(confidence_level)
        self.trust_neighbors = int(trust_neighbors)
        self.trust_rms = float(trust_rms)
        self.params = None
        self.g_mean = 0.0
        self.g_std = 1.0
        self.losses = None
        self.val_x = None
        self.val_rho = None
        self._op = None
        self.last_solve = None
        self._lib = _lib.load()
        if self._lib.slq_mlp_num_params() != 2 * HIDDEN + HIDDEN + \
                HIDDEN * HIDDEN + 2 * HIDDEN + 1:
            raise RuntimeError('libslq_hip.so has another network layout.')

    # ------------------------------------------------------------ transforms

    def _inputs(self, scales, t):
        def norm(v, lo, hi):
            a, b = numpy.log(lo), numpy.log(hi)
            return 2.0 * (numpy.log(v) - a) / (b - a) - 1.0
        x = numpy.empty((len(scales), 2), dtype=numpy.float32)
        x[:, 0] = norm(scales, *self.scale_range)
        x[:, 1] = norm(t, *self.t_range)
        return x

    def _to_g(self, y, t):
        return (1.0 + t) * (y / self.n - numpy.log1p(t))

    def _from_g(self, g, t):
        return self.n * (numpy.log1p(t) + g / (1.0 + t))

    def in_box(self, scales, t):
        s0, s1 = self.scale_range
        t0, t1 = self.t_range
        return (scales >= s0) & (scales <= s1) & (t >= t0) & (t <= t1)

    # ------------------------------------------------------------ SLQ

    def _operator(self, scale):
        # one upload of the pattern and the family's coefficients; every
        # member's values are then made on the device
        if self._op is None:
            self._op = AffineMatrixFunction(self.family.matrix(scale))
            self._op.attach_family(*self.family.coefficients())
        self._op.set_family_scale(scale)
        return self._op

    def solve(self, scales, t, seed=None):
        """SLQ estimates of logdet(K(l) + t I): one Lanczos run per distinct
        l, its Ritz values shifted to every t of that l. ``scales`` and ``t``
        are flat arrays of equal length. Returns (y, ci, seconds) with ci the
        half-width of the SLQ confidence interval."""
        scales = numpy.asarray(scales, dtype=numpy.float64)
        t = numpy.asarray(t, dtype=numpy.float64)
        y = numpy.empty(len(t))
        ci = numpy.empty(len(t))
        t0 = time.perf_counter()
        uniq, inv = numpy.unique(scales, return_inverse=True)
        values_sec = 0.0
        for k, s in enumerate(uniq):
            idx = numpy.nonzero(inv == k)[0]
            tv = time.perf_counter()
            op = self._operator(s)
            values_sec += time.perf_counter() - tv
            sd = (self.seed if seed is None else seed) + k
            val, info = logdet(
                op, parameters=t[idx], method='slq', gpu=True,
                lanczos_degree=self.lanczos_degree,
                min_num_samples=self.num_samples,
                max_num_samples=self.num_samples,
                confidence_level=self.confidence_level,
                error_rtol=0.0, seed=sd, return_info=True)
            y[idx] = val
            ci[idx] = info['error']['absolute_error']
        sec = time.perf_counter() - t0
        # time spent computing and uploading new values vs running SLQ
        self.last_solve = {'seconds': sec, 'values_seconds': values_sec,
                           'slq_seconds': sec - values_sec,
                           'lanczos_runs': len(uniq)}
        return y, ci, sec

    def generate(self, num_scales, num_t, seed, stratified=True):
        """Training or validation data: SLQ estimates at seeded queries
        (stratified over the box by default)."""
        ls, ts = sample_queries(num_scales, num_t, self.scale_range,
                                self.t_range, seed, stratified=stratified)
        scales = numpy.repeat(ls, num_t)
        t = ts.ravel()
        y, ci, sec = self.solve(scales, t, seed=seed * 1000 + 1)
        return {'scales': scales, 't': t, 'y': y, 'ci': ci, 'seconds': sec}

    # ------------------------------------------------------------ network

    def train(self, data, steps=(10000, 10000, 10000, 10000),
              lr=(3e-3, 1e-3, 3e-4, 1e-4), init_seed=0):
        """Full-batch Adam on the GPU, one stage per (steps, lr) pair (each
        stage one captured step replayed ``steps`` times). Returns seconds."""
        x = self._inputs(data['scales'], data['t'])
        g = self._to_g(data['y'], data['t'])
        self.g_mean = float(g.mean())
        self.g_std = float(g.std()) or 1.0
        target = ((g - self.g_mean) / self.g_std).astype(numpy.float32)
        p = init_params(init_seed)
        B = len(target)
        losses = []
        t0 = time.perf_counter()
        for st, rate in zip(steps, lr):
            ls = numpy.zeros(st, dtype=numpy.float32)
            _lib.check(self._lib.slq_mlp_train(
                _lib.ptr(p), _lib.ptr(x), _lib.ptr(target), B, int(st),
                ctypes.c_float(rate), _lib.ptr(ls)), 'slq_mlp_train')
            losses.append(ls)
        sec = time.perf_counter() - t0
        self.params = p
        self.losses = numpy.concatenate(losses)
        return sec

    def predict(self, scales, t, bf16=True):
        """Surrogate logdet values (float64) and the inference kernel's time
        in seconds (HIP events around the one launch)."""
        if self.params is None:
            raise RuntimeError('the surrogate is not trained.')
        scales = numpy.asarray(scales, dtype=numpy.float64)
        t = numpy.asarray(t, dtype=numpy.float64)
        x = self._inputs(scales, t)
        out = numpy.zeros(len(t), dtype=numpy.float32)
        sec = ctypes.c_double(0)
        _lib.check(self._lib.slq_mlp_predict(
            _lib.ptr(self.params), _lib.ptr(x), len(t), int(bool(bf16)),
            _lib.ptr(out), ctypes.byref(sec)), 'slq_mlp_predict')
        g = out.astype(numpy.float64) * self.g_std + self.g_mean
        return self._from_g(g, t), sec.value

    # ------------------------------------------------------------ trust

    def calibrate(self, data, bf16=True):
        """Residual ratios of the surrogate on validation SLQ data (not used
        in training); they drive ``trusted``. Returns the ratios."""
        pred, _ = self.predict(data['scales'], data['t'], bf16=bf16)
        self.val_x = self._inputs(data['scales'], data['t']).astype(
            numpy.float64)
        self.val_rho = (pred - data['y']) / data['ci']
        return self.val_rho

    def trusted(self, scales, t):
        """Mask of queries the surrogate answers without SLQ."""
        if self.val_rho is None:
            raise RuntimeError('calibrate() the surrogate first.')
        scales = numpy.asarray(scales, dtype=numpy.float64)
        t = numpy.asarray(t, dtype=numpy.float64)
        x = self._inputs(scales, t).astype(numpy.float64)
        k = min(self.trust_neighbors, len(self.val_rho))
        ok = numpy.zeros(len(t), dtype=bool)
        rho2 = self.val_rho ** 2
        for lo in range(0, len(t), 4096):
            xb = x[lo:lo + 4096]
            d2 = ((xb[:, None, :] - self.val_x[None, :, :]) ** 2).sum(axis=2)
            nn = numpy.argpartition(d2, k - 1, axis=1)[:, :k]
            rms = numpy.sqrt(rho2[nn].mean(axis=1))
            ok[lo:lo + 4096] = rms <= self.trust_rms
        return ok & self.in_box(scales, t)

    # ------------------------------------------------------------ queries

    def query(self, scales, t, judge='trust', bf16=True, seed=None):
        """logdet(K(l) + t I) for every (l, t) pair.

        judge='trust': surrogate where ``trusted``, SLQ elsewhere.
        judge='slq':   SLQ for every query; the surrogate value is kept where
                       it lies inside the SLQ confidence interval and
                       replaced by the SLQ value where it leaves it.
        Returns (y, info) with the fallback mask and the times spent.
        """
        scales = numpy.asarray(scales, dtype=numpy.float64)
        t = numpy.asarray(t, dtype=numpy.float64)
        t0 = time.perf_counter()
        pred, kernel_sec = self.predict(scales, t, bf16=bf16)
        predict_sec = time.perf_counter() - t0
        info = {'surrogate': pred, 'kernel_seconds': kernel_sec,
                'predict_seconds': predict_sec}
        if judge == 'slq':
            y_slq, ci, sec = self.solve(scales, t, seed=seed)
            fallback = numpy.abs(pred - y_slq) > ci
            y = numpy.where(fallback, y_slq, pred)
            info.update(slq=y_slq, ci=ci, slq_seconds=sec)
        elif judge == 'trust':
            t1 = time.perf_counter()
            fallback = ~self.trusted(scales, t)
            info['trust_seconds'] = time.perf_counter() - t1
            y = pred.copy()
            sec = 0.0
            if fallback.any():
                y_slq, ci, sec = self.solve(scales[fallback], t[fallback],
                                            seed=seed)
                y[fallback] = y_slq
            info['slq_seconds'] = sec
        else:
            raise ValueError('judge should be "trust" or "slq".')
        info['fallback'] = fallback
        return y, info
