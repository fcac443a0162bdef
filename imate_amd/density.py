"""``density(A, gram, p, mu, sigma, return_info, method, **options)``
(imate/density/density.py:22-398).

The spectral density of A^p at mu smoothed by a Gaussian of width sigma,
tr g(A^p) with g(x) = exp(-(x - mu)^2 / (2 sigma^2)) / (sigma sqrt(2 pi)),
by stochastic Lanczos quadrature (imate/functions/gaussian.cpp;
density/_slq_method.pyx:25-897). Only ``method='slq'`` is part of this
build. As in the reference, one call evaluates one scalar mu.
"""

import numbers

import numpy

from ._slq import slq_method


def _check_arguments(mu, sigma, return_info):
    """density/_slq_method.pyx:881-897, and a scalar mu, the only form the
    reference's Gaussian(mu, sigma) constructor accepts."""
    if (not isinstance(mu, (int, float))) and \
            (not (isinstance(mu, numpy.ndarray) and (mu.ndim == 1))) and \
            (not isinstance(mu, (list, tuple))):
        raise TypeError('"mu" should be either a real scalar or a real ' +
                        '1D array-like, such as list, tuple, or numpy array.')
    if not isinstance(mu, numbers.Real):
        raise TypeError('"mu" should be a real scalar.')
    if not isinstance(sigma, (int, float)):
        raise TypeError('"sigma" should be a scalar real number.')
    elif sigma <= 0.0:
        raise ValueError('"sigma" should be positive.')
    if not isinstance(return_info, bool):
        raise TypeError('"return_info" should be boolean.')


def density(A, gram=False, p=1.0, mu=1.0, sigma=1.0, return_info=False,
            method='eigenvalue', **options):
    if method == 'slq':
        _check_arguments(mu, sigma, return_info)
        return slq_method(('gaussian', [float(mu), float(sigma)]), A,
                          gram=gram, p=p, return_info=return_info, **options)
    if method == 'eigenvalue':
        raise NotImplementedError(
            'Method "%s" is not part of imate_amd; use method="slq".' % method)
    raise ValueError('Method "%s" is invalid.' % (method))
