"""``eigencount(A, gram, p, interval, return_info, method, **options)``
(imate/eigencount/eigencount.py:22-383).

The number of eigenvalues of A^p in [a, b], tr 1_[a,b](A^p), by stochastic
Lanczos quadrature with the indicator function (imate/functions/indicator.cpp:
0 below a or above b, else 1; eigencount/_slq_method.pyx:25-877). Only
``method='slq'`` is part of this package.
"""

import numpy

from ._slq import slq_method


def _check_arguments(interval, return_info):
    """eigencount/_slq_method.pyx:864-877."""
    if not isinstance(interval, (tuple, list, numpy.ndarray)):
        raise TypeError('"interval" should be a tuple,  list, or numpy array.')
    elif len(interval) != 2:
        raise ValueError('"interval" should be an array or list of length 2.')
    if not isinstance(return_info, bool):
        raise TypeError('"return_info" should be boolean.')


def eigencount(A, gram=False, p=1.0, interval=[0.0, 1.0], return_info=False,
               method='eigenvalue', **options):
    if method == 'slq':
        _check_arguments(interval, return_info)
        return slq_method(('indicator', [float(interval[0]),
                                         float(interval[1])]), A, gram=gram,
                          p=p, return_info=return_info, **options)
    if method == 'eigenvalue':
        raise NotImplementedError(
            'Method "%s" is not part of imate_amd; use method="slq".' % method)
    raise ValueError('Method "%s" is invalid.' % (method))
