"""``trexp(A, gram, p, coeff, return_info, method, **options)``
(imate/trexp/trexp.py:22-376).

tr exp(c A^p) by stochastic Lanczos quadrature with f(x) = exp(c x)
(imate/functions/exponential.cpp; trexp/_slq_method.pyx:24-852). Only
``method='slq'`` is part of this package.
"""

from ._slq import slq_method


def _check_return_info(return_info):
    if not isinstance(return_info, bool):
        raise TypeError('"return_info" should be boolean.')


def trexp(A, gram=False, p=1.0, coeff=1.0, return_info=False,
          method='eigenvalue', **options):
    if method == 'slq':
        _check_return_info(return_info)
        return slq_method(('exp', [float(coeff)]), A, gram=gram, p=p,
                          return_info=return_info, **options)
    if method == 'eigenvalue':
        raise NotImplementedError(
            'Method "%s" is not part of imate_amd; use method="slq".' % method)
    raise ValueError('Method "%s" is invalid.' % (method))
