"""``trace(A, gram, p, return_info, method, **options)`` (imate/trace/trace.py:391-404).

tr(A^p) by stochastic Lanczos quadrature with f(x) = x applied to theta^p
(imate/functions/identity.cpp:41-44). Only ``method='slq'`` is part of this
build.
"""

from ._slq import slq_method


def trace(A, gram=False, p=1.0, return_info=False, method='exact',
          **options):
    if method == 'slq':
        return slq_method('identity', A, gram=gram, p=p,
                          return_info=return_info, **options)
    if method in ('exact', 'eigenvalue'):
        raise NotImplementedError(
            'Method "%s" is not part of imate_amd; use method="slq".' % method)
    raise ValueError('Method "%s" is invalid.' % (method))
