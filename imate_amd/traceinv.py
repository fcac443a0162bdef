"""``traceinv(A, gram, p, return_info, method, **options)`` (imate/traceinv/traceinv.py:422-441).

tr(A^-p) by stochastic Lanczos quadrature with f(x) = 1/x applied to theta^p
(imate/functions/inverse.cpp:41-44). Only ``method='slq'`` is part of this
build.
"""

from ._slq import slq_method


def traceinv(A, gram=False, p=1, return_info=False, method='cholesky',
             **options):
    if method == 'slq':
        return slq_method('inv', A, gram=gram, p=p, return_info=return_info,
                          **options)
    if method in ('eigenvalue', 'cholesky', 'hutchinson'):
        raise NotImplementedError(
            'Method "%s" is not part of imate_amd; use method="slq".' % method)
    raise ValueError('Method "%s" is invalid.' % (method))
