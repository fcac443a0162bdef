"""``schatten(A, gram, p, return_info, method, **options)``
(imate/schatten/schatten.py:23-30,422-467): the Schatten p-norm normalised by
the size n, |A|_p = (tr(A^p)/n)^(1/p), with the p = 0 limit exp(logdet/n)
and p < 0 through traceinv."""

import numpy

from .logdet import logdet
from .trace import trace
from .traceinv import traceinv


def schatten(A, gram=False, p=2.0, return_info=False, method='eigenvalue',
             **options):
    if not isinstance(p, (int, numpy.integer, float)):
        raise ValueError('"p" should be float or integer.')

    n = A.shape[0] if hasattr(A, 'shape') else A.get_num_rows()

    if p == 0:
        output = logdet(A, gram=gram, p=1.0, return_info=return_info,
                        method=method, **options)
        value = output[0] if return_info else output
        result = numpy.exp(value / n)
    elif p > 0:
        output = trace(A, gram=gram, p=p, return_info=return_info,
                       method=method, **options)
        value = output[0] if return_info else output
        result = (value / n)**(1.0 / p)
    else:
        output = traceinv(A, gram=gram, p=numpy.abs(p),
                          return_info=return_info, method=method, **options)
        value = output[0] if return_info else output
        result = (value / n)**(1.0 / p)

    if return_info:
        return result, output[1]
    return result
