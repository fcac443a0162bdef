"""imate_amd — stochastic Lanczos quadrature (SLQ) trace estimation on AMD
Instinct MI355X (gfx950), a drop-in for imate's
``logdet / traceinv / trace / schatten / trexp / trlinfrac / eigencount /
density(A, method='slq', gpu=True, ...)``.

The numerical work runs in hand-written HIP kernels (``csrc/``) behind the
C ABI of ``include/slq_hip.h``; this package is the Python side of that
boundary, with imate's call surface, defaults, validation and ``info`` dict.
"""

from .__version__ import __version__
from .linear_operator import LinearOperator, Matrix, AffineMatrixFunction
from .logdet import logdet
from .traceinv import traceinv
from .trace import trace
from .schatten import schatten
from .trexp import trexp
from .trlinfrac import trlinfrac
from .eigencount import eigencount
from .density import density
from .interpolator import InterpolateSchatten, InterpolateLogdet, \
    InterpolateTrace
from . import sample_matrices
from ._lib import device_properties

__all__ = ['logdet', 'traceinv', 'trace', 'schatten', 'trexp', 'trlinfrac',
           'eigencount', 'density', 'InterpolateSchatten',
           'InterpolateLogdet', 'InterpolateTrace', 'LinearOperator',
           'Matrix', 'AffineMatrixFunction', 'sample_matrices',
           'device_properties', '__version__']
