"""Batched SLQ over many independent matrices (the batched form of this package).

imate estimates one operator per call; a user with many matrices loops over
``logdet(A_i, method='slq', gpu=True)``. ``slq_batch`` runs them together:
every kernel of every problem is captured into one hipGraph (one branch per
problem, so problems of uneven size run side by side) and launched once,
after an autotuner has picked each problem's launch grid. ``use_graph=False``
runs the same kernels problem after problem (the per-problem launch loop it
is compared against). Many t values of ``A + tI`` need no batching API: one
``logdet(AffineMatrixFunction(A), parameters=t, ...)`` call runs one Lanczos
per probe and shifts its Ritz values to every t.
"""

import ctypes

import numpy

from . import _lib
from .linear_operator import Matrix, LinearOperator

__all__ = ['slq_batch']

_FUNCS = {'log': _lib.FUNC_LOG, 'inv': _lib.FUNC_INV,
          'identity': _lib.FUNC_IDENT}


def slq_batch(matrices, func='log', gram=False, p=1.0, lanczos_degree=20,
              num_samples=64, orthogonalize=0, seed=0, lanczos_tol=None,
              confidence_level=0.95, outlier_significance_level=0.001,
              precision=None, use_graph=True, autotune=True):
    """Trace estimates of f(A_i^p) for every matrix A_i.

    Returns ``(trace, info)`` with ``trace[i]`` and ``info`` holding
    ``error``, ``samples`` ([count, num_samples]), ``run_time`` (launch to
    completion, seconds), ``setup_time`` and the autotuned grids (empty
    when nothing was tuned: Golub-Kahn, a rectangular matrix or
    ``autotune=False``). Tuned grids are used by this batch only; later
    estimates on the same matrices run with the default grids, so their
    results do not change bit for bit after a batch.
    """
    if func not in _FUNCS:
        raise ValueError('"func" should be one of %s.' % list(_FUNCS))
    if not isinstance(num_samples, (int, numpy.integer)) or num_samples < 1:
        raise ValueError('"num_samples" should be a positive integer.')
    if not isinstance(lanczos_degree, (int, numpy.integer)) or \
            lanczos_degree < 1:
        raise ValueError('"lanczos_degree" should be at least one.')
    prec = {None: _lib.PREC_NATIVE, 'native': _lib.PREC_NATIVE,
            'mixed': _lib.PREC_MIXED}[precision]
    lib = _lib.load()
    ops = []
    dtype = None
    for A in matrices:
        M = A if isinstance(A, LinearOperator) else Matrix(A)
        if M.get_num_parameters() != 0:
            raise ValueError('slq_batch takes plain matrices.')
        dt = M.get_data_type_name()
        dtype = dt if dtype is None else dtype
        ops.append(M)
    if not ops:
        raise ValueError('no matrices given.')
    if lanczos_tol is None:
        lanczos_tol = float(numpy.finfo(numpy.dtype(dtype.decode())).eps)
    handles = [M.get_linear_operator(gpu=True).handle for M in ops]
    count = len(handles)
    if len(set(h.value for h in handles)) != count:
        raise ValueError('each problem needs its own Matrix object.')
    grids = []
    # the autotuner times Lanczos steps, so Golub-Kahn runs (gram) and
    # rectangular matrices keep the default grids
    square = all(M.get_num_rows() == M.get_num_columns() for M in ops)
    if autotune and not gram and square:
        for h in handles:
            sg = ctypes.c_int32(0)
            ug = ctypes.c_int32(0)
            _lib.check(lib.slq_autotune(h, prec, ctypes.byref(sg),
                                        ctypes.byref(ug)), 'slq_autotune')
            grids.append((sg.value, ug.value))
    arr = (ctypes.c_void_p * count)(*[h.value for h in handles])
    trace = numpy.zeros(count)
    error = numpy.zeros(count)
    samples = numpy.zeros((count, num_samples))
    run_time = ctypes.c_double(0)
    setup_time = ctypes.c_double(0)
    _lib.check(lib.slq_estimate_batch(
        ctypes.cast(arr, ctypes.c_void_p), count, _FUNCS[func], None,
        int(gram),
        float(p), int(orthogonalize), int(seed), int(lanczos_degree),
        float(lanczos_tol), int(num_samples), float(confidence_level),
        float(outlier_significance_level), prec, int(bool(use_graph)),
        _lib.ptr(trace), _lib.ptr(error), _lib.ptr(samples),
        ctypes.byref(run_time), ctypes.byref(setup_time)),
        'slq_estimate_batch')
    info = {'error': error, 'samples': samples, 'run_time': run_time.value,
            'setup_time': setup_time.value, 'grids': grids,
            'operators': ops}
    return trace, info
