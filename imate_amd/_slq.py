"""``slq_method`` of logdet / traceinv / trace: choose f, call the estimator.

Mirrors imate/logdet/_slq_method.pyx:24-44,820-854 (f = Logarithm),
imate/traceinv/_slq_method.pyx:827 (f = Inverse) and
imate/trace/_slq_method.pyx:825 (f = Identity): same keyword arguments and
defaults; ``precision``, ``block_size`` and ``krylov`` are options of this
package whose defaults keep the reference behaviour.
"""

from ._trace_estimator import trace_estimator


def _check_return_info(return_info):
    if not isinstance(return_info, bool):
        raise TypeError('"return_info" should be boolean.')


def slq_method(func, A, gram=False, p=1.0, return_info=False,
               parameters=None, min_num_samples=10, max_num_samples=50,
               error_atol=None, error_rtol=1e-2, confidence_level=0.95,
               outlier_significance_level=0.001, lanczos_degree=20,
               lanczos_tol=None, orthogonalize=0, seed=None, num_threads=0,
               num_gpu_devices=0, gpu=False, verbose=False, plot=False,
               precision=None, block_size=0, krylov=None):
    _check_return_info(return_info)
    trace, info = trace_estimator(
        A, parameters, func, gram, p, min_num_samples, max_num_samples,
        error_atol, error_rtol, confidence_level, outlier_significance_level,
        lanczos_degree, lanczos_tol, orthogonalize, seed, num_threads,
        num_gpu_devices, verbose, plot, gpu, precision=precision,
        block_size=block_size, krylov=krylov)
    if return_info:
        return trace, info
    return trace
