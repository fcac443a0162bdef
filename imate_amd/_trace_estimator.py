"""SLQ orchestration of the drop-in: argument checks, operator coercion,
output allocation, the call into ``slq_estimate`` and the ``info`` dict.

Mirrors imate/_trace_estimator/trace_estimator.pyx:38-390 and
trace_estimator_utilities.py:24-571 (same argument semantics, defaults,
validation messages, return shapes and ``info`` schema). Both branches call
the library through its C ABI: ``gpu=True`` the HIP path (``slq_estimate``),
``gpu=False`` its host implementation (``slq_estimate_host``).
"""

import re
import time

import numpy
import scipy.sparse

from . import _lib
from .__version__ import __version__
from .linear_operator import LinearOperator, Matrix

__all__ = ['trace_estimator']

FUNCTIONS = {'log': _lib.FUNC_LOG, 'inv': _lib.FUNC_INV,
             'identity': _lib.FUNC_IDENT, 'exp': _lib.FUNC_EXP,
             'linfrac': _lib.FUNC_LINFRAC, 'indicator': _lib.FUNC_INDICATOR,
             'gaussian': _lib.FUNC_GAUSSIAN}


def function_spec(func):
    """(code, coefficients) of a function given as a name ('log', 'inv',
    'identity') or as (name, coefficients) for 'exp' [c], 'linfrac'
    [a, b, c, d], 'indicator' [a, b] and 'gaussian' [mu, sigma]."""
    if isinstance(func, str):
        name, coeff = func, None
    else:
        name, coeff = func
    if name not in FUNCTIONS:
        raise ValueError('unknown function "%s".' % name)
    fp = numpy.zeros(4, dtype=numpy.float64)
    if coeff is not None:
        c = numpy.atleast_1d(numpy.asarray(coeff, dtype=numpy.float64))
        fp[:len(c)] = c
    return FUNCTIONS[name], fp

PRECISIONS = {None: _lib.PREC_NATIVE, 'native': _lib.PREC_NATIVE,
              'mixed': _lib.PREC_MIXED, 'fp32': _lib.PREC_MIXED}


def get_operator(A, gram):
    """trace_estimator_utilities.py:24-55."""
    if isinstance(A, numpy.ndarray) or scipy.sparse.issparse(A):
        if A.ndim != 2:
            raise ValueError('Input matrix should be a 2-dimensional array.')
        elif (not gram) and (A.shape[0] != A.shape[1]):
            raise ValueError('Input matrix should be a square matrix.')
        return Matrix(A)
    elif isinstance(A, LinearOperator):
        if (not gram) and (A.get_num_rows() != A.get_num_columns()):
            raise ValueError('Input operator should have the same number ' +
                             'of rows and columns.')
        return A
    else:
        raise TypeError('The linear operator "A" should be either a ' +
                        'numpy.ndarray, a scipy.sparse array, or an ' +
                        'instance of a class derived from the ' +
                        '"LinearOperator" class.')


def get_operator_parameters(parameters, data_type_name):
    """trace_estimator_utilities.py:62-101."""
    if parameters is None:
        parameters_size = 0
    elif numpy.isscalar(parameters):
        parameters_size = 1
    elif isinstance(parameters, list):
        parameters = numpy.array(parameters, dtype=data_type_name)
        if parameters.ndim != 1:
            raise ValueError('"parameters", if given as an array, should be ' +
                             'a one-dimensional array. Current dimension of ' +
                             'parameters is %d.' % parameters.ndim)
        else:
            parameters_size = parameters.size
    elif isinstance(parameters, numpy.ndarray):
        if parameters.ndim != 1:
            raise ValueError('"parameters", if given as an array, should be ' +
                             'a one-dimensional array. Current dimension of ' +
                             'parameters is %d.' % parameters.ndim)
        else:
            parameters_size = parameters.size
    else:
        raise TypeError('"parameters" type should be either "None", a ' +
                        'scalar, a flat list, or one-dimensional ' +
                        '"numpy.ndarray.')
    return parameters, parameters_size


def check_arguments(gram, exponent, min_num_samples, max_num_samples,
                    error_atol, error_rtol, confidence_level,
                    outlier_significance_level, lanczos_degree, lanczos_tol,
                    orthogonalize, seed, num_threads, num_gpu_devices, verbose,
                    plot, gpu):
    """trace_estimator_utilities.py:108-305 (same checks and messages)."""
    if gram is None:
        raise TypeError('"gram" cannot be None.')
    elif not numpy.isscalar(gram):
        raise TypeError('"gram" should be a scalar value.')
    elif not isinstance(gram, bool):
        raise TypeError('"gram" should be boolean.')

    if exponent is None:
        raise TypeError('"exponent" cannot be None.')
    elif not numpy.isscalar(exponent):
        raise TypeError('"exponent" should be a scalar value.')
    elif isinstance(exponent, complex):
        raise TypeError('"exponent" cannot be a complex number.')

    if min_num_samples is None:
        raise ValueError('"min_num_samples" cannot be None.')
    elif not numpy.isscalar(min_num_samples):
        raise TypeError('"min_num_samples" should be a scalar value.')
    elif not isinstance(min_num_samples, (int, numpy.integer)):
        raise TypeError('"min_num_samples" should be an integer.')
    elif min_num_samples < 1:
        raise ValueError('"min_num_samples" should be at least one.')

    if max_num_samples is None:
        raise ValueError('"max_num_samples" cannot be None.')
    elif not numpy.isscalar(max_num_samples):
        raise TypeError('"max_num_samples" should be a scalar value.')
    elif not isinstance(max_num_samples, (int, numpy.integer)):
        raise TypeError('"max_num_samples" should be an integer.')
    elif max_num_samples < 1:
        raise ValueError('"max_num_samples" should be at least one.')

    if min_num_samples > max_num_samples:
        raise ValueError('"min_num_samples" cannot be greater than ' +
                         '"max_num_samples".')

    if error_atol is None:
        error_atol = 0.0
    elif not numpy.isscalar(error_atol):
        raise TypeError('"error_atol" should be a scalar value.')
    elif not isinstance(error_atol, (int, float)):
        raise TypeError('"error_atol" should be a float number.')
    elif error_atol < 0.0:
        raise ValueError('"error_atol" cannot be negative.')

    if error_rtol is None:
        error_rtol = 0.0
    elif not numpy.isscalar(error_rtol):
        raise TypeError('"error_rtol" should be a scalar value.')
    elif not isinstance(error_rtol, (int, float)):
        raise TypeError('"error_rtol" should be a float number.')
    elif error_rtol < 0.0:
        raise ValueError('"error_rtol" cannot be negative.')

    if confidence_level is None:
        raise TypeError('"confidence_level" cannot be None.')
    elif not numpy.isscalar(confidence_level):
        raise TypeError('"confidence_level" should be a scalar.')
    elif not isinstance(confidence_level, (int, float)):
        raise TypeError('"confidence_level" should be a float number.')
    elif confidence_level < 0.0 or confidence_level > 1.0:
        raise ValueError('"confidence_level" should be between 0 and 1.')

    if outlier_significance_level is None:
        raise TypeError('"outlier_significance_level" cannot be None.')
    elif not numpy.isscalar(outlier_significance_level):
        raise TypeError('"outlier_significance_level" should be a scalar.')
    elif not isinstance(outlier_significance_level,
                        (int, numpy.integer, float)):
        raise TypeError('"outlier_significance_level" must be a float number.')
    elif outlier_significance_level < 0.0 or outlier_significance_level > 1.0:
        raise ValueError(
                '"outlier_significance_level" must be in [0, 1] interval.')

    if outlier_significance_level > 1.0 - confidence_level:
        raise ValueError('The sum of "confidence_level" and ' +
                         '"outlier_significance_level" should be less than 1.')

    if lanczos_degree is None:
        raise TypeError('"lanczos_degree" cannot be None.')
    elif not numpy.isscalar(lanczos_degree):
        raise TypeError('"lanczos_degree" should be a scalar value.')
    elif not isinstance(lanczos_degree, (int, numpy.integer)):
        raise TypeError('"lanczos_degree" should be an integer.')
    elif lanczos_degree < 1:
        raise ValueError('"lanczos_degree" should be at least one.')

    if lanczos_tol is not None and not numpy.isscalar(lanczos_tol):
        raise TypeError('"lanczos_tol" should be a scalar value.')
    elif lanczos_tol is not None and not isinstance(
            lanczos_tol, (int, numpy.integer, float)):
        raise TypeError('"lanczos_tol" should be a float number.')
    elif lanczos_tol is not None and lanczos_tol < 0.0:
        raise ValueError('"lancozs_tol" cannot be negative.')

    if orthogonalize is None:
        raise TypeError('"orthogonalize" cannot be None.')
    elif not numpy.isscalar(orthogonalize):
        raise TypeError('"orthogonalize" should be a scalar value.')
    elif not isinstance(orthogonalize, (int, numpy.integer)):
        raise TypeError('"orthogonalize" should be an integer.')
    elif orthogonalize > lanczos_degree:
        raise ValueError('"orthogonalize", if positive, should be at most ' +
                         'equal to "lanczos_degree".')

    if (seed is not None) and (not numpy.isscalar(seed)):
        raise TypeError('"seed" should be a None or a scalar value.')
    elif (seed is not None) and not isinstance(seed, (int, numpy.integer)):
        raise TypeError('"seed" should be None or an integer.')

    if num_threads is None:
        raise TypeError('"num_threads" cannot be None.')
    elif not numpy.isscalar(num_threads):
        raise TypeError('"num_threads" should be a scalar value.')
    elif not isinstance(num_threads, (int, numpy.integer)):
        raise TypeError('"num_threads" should be an integer.')
    elif num_threads < 0:
        raise ValueError('"num_threads" should be a non-negative integer.')

    if num_gpu_devices is None:
        raise TypeError('"num_gpu_devices" cannot be None.')
    elif not numpy.isscalar(num_gpu_devices):
        raise TypeError('"num_gpu_devices" should be a scalar value.')
    elif not isinstance(num_gpu_devices, (int, numpy.integer)):
        raise TypeError('"num_gpu_devices" should be an integer.')
    elif num_gpu_devices < 0:
        raise ValueError('"num_gpu_devices" should be a non-negative integer.')

    if verbose is None:
        raise TypeError('"verbose" cannot be None.')
    elif not numpy.isscalar(verbose):
        raise TypeError('"verbose" should be a scalar value.')
    elif not isinstance(verbose, bool):
        raise TypeError('"verbose" should be boolean.')

    if plot is None:
        raise TypeError('"plot" cannot be None.')
    elif not numpy.isscalar(plot):
        raise TypeError('"plot" should be a scalar value.')
    elif not isinstance(plot, bool):
        raise TypeError('"plot" should be boolean.')

    if gpu is None:
        raise TypeError('"gpu" cannot be None.')
    elif not numpy.isscalar(gpu):
        raise TypeError('"gpu" should be a scalar value.')
    elif not isinstance(gpu, bool):
        raise TypeError('"gpu" should be boolean.')

    return error_atol, error_rtol


def get_machine_precision(data_type_name):
    """trace_estimator_utilities.py:312-331."""
    return float(numpy.finfo(numpy.dtype(data_type_name.decode()
                                         if isinstance(data_type_name, bytes)
                                         else data_type_name)).eps)


def find_num_inquiries(Aop, parameters_size):
    """trace_estimator_utilities.py:338-373."""
    num_operator_parameters = Aop.get_num_parameters()
    if num_operator_parameters == 0:
        num_inquiries = 1
    elif parameters_size < num_operator_parameters:
        message = """Not enough parameters are provided. The given parameter
                     size is %d, which should be larger than %d, the number of
                     parameters of the linear operator.""" % (
                        parameters_size, num_operator_parameters)
        raise ValueError(re.sub(' +', ' ', message.replace('\n', '')))
    elif parameters_size % num_operator_parameters != 0:
        message = """Parameters size (given %d) is not an integer-multiple of
                     %d, the number of the parameters of the linear operator.
                     """ % (parameters_size, num_operator_parameters)
        raise ValueError(re.sub(' +', ' ', message.replace('\n', '')))
    else:
        num_inquiries = int(parameters_size / num_operator_parameters)
    return num_inquiries


def trace_estimator(A, parameters, func, gram, exponent, min_num_samples,
                    max_num_samples, error_atol, error_rtol, confidence_level,
                    outlier_significance_level, lanczos_degree, lanczos_tol,
                    orthogonalize, seed, num_threads, num_gpu_devices,
                    verbose, plot, gpu, precision=None, block_size=0,
                    krylov=None):
    """SLQ estimate of tr f((A + tB)^p) (trace_estimator.pyx:38-390).

    ``func`` is 'log', 'inv', 'identity' or a (name, coefficients) pair
    (see ``function_spec``). ``precision`` ('native' or
    'mixed'), ``block_size`` (probes per device block, <= 64) and ``krylov``
    (None or 'single': one Lanczos recurrence per probe, imate's method;
    'block': block Lanczos over each 64-probe block with a shared, fully
    re-orthogonalised basis, ``slq_estimate_block_krylov``) are options of
    this package; their defaults reproduce the reference behaviour.
    """
    if num_threads < 1:
        import os
        num_threads = os.cpu_count() or 1

    Aop = get_operator(A, gram)
    data_type_name = Aop.get_data_type_name()
    parameters, parameters_size = get_operator_parameters(
        parameters, data_type_name.decode())
    num_inquiries = find_num_inquiries(Aop, parameters_size)

    error_atol, error_rtol = check_arguments(
        gram, exponent, min_num_samples, max_num_samples, error_atol,
        error_rtol, confidence_level, outlier_significance_level,
        lanczos_degree, lanczos_tol, orthogonalize, seed, num_threads,
        num_gpu_devices, verbose, plot, gpu)

    if precision not in PRECISIONS:
        raise ValueError('"precision" should be None, "native" or "mixed".')
    if not isinstance(block_size, (int, numpy.integer)) or \
            block_size < 0 or block_size > 64:
        raise ValueError('"block_size" should be an integer in [0, 64].')

    if seed is None:
        seed = -1

    if lanczos_tol is None:
        lanczos_tol = get_machine_precision(data_type_name)

    if plot:
        raise NotImplementedError('plot=True (matplotlib convergence plot) '
                                  'is not part of imate_amd.')
    if krylov not in (None, 'single', 'block'):
        raise ValueError('"krylov" should be None, "single" or "block".')
    block_krylov = krylov == 'block'
    if block_krylov and (not gpu or gram):
        raise ValueError('krylov="block" needs gpu=True and gram=False.')
    if not gpu and PRECISIONS[precision] != _lib.PREC_NATIVE:
        raise ValueError('"precision" other than "native" needs gpu=True.')
    if gpu and data_type_name == b'float128':
        raise TypeError('When gpu is enabled, data type should be ' +
                        '"float32" or "float64".')

    out_dtype = numpy.dtype(data_type_name.decode())
    q = num_inquiries
    trace = numpy.empty((q,), dtype=numpy.float64)
    error = numpy.empty((q,), dtype=numpy.float64)
    samples = numpy.empty((max_num_samples, q), dtype=numpy.float64)
    samples[:, :] = numpy.nan
    processed_samples_indices = numpy.zeros((max_num_samples,),
                                            dtype=numpy.int32)
    num_samples_used = numpy.zeros((q,), dtype=numpy.int32)
    num_outliers = numpy.zeros((q,), dtype=numpy.int32)
    converged = numpy.zeros((q,), dtype=numpy.int32)
    all_converged = numpy.zeros((1,), dtype=numpy.int32)
    alg_wall_times = numpy.zeros((1,), dtype=numpy.float64)

    init_wall_time = time.perf_counter()
    init_proc_time = time.process_time()

    lib = _lib.load()
    if Aop.get_num_parameters() > 0:
        params = numpy.ascontiguousarray(
            numpy.atleast_1d(numpy.asarray(parameters, dtype=numpy.float64)))
    else:
        params = None
    fcode, fparams = function_spec(func)

    if gpu:
        gpu_Aop = Aop.get_linear_operator(gpu=True,
                                          num_gpu_devices=num_gpu_devices)
        props = gpu_Aop.get_device_properties()
        num_gpu_multiprocessors = props['num_multiprocessors'][0]
        num_gpu_threads_per_multiprocessor = \
            props['num_threads_per_multiprocessor'][0]
        # one device: the reproducible configuration of the reference
        num_gpu_devices = 1
    if gpu and block_krylov:
        _lib.check(lib.slq_estimate_block_krylov(
            gpu_Aop.handle, _lib.ptr(params), q, fcode, _lib.ptr(fparams),
            float(exponent), int(seed), int(lanczos_degree),
            int(min_num_samples), int(max_num_samples), float(error_atol),
            float(error_rtol), float(confidence_level),
            float(outlier_significance_level), PRECISIONS[precision],
            _lib.ptr(trace), _lib.ptr(error), _lib.ptr(samples),
            _lib.ptr(processed_samples_indices), _lib.ptr(num_samples_used),
            _lib.ptr(num_outliers), _lib.ptr(converged),
            _lib.ptr(all_converged), _lib.ptr(alg_wall_times)),
            'slq_estimate_block_krylov')
    elif gpu:
        _lib.check(lib.slq_estimate(
            gpu_Aop.handle, _lib.ptr(params), q, fcode, _lib.ptr(fparams),
            int(gram),
            float(exponent), int(orthogonalize), int(seed),
            int(lanczos_degree),
            float(lanczos_tol), int(min_num_samples), int(max_num_samples),
            float(error_atol), float(error_rtol), float(confidence_level),
            float(outlier_significance_level), int(block_size),
            PRECISIONS[precision], _lib.ptr(trace), _lib.ptr(error),
            _lib.ptr(samples), _lib.ptr(processed_samples_indices),
            _lib.ptr(num_samples_used), _lib.ptr(num_outliers),
            _lib.ptr(converged), _lib.ptr(all_converged),
            _lib.ptr(alg_wall_times)), 'slq_estimate')
    else:
        # CPU path (trace_estimator.pyx:246-277): the library's host
        # implementation over num_threads threads; no HIP call
        cpu_Aop = Aop.get_linear_operator(gpu=False)
        num_gpu_multiprocessors = 0
        num_gpu_threads_per_multiprocessor = 0
        _lib.check(lib.slq_estimate_host(
            cpu_Aop.handle, _lib.ptr(params), q, fcode, _lib.ptr(fparams),
            int(gram), float(exponent), int(orthogonalize), int(seed),
            int(lanczos_degree), float(lanczos_tol), int(min_num_samples),
            int(max_num_samples), float(error_atol), float(error_rtol),
            float(confidence_level), float(outlier_significance_level),
            int(num_threads), _lib.ptr(trace), _lib.ptr(error),
            _lib.ptr(samples), _lib.ptr(processed_samples_indices),
            _lib.ptr(num_samples_used), _lib.ptr(num_outliers),
            _lib.ptr(converged), _lib.ptr(all_converged),
            _lib.ptr(alg_wall_times)), 'slq_estimate_host')

    tot_wall_time = time.perf_counter() - init_wall_time
    cpu_proc_time = time.process_time() - init_proc_time

    trace = trace.astype(out_dtype)
    error = error.astype(out_dtype)
    samples = samples.astype(out_dtype)

    info = {
        'matrix':
        {
            'data_type': data_type_name,
            'gram': gram,
            'exponent': exponent,
            'size': (Aop.get_num_rows(), Aop.get_num_columns()),
            'sparse': Aop.is_sparse(),
            'nnz': Aop.get_nnz(),
            'density': Aop.get_density(),
            'num_inquiries': num_inquiries,
            'num_operator_parameters': Aop.get_num_parameters(),
            'parameters': parameters
        },
        'error':
        {
            'absolute_error': None,
            'relative_error': None,
            'error_atol': error_atol,
            'error_rtol': error_rtol,
            'confidence_level': confidence_level,
            'outlier_significance_level': outlier_significance_level
        },
        'convergence':
        {
            'converged': None,
            'all_converged': bool(all_converged[0]),
            'min_num_samples': min_num_samples,
            'max_num_samples': max_num_samples,
            'num_samples_used': None,
            'num_outliers': None,
            'samples': None,
            'samples_mean': None,
            'samples_processed_order': processed_samples_indices
        },
        'time':
        {
            'tot_wall_time': tot_wall_time,
            'alg_wall_time': float(alg_wall_times[0]),
            'cpu_proc_time': cpu_proc_time
        },
        'device':
        {
            'num_cpu_threads': num_threads,
            'num_gpu_devices': num_gpu_devices,
            'num_gpu_multiprocessors': num_gpu_multiprocessors,
            'num_gpu_threads_per_multiprocessor':
                num_gpu_threads_per_multiprocessor
        },
        'solver':
        {
            'version': __version__,
            'lanczos_degree': lanczos_degree,
            'lanczos_tol': lanczos_tol,
            'orthogonalize': orthogonalize,
            'seed': seed,
            'method': 'slq',
        }
    }

    output_is_array = not ((parameters is None) or numpy.isscalar(parameters))
    if output_is_array:
        info['error']['absolute_error'] = error
        info['error']['relative_error'] = \
            error / (numpy.abs(trace) + numpy.finfo(float).eps)
        info['convergence']['converged'] = converged.astype(bool)
        info['convergence']['num_samples_used'] = num_samples_used
        info['convergence']['num_outliers'] = num_outliers
        info['convergence']['samples'] = samples
        info['convergence']['samples_mean'] = trace
    else:
        info['error']['absolute_error'] = error[0]
        info['error']['relative_error'] = \
            error[0] / (numpy.abs(trace[0]) + numpy.finfo(float).eps)
        info['convergence']['converged'] = bool(converged[0])
        info['convergence']['num_samples_used'] = num_samples_used[0]
        info['convergence']['num_outliers'] = num_outliers[0]
        info['convergence']['samples'] = samples[:, 0]
        info['convergence']['samples_mean'] = trace[0]

    if verbose:
        from ._summary import print_summary
        print_summary(info)

    if output_is_array:
        return trace, info
    return trace[0], info
