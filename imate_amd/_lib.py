"""ctypes binding of ``libslq_hip.so`` (the C ABI declared in
``include/slq_hip.h``).

The library is loaded from this package directory only (it is built in place
by ``make -C imate_amd/csrc``). There is no fallback: when the library or a
HIP device is missing, the GPU entry points raise, as imate raises when it was
built without CUDA (imate/_trace_estimator/trace_estimator.pyx:190-197).
"""

import ctypes
import os

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libslq_hip.so')

SLQ_OK = 0
SLQ_E_ARG = -1
SLQ_E_NODEVICE = -2
SLQ_E_OOM = -3
SLQ_E_HIP = -4
SLQ_E_NOTSPD = -5

SLQ_F64 = 0
SLQ_F32 = 1
SLQ_F128 = 2

FUNC_LOG = 0
FUNC_INV = 1
FUNC_IDENT = 2
FUNC_EXP = 3
FUNC_LINFRAC = 4
FUNC_INDICATOR = 5
FUNC_GAUSSIAN = 6

PREC_NATIVE = 0
PREC_MIXED = 1

# Every symbol of include/slq_hip.h, with its ctypes signature.
_P = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
SIGNATURES = {
    'slq_last_error': (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    'slq_device_properties': (ctypes.c_int, [_P, _P, _P, ctypes.c_char_p,
                                             ctypes.c_size_t]),
    'slq_op_create_csr': (ctypes.c_int, [_i64, _i64, _i64, _P, _P, _P,
                                         ctypes.c_int, _P]),
    'slq_op_create_csc': (ctypes.c_int, [_i64, _i64, _i64, _P, _P, _P,
                                         ctypes.c_int, _P]),
    'slq_op_create_affine': (ctypes.c_int, [_P, _P, _P]),
    'slq_op_create_dense': (ctypes.c_int, [_i64, _i64, _P, ctypes.c_int, _P]),
    'slq_op_destroy': (ctypes.c_int, [_P]),
    'slq_op_set_values': (ctypes.c_int, [_P, _P]),
    'slq_op_family_attach': (ctypes.c_int, [_P, _P, _P, _P]),
    'slq_op_family_values': (ctypes.c_int, [_P, _f64]),
    'slq_op_create_sparse_gaussian': (ctypes.c_int, [_i64, _i64, _i32, _i64,
                                                     ctypes.c_int, _P]),
    'slq_op_get_csr': (ctypes.c_int, [_P, _P, _P, _P]),
    'slq_op_shape': (ctypes.c_int, [_P, _P, _P, _P, _P]),
    'slq_estimate': (ctypes.c_int, [
        _P, _P, _i32, _i32, _P, _i32, _f64, _i32, _i64, _i32, _f64, _i32, _i32,
        _f64, _f64, _f64, _f64, _i32, _i32, _P, _P, _P, _P, _P, _P, _P, _P,
        _P]),
    'slq_estimate_noalloc': (ctypes.c_int, [
        _P, _P, _i32, _i32, _P, _i32, _f64, _i32, _i64, _i32, _f64, _i32, _i32,
        _f64, _f64, _f64, _f64, _i32, _i32, _P, _P, _P, _P, _P, _P, _P, _P,
        _P]),
    'slq_estimate_block_krylov': (ctypes.c_int, [
        _P, _P, _i32, _i32, _P, _f64, _i64, _i32, _i32, _i32, _f64, _f64, _f64,
        _f64, _i32, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    'slq_rademacher': (ctypes.c_int, [_i64, _i64, _i64, _i32, ctypes.c_int,
                                      _P, _P]),
    'slq_op_spmm': (ctypes.c_int, [_P, _f64, _i32, _i32, _P, _P]),
    'slq_recurrences': (ctypes.c_int, [_P, _f64, _i32, _i64, _i64, _i32, _i32,
                                       _f64, _i32, _i32, _P, _P, _P]),
    'slq_quadrature': (ctypes.c_int, [_P, _P, _P, _i32, _i32, _i32, _i32,
                                      _P, _f64, _f64, _P, _i32, _i32, _P, _P,
                                      _P]),
    'slq_profile_enable': (ctypes.c_int, [_P, _i32]),
    'slq_profile_read': (ctypes.c_int, [_P, _P, _P, _P, _P]),
    'slq_profile_steps': (ctypes.c_int, [_P, _P]),
    'slq_profile_gemm': (ctypes.c_int, [_P, _P, _P, _P]),
    'slq_copy_bandwidth': (ctypes.c_int, [_i64, _i32, _P]),
    'slq_estimate_batch': (ctypes.c_int, [
        _P, _i32, _i32, _P, _i32, _f64, _i32, _i64, _i32, _f64, _i32, _f64, _f64,
        _i32, _i32, _P, _P, _P, _P, _P]),
    'slq_autotune': (ctypes.c_int, [_P, _i32, _P, _P]),
    'slq_op_set_launch': (ctypes.c_int, [_P, _i32, _i32, _i32, _i32]),
    'slq_op_get_launch': (ctypes.c_int, [_P, _P, _P, _P, _P]),
    'slq_mlp_num_params': (ctypes.c_int, []),
    'slq_mlp_grad': (ctypes.c_int, [_P, _P, _P, _i32, _P, _P]),
    'slq_mlp_train': (ctypes.c_int, [_P, _P, _P, _i32, _i32, ctypes.c_float,
                                     _P]),
    'slq_mlp_predict': (ctypes.c_int, [_P, _P, _i32, _i32, _P, _P]),
    'slq_host_op_create_csr': (ctypes.c_int, [_i64, _i64, _i64, _P, _P, _P,
                                              _i32, _P]),
    'slq_host_op_create_csc': (ctypes.c_int, [_i64, _i64, _i64, _P, _P, _P,
                                              _i32, _P]),
    'slq_host_op_create_dense': (ctypes.c_int, [_i64, _i64, _P, _i32, _P]),
    'slq_host_op_create_affine': (ctypes.c_int, [_P, _P, _P]),
    'slq_host_op_destroy': (ctypes.c_int, [_P]),
    'slq_estimate_host': (ctypes.c_int, [
        _P, _P, _i32, _i32, _P, _i32, _f64, _i32, _i64, _i32, _f64, _i32, _i32,
        _f64, _f64, _f64, _f64, _i32, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
}

_lib = None


def load():
    """Load the library (once). Raises ImportError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            'libslq_hip.so is not built. Build it with '
            '"make -C imate_amd/csrc" (hipcc, gfx950) before using gpu=True.')
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().slq_last_error(buf, len(buf))
    return buf.value.decode(errors='replace')


def check(code, what):
    """Map a status code to the exception type imate raises for it."""
    if code == SLQ_OK:
        return
    msg = '%s: %s' % (what, last_error())
    if code == SLQ_E_ARG:
        raise ValueError(msg)
    if code == SLQ_E_OOM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def ptr(a):
    """Address of a numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


def device_properties():
    """``{'num_devices', 'num_multiprocessors', 'num_threads_per_multiprocessor'}``
    like pycuLinearOperator.get_device_properties (py_cu_linear_operator.pyx:195)."""
    lib = load()
    nd = ctypes.c_int(0)
    nm = ctypes.c_int(0)
    nt = ctypes.c_int(0)
    name = ctypes.create_string_buffer(128)
    check(lib.slq_device_properties(ctypes.byref(nd), ctypes.byref(nm),
                                    ctypes.byref(nt), name, len(name)),
          'slq_device_properties')
    return {'num_devices': nd.value,
            'num_multiprocessors': [nm.value] * nd.value,
            'num_threads_per_multiprocessor': [nt.value] * nd.value,
            'name': name.value.decode()}


def dtype_code(dtype):
    dtype = numpy.dtype(dtype)
    if dtype == numpy.float64:
        return SLQ_F64
    if dtype == numpy.float32:
        return SLQ_F32
    raise TypeError('When gpu is enabled, data type should be "float32" or '
                    '"float64".')


def host_dtype_code(dtype):
    """Value type code of a host (gpu=False) handle: float32, float64 or
    float128 (long double), the types imate's CPU path accepts."""
    dtype = numpy.dtype(dtype)
    if dtype == numpy.float64:
        return SLQ_F64
    if dtype == numpy.float32:
        return SLQ_F32
    if dtype == numpy.longdouble:
        return SLQ_F128
    raise TypeError('Data type should be "float32", "float64", or '
                    '"float128".')
