"""ORACLE — test infrastructure only.

CPU restatement of imate's stochastic Lanczos quadrature (SLQ), used by the
tests, by ``__graft_entry__.smoke()`` and by ``bench.py``'s ``cpu_baseline``
leg, and only to check results (or as the timed CPU baseline). The product package
``imate_amd`` never imports this module.

Split of the work:

* ``libslq_oracle.so`` (``slq_oracle.c``): RNG, operators, Gram-Schmidt,
  Lanczos and Golub-Kahn recurrences, restated line by line from the
  reference's C++ (file:line in that source).
* this module: the per-sample driver of
  ``_c_trace_estimator/c_trace_estimator.cpp:410-654`` (eigensolve with LAPACK
  ``?stev`` / ``?bdsdc`` through SciPy's ``cython_lapack``, which is the very
  routine the reference binds in ``_c_trace_estimator/lapack_api.pxi:14-99``;
  Ritz shift for ``A + tI``; quadrature ``n * sum tau^2 f(theta^p)``), the
  Monte-Carlo loop of ``c_trace_estimator.cpp:191-307`` with one thread, and
  ``ConvergenceTools`` (``convergence_tools.cpp:94-376``).

Pinning: see the header of ``slq_oracle.c``. The LAPACK routines are the
reference's own third-party dependency (SciPy's bundled LAPACK, version not
pinned by the reference); their use here is identical to the reference's.
"""

import ctypes
import math
import os

import numpy
import scipy.sparse

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None


# ---------------------------------------------------------------- loading

def lib():
    """The restatement library (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'libslq_oracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle library missing: run make -C oracle')
        _LIB = ctypes.CDLL(path)
        _declare(_LIB)
    return _LIB


def ref_lib():
    """The reference's own CPU sources built into ``oracle/_ref`` or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, '_ref', 'libimate_ref.so')
        if not os.path.exists(path):
            return None
        _REF = ctypes.CDLL(path)
    return _REF


class _CSR(ctypes.Structure):
    _fields_ = [('nrows', ctypes.c_int64), ('ncols', ctypes.c_int64),
                ('indptr', ctypes.c_void_p), ('indices', ctypes.c_void_p),
                ('data', ctypes.c_void_p)]


class _Op(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('t', ctypes.c_double),
                ('A', _CSR), ('B', _CSR)]


OP_MATRIX, OP_AFFINE_I, OP_AFFINE_B = 0, 1, 2

_P = ctypes.c_void_p
_u64p = ctypes.POINTER(ctypes.c_uint64)


def _declare(L):
    L.orc_thread_state.argtypes = [ctypes.c_int64, ctypes.c_int32, _P]
    L.orc_splitmix.argtypes = [ctypes.c_uint64, ctypes.c_int32, _P]
    L.orc_xoshiro_words.argtypes = [_P, ctypes.c_int64, _P]
    L.orc_xoshiro_jump.argtypes = [_P]
    L.orc_erf_inv.argtypes = [ctypes.c_double]
    L.orc_erf_inv.restype = ctypes.c_double
    for sfx, real in (('f64', ctypes.c_double), ('f32', ctypes.c_float)):
        getattr(L, 'orc_rademacher_' + sfx).argtypes = [
            _P, ctypes.c_int64, ctypes.c_int32, _P]
        getattr(L, 'orc_op_dot_' + sfx).argtypes = [_P, _P, _P]
        getattr(L, 'orc_op_tdot_' + sfx).argtypes = [_P, _P, _P]
        getattr(L, 'orc_dot_' + sfx).argtypes = [_P, _P, ctypes.c_int64]
        getattr(L, 'orc_dot_' + sfx).restype = real
        getattr(L, 'orc_norm_' + sfx).argtypes = [_P, ctypes.c_int64]
        getattr(L, 'orc_norm_' + sfx).restype = real
        getattr(L, 'orc_gram_schmidt_' + sfx).argtypes = [
            _P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
            ctypes.c_int32, _P]
        getattr(L, 'orc_recurrence_' + sfx).argtypes = [
            _P, ctypes.c_int32, _P, ctypes.c_int32, real, ctypes.c_int32,
            _P, _P]
        getattr(L, 'orc_recurrence_' + sfx).restype = ctypes.c_int32
        getattr(L, 'orc_slq_recurrences_' + sfx).argtypes = [
            _P, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
            ctypes.c_int32, ctypes.c_int32, real, ctypes.c_int32, _P, _P, _P]
        getattr(L, 'orc_slq_recurrences_' + sfx).restype = ctypes.c_int32


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _sfx(dtype):
    dtype = numpy.dtype(dtype)
    if dtype == numpy.float64:
        return 'f64'
    if dtype == numpy.float32:
        return 'f32'
    raise TypeError('oracle supports float32 and float64, not %s' % dtype)


# ---------------------------------------------------------------- operator

class Operator(object):
    """``A``, ``A + tI`` or ``A + tB`` over CSR arrays the oracle can read.

    Keeps the converted arrays alive while the C structure points at them.
    """

    def __init__(self, A, dtype=None, B=None, affine=False, t=0.0):
        A = scipy.sparse.csr_matrix(A)
        A.sort_indices()
        self.dtype = numpy.dtype(dtype if dtype is not None else A.dtype)
        self.shape = A.shape
        self._keep = []
        self.op = _Op()
        self.op.kind = OP_MATRIX if not affine else (
            OP_AFFINE_I if B is None else OP_AFFINE_B)
        self.op.t = float(t)
        self.op.A = self._csr(A)
        if B is not None:
            B = scipy.sparse.csr_matrix(B)
            B.sort_indices()
            self.op.B = self._csr(B)

    def _csr(self, M):
        ptr = numpy.ascontiguousarray(M.indptr, dtype=numpy.int64)
        idx = numpy.ascontiguousarray(M.indices, dtype=numpy.int32)
        dat = numpy.ascontiguousarray(M.data, dtype=self.dtype)
        self._keep += [ptr, idx, dat]
        c = _CSR()
        c.nrows, c.ncols = M.shape
        c.indptr, c.indices, c.data = _ptr(ptr), _ptr(idx), _ptr(dat)
        return c

    def set_t(self, t):
        self.op.t = float(t)

    @property
    def ref(self):
        return ctypes.byref(self.op)


def op_dot(op, x):
    x = numpy.ascontiguousarray(x, dtype=op.dtype)
    y = numpy.empty(op.shape[0], dtype=op.dtype)
    getattr(lib(), 'orc_op_dot_' + _sfx(op.dtype))(op.ref, _ptr(x), _ptr(y))
    return y


def op_tdot(op, x):
    x = numpy.ascontiguousarray(x, dtype=op.dtype)
    y = numpy.empty(op.shape[1], dtype=op.dtype)
    getattr(lib(), 'orc_op_tdot_' + _sfx(op.dtype))(op.ref, _ptr(x), _ptr(y))
    return y


# ---------------------------------------------------------------- RNG

def thread_state(seed, thread_id=0):
    s = numpy.zeros(4, dtype=numpy.uint64)
    lib().orc_thread_state(int(seed), int(thread_id), _ptr(s))
    return s


def splitmix(state, count):
    out = numpy.zeros(count, dtype=numpy.uint64)
    lib().orc_splitmix(ctypes.c_uint64(int(state)), int(count), _ptr(out))
    return out


def xoshiro_words(state, count):
    """Advance ``state`` (uint64[4], in place) and return ``count`` words."""
    out = numpy.zeros(count, dtype=numpy.uint64)
    lib().orc_xoshiro_words(_ptr(state), int(count), _ptr(out))
    return out


def xoshiro_jump(state):
    lib().orc_xoshiro_jump(_ptr(state))


def rademacher(seed, n, count, dtype=numpy.float64, thread_id=0):
    """``count`` consecutive probes of length ``n`` from one thread stream."""
    s = thread_state(seed, thread_id)
    out = numpy.empty((count, n), dtype=dtype)
    getattr(lib(), 'orc_rademacher_' + _sfx(dtype))(
        _ptr(s), int(n), int(count), _ptr(out))
    return out


def erf_inv(x):
    return lib().orc_erf_inv(float(x))


# ---------------------------------------------------------------- recurrences

def recurrence(op, v, m, tol, ortho, gram=False):
    """One Lanczos (or Golub-Kahn) run from ``v``: (alpha, beta, size)."""
    dt = op.dtype
    v = numpy.ascontiguousarray(v, dtype=dt)
    a = numpy.zeros(m, dtype=dt)
    b = numpy.zeros(m, dtype=dt)
    k = getattr(lib(), 'orc_recurrence_' + _sfx(dt))(
        op.ref, int(bool(gram)), _ptr(v), int(m), tol, int(ortho),
        _ptr(a), _ptr(b))
    if k < 0:
        raise MemoryError('oracle recurrence failed to allocate')
    return a, b, k


def recurrences(op, seed, num_samples, m, tol, ortho, gram=False,
                num_threads=1):
    """Recurrences of ``num_samples`` probes (thread-interleaved streams)."""
    dt = op.dtype
    a = numpy.zeros((num_samples, m), dtype=dt)
    b = numpy.zeros((num_samples, m), dtype=dt)
    sizes = numpy.zeros(num_samples, dtype=numpy.int32)
    r = getattr(lib(), 'orc_slq_recurrences_' + _sfx(dt))(
        op.ref, int(bool(gram)), int(seed), int(num_threads),
        int(num_samples), int(m), tol, int(ortho), _ptr(a), _ptr(b),
        _ptr(sizes))
    if r != 0:
        raise MemoryError('oracle recurrences failed to allocate')
    return a, b, sizes


# ---------------------------------------------------------------- LAPACK

def _capsule(name, argtypes):
    import scipy.linalg.cython_lapack as cl
    cap = cl.__pyx_capi__[name]
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = ctypes.c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    addr = get_ptr(cap, get_name(cap))
    return ctypes.CFUNCTYPE(None, *argtypes)(addr)


_LAPACK = {}


def _lapack(name):
    if name not in _LAPACK:
        P = ctypes.c_void_p
        if name.endswith('stev'):
            args = [P] * 8
        else:
            args = [P] * 14
        _LAPACK[name] = _capsule(name, args)
    return _LAPACK[name]


def eigh_tridiagonal(alpha, beta, k):
    """``Diagonalization::eigh_tridiagonal`` (diagonalization.cpp:88-118):
    ``?stev('V')`` on the leading k x k block; returns (theta, tau) with tau
    the first component of each eigenvector (c_trace_estimator.cpp:558-562)."""
    dt = alpha.dtype
    d = numpy.array(alpha[:k], dtype=dt)
    e = numpy.array(beta[:max(k, 1)], dtype=dt)
    z = numpy.zeros(k * k, dtype=dt)
    work = numpy.zeros(max(2 * k - 2, 1), dtype=dt)
    jobz = ctypes.c_char(b'V')
    n = ctypes.c_int(k)
    info = ctypes.c_int(0)
    fn = _lapack('dstev' if dt == numpy.float64 else 'sstev')
    fn(ctypes.byref(jobz), ctypes.byref(n), _ptr(d), _ptr(e), _ptr(z),
       ctypes.byref(n), _ptr(work), ctypes.byref(info))
    if info.value != 0:
        raise RuntimeError('?stev info=%d' % info.value)
    tau = z[numpy.arange(k) * k]
    return d, tau


def svd_bidiagonal(alpha, beta, k):
    """``Diagonalization::svd_bidiagonal`` (diagonalization.cpp:189-231):
    ``?bdsdc('U', 'I')``; theta = sigma^2 and tau = Vt[i] (the first component
    of the i-th right singular vector), c_trace_estimator.cpp:525-534."""
    dt = alpha.dtype
    d = numpy.array(alpha[:k], dtype=dt)
    e = numpy.array(beta[:max(k, 1)], dtype=dt)
    u = numpy.zeros(k * k, dtype=dt)
    vt = numpy.zeros(k * k, dtype=dt)
    work = numpy.zeros(3 * k * k + 4 * k, dtype=dt)
    iwork = numpy.zeros(8 * k, dtype=numpy.int32)
    uplo = ctypes.c_char(b'U')
    compq = ctypes.c_char(b'I')
    n = ctypes.c_int(k)
    info = ctypes.c_int(0)
    fn = _lapack('dbdsdc' if dt == numpy.float64 else 'sbdsdc')
    fn(ctypes.byref(uplo), ctypes.byref(compq), ctypes.byref(n), _ptr(d),
       _ptr(e), _ptr(u), ctypes.byref(n), _ptr(vt), ctypes.byref(n), None,
       None, _ptr(work), _ptr(iwork), ctypes.byref(info))
    if info.value != 0:
        raise RuntimeError('?bdsdc info=%d' % info.value)
    theta = d * d
    tau = vt[:k]
    return theta, tau


# ---------------------------------------------------------------- quadrature

FUNCTIONS = ('log', 'inv', 'identity', 'exp', 'linfrac', 'indicator',
             'gaussian')

# M_SQRT1_2 and M_2_SQRTPI of <math.h>, as imate's gaussian.cpp uses them
_SQRT1_2 = 0.70710678118654752440
_2_SQRTPI = 1.12837916709551257390


def _spec(func):
    """(name, coefficients) of ``func``: a name, or (name, coefficients)."""
    if isinstance(func, str):
        return func, None
    return func[0], [float(c) for c in func[1]]


def _fn64(name, x, fp=None):
    """functions/{logarithm,inverse,identity,exponential,homographic,
    indicator,gaussian}.cpp on a double, with the C library's behaviour at
    the poles."""
    if name == 'log':
        if x > 0.0:
            return math.log(x)
        return -math.inf if x == 0.0 else math.nan
    if name == 'inv':
        if x == 0.0:
            return math.copysign(math.inf, x)
        return 1.0 / x
    if name == 'exp':
        # exponential.cpp:44-47: exp(lambda * coeff)
        try:
            return math.exp(x * fp[0])
        except OverflowError:
            return math.inf
    if name == 'linfrac':
        # homographic.cpp:53-61: (a lambda + b) / (c lambda + d)
        num = fp[0] * x + fp[1]
        den = fp[2] * x + fp[3]
        if den == 0.0:
            return math.nan if num == 0.0 else math.copysign(math.inf, num)
        return num / den
    if name == 'indicator':
        # indicator.cpp:51-61: 0 outside [a, b]
        return 0.0 if (x < fp[0] or x > fp[1]) else 1.0
    if name == 'gaussian':
        # gaussian.cpp:59-63
        z = (x - fp[0]) / fp[1]
        return (0.5 * _SQRT1_2 * _2_SQRTPI / fp[1]) * math.exp(-0.5 * z * z)
    return x


def _fn32(name, x, fp=None):
    """The same in single precision (the float overloads)."""
    f32 = numpy.float32
    if name == 'log':
        return numpy.log(x)
    if name == 'inv':
        return f32(1.0) / x
    if name == 'exp':
        return numpy.exp(x * f32(fp[0]))
    if name == 'linfrac':
        a, b, c, d = (f32(v) for v in fp)
        return (a * x + b) / (c * x + d)
    if name == 'indicator':
        return f32(0.0) if (x < fp[0] or x > fp[1]) else f32(1.0)
    if name == 'gaussian':
        mu, sigma = f32(fp[0]), f32(fp[1])
        z = (x - mu) / sigma
        return f32((0.5 * _SQRT1_2 * _2_SQRTPI / sigma) *
                   numpy.exp(f32(-0.5) * z * z))
    return x


def _pow64(x, p):
    try:
        return math.pow(x, p)
    except (ValueError, OverflowError, ZeroDivisionError):
        if x == 0.0 and p < 0:
            return math.inf
        return math.nan


def quadrature(theta, tau, k, func, exponent, n, dtype=numpy.float64):
    """``trace_estimate = n * sum_{i<k} tau_i^2 f(theta_i^p)``,
    c_trace_estimator.cpp:597-621 (sequential sum in the matrix dtype).
    ``func`` is a name or (name, coefficients)."""
    name, fp = _spec(func)
    if numpy.dtype(dtype) == numpy.float64:
        s = 0.0
        for i in range(k):
            s += float(tau[i]) * float(tau[i]) * _fn64(
                name, _pow64(float(theta[i]), float(exponent)), fp)
        return n * s
    f32 = numpy.float32
    with numpy.errstate(all='ignore'):
        s = f32(0.0)
        for i in range(k):
            x = numpy.power(f32(theta[i]), f32(exponent))
            fx = _fn32(name, x, fp)
            s = f32(s + f32(f32(tau[i]) * f32(tau[i])) * fx)
        return f32(n) * s


# ---------------------------------------------------------------- convergence

def check_convergence(samples, min_num_samples, processed, num_processed,
                      confidence_level, error_atol, error_rtol, error,
                      num_samples_used, converged):
    """``ConvergenceTools::check_convergence`` (convergence_tools.cpp:94-194).
    Updates error / num_samples_used / converged in place."""
    q = samples.shape[1]
    if num_processed < min_num_samples:
        error[:] = math.inf
        converged[:] = 0
        num_samples_used[:] = num_processed
        return 0
    z = math.sqrt(2) * erf_inv(float(confidence_level))
    for j in range(q):
        if converged[j] == 0:
            s = 0.0
            for i in range(num_processed):
                s += float(samples[processed[i], j])
            mean = s / num_processed
            if num_processed > 1:
                s = 0.0
                for i in range(num_processed):
                    d = float(samples[processed[i], j])
                    s += (d - mean) * (d - mean)
                std = math.sqrt(s / (num_processed - 1.0))
            else:
                std = math.inf
            error[j] = z * std / math.sqrt(num_processed)
            if error[j] < max(error_atol, error_rtol * mean):
                converged[j] = 1
            num_samples_used[j] = num_processed
    return int(all(converged[j] != 0 for j in range(q)))


def average_estimates(samples, confidence_level, outlier_significance_level,
                      num_samples_used, processed, num_outliers, trace, error):
    """``ConvergenceTools::average_estimates`` (convergence_tools.cpp:255-376)."""
    q = samples.shape[1]
    ez = math.sqrt(2) * erf_inv(float(confidence_level))
    oz = math.sqrt(2.0) * erf_inv(1.0 - float(outlier_significance_level))
    for j in range(q):
        N = int(num_samples_used[j])
        out = [0] * N
        num_outliers[j] = 0
        s = 0.0
        for i in range(N):
            s += float(samples[processed[i], j])
        mean = s / N if N > 0 else math.nan
        if N > 1:
            s = 0.0
            for i in range(N):
                d = float(samples[processed[i], j]) - mean
                s += d * d
            std = math.sqrt(s / (N - 1.0))
        else:
            std = math.inf
        half = oz * std
        for i in range(N):
            d = float(samples[processed[i], j]) - mean
            if abs(d) > half:
                out[i] = 1
                num_outliers[j] += 1
        no = int(num_outliers[j])
        s = 0.0
        for i in range(N):
            if out[i] == 0:
                s += float(samples[processed[i], j])
        mean = s / (N - no) if N - no > 0 else math.nan
        if N > 1 + no:
            s = 0.0
            for i in range(N):
                if out[i] == 0:
                    d = float(samples[processed[i], j]) - mean
                    s += d * d
            std = math.sqrt(s / (N - no - 1.0))
        else:
            std = math.inf
        trace[j] = mean
        error[j] = ez * std / math.sqrt(N - no) if N - no > 0 else math.nan


# ---------------------------------------------------------------- SLQ driver

def sample_estimates(op, probe, func, gram, exponent, m, tol, ortho,
                     parameters=None, converged=None):
    """One Monte-Carlo sample (``_c_stochastic_lanczos_quadrature``,
    c_trace_estimator.cpp:410-654) for every inquiry. Returns the estimates
    (NaN for skipped inquiries) and the recurrence sizes."""
    dt = op.dtype
    kind = op.op.kind
    params = [None] if parameters is None else list(parameters)
    q = len(params)
    if converged is None:
        converged = numpy.zeros(q, dtype=numpy.int32)
    relation = (kind == OP_AFFINE_I)
    required = 1 if relation else q
    n = op.shape[1] if gram else op.shape[0]
    theta = [None] * q
    tau = [None] * q
    sizes = [0] * q
    for j in range(required):
        if converged[j] == 1 and required == q:
            continue
        if params[j] is not None:
            op.set_t(params[j])
        a, b, k = recurrence(op, probe, m, tol, ortho, gram=gram)
        if gram:
            theta[j], tau[j] = svd_bidiagonal(a, b, k)
        else:
            theta[j], tau[j] = eigh_tridiagonal(a, b, k)
        sizes[j] = k
    if relation and q > 1:
        converged[:] = 0
        for j in range(1, q):
            sizes[j] = sizes[0]
            th = numpy.empty_like(theta[0])
            for i in range(sizes[0]):
                # cAffineMatrixFunction::get_eigenvalue: theta - t0 + tj
                th[i] = theta[0][i] - dt.type(params[0]) + dt.type(params[j])
            theta[j] = th
            tau[j] = tau[0]
    est = numpy.full(q, numpy.nan, dtype=dt)
    for j in range(q):
        if converged[j] == 1:
            continue
        est[j] = quadrature(theta[j], tau[j], sizes[j], func, exponent, n, dt)
    return est, sizes


def slq(A, func='log', gram=False, exponent=1.0, parameters=None, B=None,
        affine=False, min_num_samples=10, max_num_samples=50,
        error_atol=0.0, error_rtol=1e-2, confidence_level=0.95,
        outlier_significance_level=0.001, lanczos_degree=20,
        lanczos_tol=None, orthogonalize=0, seed=0, dtype=None):
    """Full SLQ with one thread, following c_trace_estimator.cpp:191-307.

    ``affine=True`` makes ``A + tI`` (or ``A + tB`` when B is given) with
    ``parameters`` the t values. Returns a dict with the outputs of
    ``pyc_trace_estimator`` (trace, error, samples, processed order, ...).
    """
    if affine and parameters is None:
        raise ValueError('affine operators need parameters')
    op = Operator(A, dtype=dtype, B=B, affine=affine)
    dt = op.dtype
    if parameters is not None and numpy.isscalar(parameters):
        parameters = [parameters]
    q = 1 if parameters is None else len(parameters)
    m = int(lanczos_degree)
    tol = float(numpy.finfo(dt).eps) if lanczos_tol is None else lanczos_tol
    n = op.shape[1] if gram else op.shape[0]

    samples = numpy.full((max_num_samples, q), numpy.nan, dtype=dt)
    processed = numpy.zeros(max_num_samples, dtype=numpy.int32)
    num_used = numpy.zeros(q, dtype=numpy.int32)
    num_outliers = numpy.zeros(q, dtype=numpy.int32)
    converged = numpy.zeros(q, dtype=numpy.int32)
    trace = numpy.zeros(q, dtype=dt)
    error = numpy.zeros(q, dtype=dt)
    sizes = numpy.zeros((max_num_samples, q), dtype=numpy.int32)

    state = thread_state(seed, 0)
    probe = numpy.empty(n, dtype=dt)
    all_conv = 0
    nproc = 0
    sfx = _sfx(dt)
    for i in range(max_num_samples):
        if all_conv:
            break
        getattr(lib(), 'orc_rademacher_' + sfx)(_ptr(state), int(n), 1,
                                                _ptr(probe))
        est, sz = sample_estimates(op, probe, func, gram, exponent, m, tol,
                                   orthogonalize, parameters, converged)
        samples[i, :] = est
        sizes[i, :] = sz
        processed[nproc] = i
        nproc += 1
        all_conv = check_convergence(
            samples, min_num_samples, processed, nproc, confidence_level,
            error_atol, error_rtol, error, num_used, converged)
    average_estimates(samples, confidence_level, outlier_significance_level,
                      num_used, processed, num_outliers, trace, error)
    return dict(trace=trace, error=error, samples=samples,
                processed_samples_indices=processed,
                num_samples_used=num_used, num_outliers=num_outliers,
                converged=converged, all_converged=all_conv, sizes=sizes)


def first_samples(A, func='log', lanczos_degree=20, count=16, seed=0,
                  gram=False, exponent=1.0, orthogonalize=0, dtype=None,
                  threads=16):
    """Per-sample estimates of the first ``count`` samples of the one-thread
    (reproducible) stream, i.e. ``slq(...)['samples'][:count, 0]`` of a plain
    matrix, with the samples' recurrences spread over ``threads`` host
    threads (each runs ``sample_estimates`` on its probe; the C calls release
    the interpreter lock). For sizes where the one-thread driver would take
    minutes."""
    from concurrent.futures import ThreadPoolExecutor
    op = Operator(A, dtype=dtype)
    dt = op.dtype
    n = op.shape[1] if gram else op.shape[0]
    tol = float(numpy.finfo(dt).eps)
    probes = rademacher(seed, n, count, dtype=dt)

    def one(k):
        est, sizes = sample_estimates(op, probes[k], func, gram, exponent,
                                      int(lanczos_degree), tol,
                                      int(orthogonalize))
        return est[0], sizes[0]

    with ThreadPoolExecutor(max_workers=max(1, int(threads))) as ex:
        res = list(ex.map(one, range(count)))
    return (numpy.array([r[0] for r in res]),
            numpy.array([r[1] for r in res], dtype=numpy.int32))
