"""Linear operators of the drop-in: ``LinearOperator``, ``Matrix`` and
``AffineMatrixFunction``, with the interface of
imate/linear_operator/{linear_operator,matrix,affine_matrix_function}.py.

The operator is uploaded lazily, once, on first GPU use (like
``Matrix.initialize``, matrix.py:286-352) and reused by every later call, so
passing a ``Matrix`` instead of a scipy matrix keeps A resident in HBM across
calls (the usage imate recommends, logdet/_slq_method.pyx:218-222).

``get_linear_operator(gpu=False)`` makes the host copy used by the CPU
path (``slq_estimate_host``), like imate's pycMatrix / pycAffineMatrixFunction
(py_c_matrix.pyx:119-215, py_c_affine_matrix_function.pyx:172-305).
"""

import ctypes

import numpy
import scipy.sparse

from . import _lib

__all__ = ['LinearOperator', 'Matrix', 'AffineMatrixFunction', 'DeviceMatrix']


class HostOperator(object):
    """Owner of one ``slq_host_op*`` handle (host copy of the operator for
    the CPU path)."""

    def __init__(self, handle, parents=()):
        self.handle = handle
        self._parents = parents

    @staticmethod
    def from_matrix(A):
        """Host copy of A: dense for a numpy array (imate's cDenseMatrix),
        CSC for a scipy CSC matrix (cCSCMatrix), CSR otherwise
        (cCSRMatrix)."""
        lib = _lib.load()
        if A.ndim != 2:
            raise ValueError('Input matrix should be a 2-dimensional array.')
        code = _lib.host_dtype_code(A.dtype)
        h = ctypes.c_void_p()
        if not scipy.sparse.issparse(A):
            D = numpy.ascontiguousarray(A)
            _lib.check(lib.slq_host_op_create_dense(
                D.shape[0], D.shape[1], _lib.ptr(D), code, ctypes.byref(h)),
                'slq_host_op_create_dense')
            return HostOperator(h)
        if A.format == 'csc':
            fn, name, M = lib.slq_host_op_create_csc, 'slq_host_op_create_csc', A
        else:
            fn, name, M = (lib.slq_host_op_create_csr, 'slq_host_op_create_csr',
                           scipy.sparse.csr_matrix(A))
        indptr = numpy.ascontiguousarray(M.indptr, dtype=numpy.int64)
        indices = numpy.ascontiguousarray(M.indices, dtype=numpy.int32)
        data = numpy.ascontiguousarray(M.data, dtype=A.dtype)
        _lib.check(fn(M.shape[0], M.shape[1], M.nnz, _lib.ptr(indptr),
                      _lib.ptr(indices), _lib.ptr(data), code,
                      ctypes.byref(h)), name)
        return HostOperator(h)

    @staticmethod
    def affine(A_host, B_host=None):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.slq_host_op_create_affine(
            A_host.handle, B_host.handle if B_host is not None else None,
            ctypes.byref(h)), 'slq_host_op_create_affine')
        return HostOperator(h, parents=(A_host, B_host))

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.slq_host_op_destroy(h)
            self.handle = None


class DeviceOperator(object):
    """Owner of one ``slq_op*`` handle (device copy of the operator)."""

    def __init__(self, handle, parents=()):
        self.handle = handle
        self._parents = parents  # keep referenced handles alive

    @staticmethod
    def from_matrix(A, dense_ok=True):
        """Device copy of A: a dense handle for a numpy array (the matrix
        cores multiply it as is, imate's cuDenseMatrix), a CSC handle for a
        scipy CSC matrix (imate's cuCSCMatrix), a CSR handle for any other
        scipy sparse matrix or when ``dense_ok`` is False."""
        lib = _lib.load()
        if A.ndim != 2:
            raise ValueError('Input matrix should be a 2-dimensional array.')
        code = _lib.dtype_code(A.dtype)
        if dense_ok and not scipy.sparse.issparse(A):
            D = numpy.ascontiguousarray(A)
            h = ctypes.c_void_p()
            _lib.check(lib.slq_op_create_dense(
                D.shape[0], D.shape[1], _lib.ptr(D), code, ctypes.byref(h)),
                'slq_op_create_dense')
            return DeviceOperator(h)
        if scipy.sparse.issparse(A) and A.format == 'csc':
            # CSC handle (imate's cuCSCMatrix); the library converts it once
            indptr = numpy.ascontiguousarray(A.indptr, dtype=numpy.int64)
            indices = numpy.ascontiguousarray(A.indices, dtype=numpy.int32)
            data = numpy.ascontiguousarray(A.data, dtype=A.dtype)
            h = ctypes.c_void_p()
            _lib.check(lib.slq_op_create_csc(
                A.shape[0], A.shape[1], A.nnz, _lib.ptr(indptr),
                _lib.ptr(indices), _lib.ptr(data), code, ctypes.byref(h)),
                'slq_op_create_csc')
            return DeviceOperator(h)
        if scipy.sparse.issparse(A):
            M = scipy.sparse.csr_matrix(A)
        else:
            M = scipy.sparse.csr_matrix(numpy.asarray(A))
        if not M.has_sorted_indices:
            M = M.copy()
            M.sort_indices()
        indptr = numpy.ascontiguousarray(M.indptr, dtype=numpy.int64)
        indices = numpy.ascontiguousarray(M.indices, dtype=numpy.int32)
        data = numpy.ascontiguousarray(M.data, dtype=A.dtype)
        h = ctypes.c_void_p()
        _lib.check(lib.slq_op_create_csr(
            M.shape[0], M.shape[1], M.nnz, _lib.ptr(indptr),
            _lib.ptr(indices), _lib.ptr(data), code, ctypes.byref(h)),
            'slq_op_create_csr')
        return DeviceOperator(h)

    @staticmethod
    def affine(A_dev, B_dev=None):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.slq_op_create_affine(
            A_dev.handle, B_dev.handle if B_dev is not None else None,
            ctypes.byref(h)), 'slq_op_create_affine')
        return DeviceOperator(h, parents=(A_dev, B_dev))

    def set_values(self, data):
        lib = _lib.load()
        _lib.check(lib.slq_op_set_values(self.handle, _lib.ptr(data)),
                   'slq_op_set_values')

    def family_attach(self, d, w, c):
        """Per-entry coefficients of a value family (CSR value order)."""
        lib = _lib.load()
        arrs = [numpy.ascontiguousarray(x, dtype=numpy.float64)
                for x in (d, w, c)]
        _lib.check(lib.slq_op_family_attach(self.handle, *[_lib.ptr(x)
                                                           for x in arrs]),
                   'slq_op_family_attach')

    def family_values(self, scale):
        """Values w exp(-d / scale) + c made on the device."""
        lib = _lib.load()
        _lib.check(lib.slq_op_family_values(self.handle, float(scale)),
                   'slq_op_family_values')

    def get_device_properties(self):
        return _lib.device_properties()

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.slq_op_destroy(h)
            self.handle = None


class LinearOperator(object):
    """Base class (imate/linear_operator/linear_operator.py:14-207)."""

    def __init__(self):
        self.Aop = None
        self.cpu_Aop = None
        self.gpu_Aop = None
        self.gpu = False
        self.num_gpu_devices = 0
        self.initialized_on_cpu = False
        self.initialized_on_gpu = False
        self.data_type_name = None
        self.num_parameters = None

    def initialize(self, gpu, num_gpu_devices):
        raise NotImplementedError('This method should be called by a derived' +
                                  ' class.')

    def get_num_rows(self):
        raise NotImplementedError('This method should be called by a derived' +
                                  ' class.')

    def get_num_columns(self):
        raise NotImplementedError('This method should be called by a derived' +
                                  ' class.')

    def is_sparse(self):
        raise NotImplementedError('This method should be called by a derived' +
                                  ' class.')

    def get_nnz(self):
        raise NotImplementedError('This method should be called by a derived' +
                                  ' class.')

    def get_density(self):
        raise NotImplementedError('This method should be called by a derived' +
                                  ' class.')

    def get_linear_operator(self, gpu=False, num_gpu_devices=0):
        """The device operator (linear_operator.py:114-136)."""
        self.initialize(gpu, num_gpu_devices)
        if gpu:
            if self.gpu_Aop is None or not self.initialized_on_gpu:
                raise RuntimeError('Matrix is not initialized on gpu.')
            self.Aop = self.gpu_Aop
        else:
            if self.cpu_Aop is None or not self.initialized_on_cpu:
                raise RuntimeError('Matrix is not initialized on cpu.')
            self.Aop = self.cpu_Aop
        return self.Aop

    def set_data_type_name(self, A):
        if A.dtype == numpy.float32:
            self.data_type_name = b'float32'
        elif A.dtype == numpy.float64:
            self.data_type_name = b'float64'
        elif str(A.dtype) in ('float128', 'longdouble'):
            self.data_type_name = b'float128'
        else:
            raise TypeError('Data type should be "float32", "float64", or ' +
                            '"float128".')

    def get_data_type_name(self):
        if self.data_type_name is None:
            raise RuntimeError('"data_type_name" is None.')
        return self.data_type_name

    def get_num_parameters(self):
        if self.num_parameters is None:
            raise RuntimeError('"num_parameters" is None.')
        return self.num_parameters

    def _initialize_cpu(self):
        raise NotImplementedError('This method should be called by a '
                                  'derived class.')

    def _nnz(self, A):
        if scipy.sparse.issparse(A):
            return A.nnz
        return A.shape[0] * A.shape[1]


def _new_values(A, data):
    """Copy of the CSR matrix A with its stored values replaced by data."""
    if not (scipy.sparse.isspmatrix_csr(A) and A.has_sorted_indices):
        raise ValueError('set_values needs A as a CSR matrix with sorted '
                         'indices.')
    data = numpy.ascontiguousarray(data, dtype=A.dtype)
    if data.shape != (A.nnz,):
        raise ValueError('data should hold the %d stored values of A.' % A.nnz)
    return scipy.sparse.csr_matrix((data, A.indices, A.indptr),
                                   shape=A.shape), data


class Matrix(LinearOperator):
    """A matrix as a linear operator (imate/linear_operator/matrix.py)."""

    def __init__(self, A):
        super(Matrix, self).__init__()
        self.A = A
        self.num_parameters = 0
        self.set_data_type_name(self.A)

    def initialize(self, gpu, num_gpu_devices):
        if gpu:
            if not self.initialized_on_gpu:
                self.gpu_Aop = DeviceOperator.from_matrix(self.A)
                self.initialized_on_gpu = True
        else:
            self._initialize_cpu()
        self.gpu = gpu
        self.num_gpu_devices = num_gpu_devices

    def _initialize_cpu(self):
        if not self.initialized_on_cpu:
            self.cpu_Aop = HostOperator.from_matrix(self.A)
            self.initialized_on_cpu = True

    def set_values(self, data):
        """Replace the stored values of A (same sparsity pattern, CSR order)
        in place, also on the device, without re-uploading the pattern or
        rebuilding the tiled layout: for families of matrices that share a
        pattern. No counterpart in imate, which builds a new Matrix."""
        self.A, data = _new_values(self.A, data)
        if self.initialized_on_gpu:
            self.gpu_Aop.set_values(data)
        # the host copy is made again on its next use
        self.cpu_Aop = None
        self.initialized_on_cpu = False

    def get_num_rows(self):
        return self.A.shape[0]

    def get_num_columns(self):
        return self.A.shape[1]

    def is_sparse(self):
        return scipy.sparse.issparse(self.A)

    def get_nnz(self):
        return self._nnz(self.A)

    def get_density(self):
        if self.is_sparse():
            return self.get_nnz() / \
                (self.get_num_rows() * self.get_num_columns())
        return 1.0


class AffineMatrixFunction(LinearOperator):
    """``A + t B`` with parameter t; ``B=None`` means the identity, for which
    one Lanczos run per probe serves every t by a Ritz shift
    (imate/linear_operator/affine_matrix_function.py)."""

    def __init__(self, A, B=None):
        super(AffineMatrixFunction, self).__init__()
        self.A = A
        self.B = B
        self.num_parameters = 1
        self.set_data_type_name(self.A)
        self._A_dev = None
        self._B_dev = None

    def initialize(self, gpu, num_gpu_devices):
        if gpu:
            if not self.initialized_on_gpu:
                # A + tB with a general B runs the fused CSR SpMM of A and B
                dense_ok = self.B is None
                self._A_dev = DeviceOperator.from_matrix(self.A, dense_ok)
                if self.B is not None:
                    B = self.B
                    if B.dtype != self.A.dtype:
                        B = B.astype(self.A.dtype)
                    self._B_dev = DeviceOperator.from_matrix(B, False)
                self.gpu_Aop = DeviceOperator.affine(self._A_dev, self._B_dev)
                self.initialized_on_gpu = True
        else:
            self._initialize_cpu()
        self.gpu = gpu
        self.num_gpu_devices = num_gpu_devices

    def _initialize_cpu(self):
        if not self.initialized_on_cpu:
            A_host = HostOperator.from_matrix(self.A)
            B_host = None
            if self.B is not None:
                B = self.B
                if B.dtype != self.A.dtype:
                    B = B.astype(self.A.dtype)
                B_host = HostOperator.from_matrix(B)
            self.cpu_Aop = HostOperator.affine(A_host, B_host)
            self.initialized_on_cpu = True

    def set_values(self, data):
        """Replace the stored values of A (not B) in place; see
        ``Matrix.set_values``."""
        self.A, data = _new_values(self.A, data)
        self.cpu_Aop = None
        self.initialized_on_cpu = False
        if self.initialized_on_gpu:
            self._A_dev.set_values(data)

    def attach_family(self, d, w, c):
        """A value family of A (one pattern; per-entry coefficients d, w, c
        in A's CSR value order): ``set_family_scale(l)`` then makes the
        values w exp(-d / l) + c on the device, with no host values and no
        upload. No counterpart in imate."""
        if not (scipy.sparse.isspmatrix_csr(self.A) and
                self.A.has_sorted_indices):
            raise ValueError('a value family needs A as a CSR matrix with '
                             'sorted indices.')
        self.initialize(True, 0)
        self._A_dev.family_attach(d, w, c)

    def set_family_scale(self, scale):
        """New device values of A from its family at ``scale``. The host
        copy of A keeps its old values (the device copy is the operator)."""
        self._A_dev.family_values(scale)

    def get_num_rows(self):
        return self.A.shape[0]

    def get_num_columns(self):
        return self.A.shape[1]

    def is_sparse(self):
        return scipy.sparse.issparse(self.A)

    def get_nnz(self):
        return self._nnz(self.A)

    def get_density(self):
        if self.is_sparse():
            return self.get_nnz() / \
                (self.get_num_rows() * self.get_num_columns())
        return 1.0


class DeviceMatrix(LinearOperator):
    """A matrix that exists only in device memory (for instance generated on
    the device, ``sample_matrices.sparse_gaussian_device``): the ``Matrix``
    interface without a host copy. ``to_csr()`` copies it back."""

    def __init__(self, device_operator):
        super(DeviceMatrix, self).__init__()
        lib = _lib.load()
        nr = ctypes.c_int64(0)
        nc = ctypes.c_int64(0)
        nnz = ctypes.c_int64(0)
        dt = ctypes.c_int32(0)
        _lib.check(lib.slq_op_shape(device_operator.handle, ctypes.byref(nr),
                                    ctypes.byref(nc), ctypes.byref(nnz),
                                    ctypes.byref(dt)), 'slq_op_shape')
        self.shape = (nr.value, nc.value)
        self.nnz = nnz.value
        self.dtype = numpy.dtype(numpy.float64 if dt.value == _lib.SLQ_F64
                                 else numpy.float32)
        self.data_type_name = self.dtype.name.encode()
        self.num_parameters = 0
        self.gpu_Aop = device_operator
        self.initialized_on_gpu = True

    def initialize(self, gpu, num_gpu_devices):
        if not gpu:
            self._initialize_cpu()
        self.gpu = gpu
        self.num_gpu_devices = num_gpu_devices

    def _initialize_cpu(self):
        # the CPU path needs a host copy: one copy back from the device
        if not self.initialized_on_cpu:
            self.cpu_Aop = HostOperator.from_matrix(self.to_csr())
            self.initialized_on_cpu = True

    def get_num_rows(self):
        return self.shape[0]

    def get_num_columns(self):
        return self.shape[1]

    def is_sparse(self):
        return True

    def get_nnz(self):
        return self.nnz

    def get_density(self):
        return self.nnz / (self.shape[0] * self.shape[1])

    def to_csr(self):
        """The matrix as a scipy CSR matrix (a copy from the device)."""
        lib = _lib.load()
        indptr = numpy.empty(self.shape[0] + 1, dtype=numpy.int64)
        indices = numpy.empty(self.nnz, dtype=numpy.int32)
        data = numpy.empty(self.nnz, dtype=self.dtype)
        _lib.check(lib.slq_op_get_csr(self.gpu_Aop.handle, _lib.ptr(indptr),
                                      _lib.ptr(indices), _lib.ptr(data)),
                   'slq_op_get_csr')
        return scipy.sparse.csr_matrix((data, indices, indptr),
                                       shape=self.shape)
