"""Sample matrices for tests and benchmarks.

* ``toeplitz(a, b, size, gram, format, dtype)`` and its closed forms
  ``toeplitz_logdet / traceinv / trace / schatten`` — the reference's test
  matrices (imate/sample_matrices/toeplitz.py:25-202,273-635): the upper
  bidiagonal matrix with a on the diagonal and b above it, or its Gramian.
* ``sparse_correlation_matrix(...)`` — a seeded, spatially ordered sparse
  correlation matrix (the construction of the benchmark workloads).
  The reference's own ``correlation_matrix`` is time-seeded with an O(n^2)
  pair loop; this one is seeded, O(n k) with a KD-tree, and positive definite
  by construction:

      K_ij = exp(-d_ij / l) * w(d_ij / R),   R = l * ln(1/tau),
      w(r) = (1 - r)^4 (4 r + 1) for r < 1, else 0   (Wendland, C^2)

  The compactly supported Wendland factor is a positive-definite kernel in
  up to three dimensions, so the Schur product with the exponential kernel
  keeps K positive semi-definite while giving exactly the sparsity of the
  hard threshold K >= tau; the diagonal is 1 + jitter. Points are uniform in
  [0, 1]^2 (numpy default_rng(seed)) and sorted along a Hilbert curve
  (``order='hilbert'``, the default) or a Morton (Z-order) curve
  (``order='morton'``), so that neighbouring rows share columns — the
  property that keeps the SpMM's gathers in cache. The Hilbert curve never
  jumps, so 16 consecutive rows span fewer columns: at 50 non-zeros per row
  their union is 118 columns against 131 in Z-order. Both orders give the
  same matrix up to a symmetric permutation (the same spectrum). The length
  scale for about k non-zeros per row is l = sqrt(k / (pi n)) / ln(1/tau).
* ``CorrelationFamily(n, support, ...)`` — the matrices above over the
  length scale l with the support radius R held fixed, so that every member
  shares one sparsity pattern (K_ij = exp(-d_ij / l) w(d_ij / R), still
  positive definite) and a new l only needs new values.
* ``sparse_gaussian(nrows, ncols, density, seed)`` — rows with
  round(density * ncols) distinct uniform columns and N(0, 1) values (the
  rectangular workload of gram=True), built on the host.
* ``sparse_gaussian_device(nrows, ncols, density, seed)`` — the large
  rectangular workload built on the device (a ``DeviceMatrix``): row r has
  one entry in each of k = round(density * ncols) column strata, at a hashed
  column of the stratum, with a hashed N(0, 1) value.
"""

import numpy
import scipy.sparse
from scipy.sparse import coo_matrix

__all__ = ['toeplitz', 'toeplitz_trace', 'toeplitz_traceinv',
           'toeplitz_logdet', 'toeplitz_schatten', 'correlation_scale',
           'sparse_correlation_matrix', 'CorrelationFamily',
           'sparse_gaussian', 'sparse_gaussian_device', 'morton_order',
           'hilbert_order']


# ---------------------------------------------------------------- Toeplitz

def toeplitz(a, b, size=20, gram=False, format='csr', dtype=r'float64'):
    """Upper bidiagonal Toeplitz matrix (toeplitz.py:25-202)."""
    if size < 1:
        raise ValueError('"size" should be a positive integer.')
    if format not in ('csr', 'csc'):
        raise ValueError('"format" should be either "csr" or "csc".')
    if dtype not in (r'float32', r'float64', r'float128'):
        raise TypeError('"dtype" should be either "float32", "float64", or ' +
                        '"float128".')
    n = int(size)
    rows = numpy.concatenate([numpy.arange(n), numpy.arange(n - 1)])
    cols = numpy.concatenate([numpy.arange(n), numpy.arange(1, n)])
    vals = numpy.concatenate([numpy.full(n, a, dtype=dtype),
                              numpy.full(n - 1, b, dtype=dtype)])
    A = coo_matrix((vals, (rows, cols)), shape=(n, n))
    if gram:
        A = A.T @ A
    return A.tocsc() if format == 'csc' else A.tocsr()


def toeplitz_trace(a, b, size, gram=False):
    """toeplitz.py:273-348."""
    if gram:
        return a**2 + (size - 1) * (a**2 + b**2)
    return size * a


def toeplitz_traceinv(a, b, size, gram=False):
    """toeplitz.py:355-460 (the reference's closed form of tr((A^T A)^-1))."""
    if not gram:
        return size / a
    if a == b:
        return size * (size + 1) / (2.0 * a**2)
    q = b / a
    if size < 200:
        return (1.0 / (a**2 - b**2)) * \
            (size - (q**2) * ((q**(2.0 * size) - 1.0) / (q**2 - 1.0)))
    return (1.0 / (a**2 - b**2)) * (size - ((q**2) / (1.0 - q**2)))


def toeplitz_logdet(a, b, size, gram=False):
    """toeplitz.py:467-540: n log a, doubled for the Gramian."""
    ld = size * numpy.log(a)
    return 2.0 * ld if gram else ld


def toeplitz_schatten(a, b, size, p=2):
    """toeplitz.py:547-635 (the reference's large-n value)."""
    return a


# ---------------------------------------------------------------- correlation

def morton_order(points, bits=16):
    """Permutation sorting 2-D points in [0, 1]^2 along the Z-order curve."""
    q = numpy.clip((points * (1 << bits)).astype(numpy.int64), 0,
                   (1 << bits) - 1).astype(numpy.uint64)

    def spread(v):
        v = v & numpy.uint64(0xFFFFFFFF)
        v = (v | (v << numpy.uint64(16))) & numpy.uint64(0x0000FFFF0000FFFF)
        v = (v | (v << numpy.uint64(8))) & numpy.uint64(0x00FF00FF00FF00FF)
        v = (v | (v << numpy.uint64(4))) & numpy.uint64(0x0F0F0F0F0F0F0F0F)
        v = (v | (v << numpy.uint64(2))) & numpy.uint64(0x3333333333333333)
        v = (v | (v << numpy.uint64(1))) & numpy.uint64(0x5555555555555555)
        return v

    key = spread(q[:, 0]) | (spread(q[:, 1]) << numpy.uint64(1))
    return numpy.argsort(key, kind='stable')


def hilbert_order(points, bits=16):
    """Permutation sorting 2-D points in [0, 1]^2 along the Hilbert curve.

    Each point is snapped to a 2^bits x 2^bits grid and given its distance d
    along the curve: from the top bit down, the quadrant (rx, ry) adds
    s^2 ((3 rx) xor ry) to d, and the coordinates are reflected and swapped
    into that quadrant's frame (the curve's standard recursive definition:
    Hilbert, 1891; Butz, IEEE Trans. Computers C-20, 1971). Consecutive grid
    cells along the curve are edge neighbours.
    """
    side = 1 << bits
    q = numpy.clip((points * side).astype(numpy.int64), 0, side - 1)
    x = q[:, 0].copy()
    y = q[:, 1].copy()
    d = numpy.zeros(len(x), dtype=numpy.int64)
    s = side >> 1
    while s > 0:
        rx = (x & s) > 0
        ry = (y & s) > 0
        d += s * s * ((3 * rx.astype(numpy.int64)) ^ ry.astype(numpy.int64))
        low = ~ry
        flip = low & rx
        x[flip] = side - 1 - x[flip]
        y[flip] = side - 1 - y[flip]
        x[low], y[low] = y[low], x[low].copy()
        s >>= 1
    return numpy.argsort(d, kind='stable')


_ORDERS = {'hilbert': hilbert_order, 'morton': morton_order}


def _spatial_order(points, order):
    if order not in _ORDERS:
        raise ValueError('order must be one of %s.' % sorted(_ORDERS))
    return _ORDERS[order](points)


def correlation_scale(n, nnz_per_row, threshold=0.03):
    """Length scale l giving about ``nnz_per_row`` entries per row."""
    return numpy.sqrt(nnz_per_row / (numpy.pi * n)) / numpy.log(1.0 / threshold)


def sparse_correlation_matrix(n, scale=None, nnz_per_row=40, threshold=0.03,
                              jitter=1e-3, seed=0, dtype='float64',
                              taper=True, order='hilbert'):
    """Seeded sparse SPD correlation matrix (CSR, sorted indices).

    ``scale`` is the kernel length scale l (derived from ``nnz_per_row``
    when None). ``taper=False`` gives the plain hard threshold
    exp(-d/l) >= tau, which may be indefinite for dense neighbourhoods.
    ``order`` is the curve the points are sorted along ('hilbert' or
    'morton').
    """
    from scipy.spatial import cKDTree
    if scale is None:
        scale = correlation_scale(n, nnz_per_row, threshold)
    rng = numpy.random.default_rng(seed)
    pts = rng.random((n, 2))
    pts = pts[_spatial_order(pts, order)]
    R = scale * numpy.log(1.0 / threshold)
    tree = cKDTree(pts)
    pairs = tree.query_pairs(R, output_type='ndarray')
    i, j = pairs[:, 0], pairs[:, 1]
    d = numpy.sqrt(((pts[i] - pts[j])**2).sum(axis=1))
    v = numpy.exp(-d / scale)
    if taper:
        r = d / R
        v = v * (1.0 - r)**4 * (4.0 * r + 1.0)
    rows = numpy.concatenate([i, j, numpy.arange(n)])
    cols = numpy.concatenate([j, i, numpy.arange(n)])
    vals = numpy.concatenate([v, v, numpy.full(n, 1.0 + jitter)])
    A = scipy.sparse.csr_matrix((vals.astype(dtype), (rows, cols)),
                                shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A


class CorrelationFamily(object):
    """Correlation matrices over the length scale l with one fixed pattern.

    Points as in ``sparse_correlation_matrix`` (seeded, sorted along the
    curve ``order``);
    pairs within the support radius ``support`` are the pattern, and

        K_ij(l) = exp(-d_ij / l) * w(d_ij / support),  K_ii = 1 + jitter,

    with the Wendland factor w, so every member is positive definite.
    ``matrix(l)`` returns the CSR matrix (sorted indices, no duplicates);
    ``values(l)`` only its value array, in the same order.
    """

    def __init__(self, n, support, jitter=1e-3, seed=0, dtype='float64',
                 order='hilbert'):
        from scipy.spatial import cKDTree
        rng = numpy.random.default_rng(seed)
        pts = rng.random((n, 2))
        pts = pts[_spatial_order(pts, order)]
        pairs = cKDTree(pts).query_pairs(support, output_type='ndarray')
        i, j = pairs[:, 0], pairs[:, 1]
        d = numpy.sqrt(((pts[i] - pts[j])**2).sum(axis=1))
        rows = numpy.concatenate([i, j, numpy.arange(n)])
        cols = numpy.concatenate([j, i, numpy.arange(n)])
        dist = numpy.concatenate([d, d, numpy.full(n, -1.0)])
        order = numpy.lexsort((cols, rows))
        self.n = n
        self.support = float(support)
        self.jitter = float(jitter)
        self.dtype = numpy.dtype(dtype)
        self.indices = cols[order].astype(numpy.int32)
        self.indptr = numpy.zeros(n + 1, dtype=numpy.int64)
        numpy.cumsum(numpy.bincount(rows, minlength=n), out=self.indptr[1:])
        dist = dist[order]
        self._diag = dist < 0
        self._d = numpy.where(self._diag, 0.0, dist)
        r = self._d / self.support
        self._w = (1.0 - r)**4 * (4.0 * r + 1.0)
        self.nnz = len(self.indices)

    def coefficients(self):
        """(d, w, c) in CSR value order with every value equal to
        w exp(-d / l) + c: d the distance and w the Wendland factor off the
        diagonal, c = 1 + jitter on it (d = w = 0 there); for a device value
        family (``AffineMatrixFunction.attach_family``)."""
        w = numpy.where(self._diag, 0.0, self._w)
        c = numpy.where(self._diag, 1.0 + self.jitter, 0.0)
        return self._d, w, c

    def values(self, scale):
        v = numpy.exp(-self._d / scale) * self._w
        v[self._diag] = 1.0 + self.jitter
        return v.astype(self.dtype)

    def matrix(self, scale):
        return scipy.sparse.csr_matrix(
            (self.values(scale), self.indices, self.indptr),
            shape=(self.n, self.n))

    def frobenius_squared(self, scale):
        """tr(K(l)^2): the sum of the squared entries."""
        v = self.values(scale).astype(numpy.float64)
        return float(numpy.dot(v, v))


def sparse_gaussian(nrows, ncols, density, seed=0, dtype='float64'):
    """Rows of round(density * ncols) distinct uniform columns, N(0,1) values."""
    rng = numpy.random.default_rng(seed)
    k = max(1, int(round(density * ncols)))
    cols = numpy.empty((nrows, k), dtype=numpy.int64)
    for r in range(nrows):
        cols[r] = numpy.sort(rng.choice(ncols, size=k, replace=False))
    vals = rng.standard_normal((nrows, k)).astype(dtype)
    indptr = numpy.arange(0, nrows * k + 1, k, dtype=numpy.int64)
    return scipy.sparse.csr_matrix((vals.ravel(), cols.ravel(), indptr),
                                   shape=(nrows, ncols))


def sparse_gaussian_device(nrows, ncols, density, seed=0, dtype='float64'):
    """Seeded sparse Gaussian matrix generated in device memory
    (``slq_op_create_sparse_gaussian``): k = round(density * ncols) entries
    per row, one in each column stratum [floor(s ncols / k),
    floor((s + 1) ncols / k)), N(0, 1) values. Returns a ``DeviceMatrix``."""
    import ctypes
    from . import _lib
    from .linear_operator import DeviceOperator, DeviceMatrix
    k = max(1, int(round(density * ncols)))
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.slq_op_create_sparse_gaussian(
        int(nrows), int(ncols), k, int(seed), _lib.dtype_code(dtype),
        ctypes.byref(h)), 'slq_op_create_sparse_gaussian')
    return DeviceMatrix(DeviceOperator(h))
