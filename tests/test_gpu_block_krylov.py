"""The block Lanczos form (krylov='block', slq_estimate_block_krylov)
against the NumPy statement in tests/block_lanczos_ref.py on the same
probes: per sample, 1e-9 relative in FP64 (the two differ in the order of
the GEMM sums and in the eigensolver), 1e-4 for FP32 vectors and basis
(matrix-core FP32 GEMMs), and exactly for f(x) = x^2, which block Gauss
quadrature integrates exactly from two block steps."""

import numpy
import pytest
import scipy.sparse

import imate_amd
from imate_amd.sample_matrices import sparse_correlation_matrix
from oracle import oracle
from tests.block_lanczos_ref import samples as ref_samples

pytestmark = pytest.mark.gpu

B = 64


def _ref(A, seed, count, m, f, exponent=1.0, params=None):
    n = A.shape[0]
    nb = (count + B - 1) // B
    Z = oracle.rademacher(seed, n, nb * B).T
    rows = [ref_samples(A, Z[:, k * B:(k + 1) * B], m, f, exponent, params)
            for k in range(nb)]
    return numpy.vstack(rows)[:count]


def _run(fn, A, m, s, seed, **kw):
    val, info = fn(A, method='slq', gpu=True, krylov='block',
                   lanczos_degree=m, min_num_samples=s, max_num_samples=s,
                   seed=seed, return_info=True, **kw)
    got = numpy.asarray(info['convergence']['samples'],
                        dtype=numpy.float64).reshape(s, -1)
    return val, got


def _rel(got, want):
    return float(numpy.max(numpy.abs(got - want) /
                           numpy.maximum(1.0, numpy.abs(want))))


@pytest.fixture(scope='module')
def c1():
    return sparse_correlation_matrix(4096, nnz_per_row=40, seed=0)


@pytest.mark.parametrize('m', [4, 12])
def test_logdet_matches_numpy(c1, m):
    val, got = _run(imate_amd.logdet, c1, m, 128, 5)
    want = _ref(c1, 5, 128, m, numpy.log)
    assert _rel(got, want) < 1e-9
    assert abs(val - want.mean()) <= 1e-9 * abs(want.mean())


def test_traceinv_and_partial_last_block(c1):
    # 100 samples: the second block still runs all 64 probes
    val, got = _run(imate_amd.traceinv, c1, 8, 100, 6)
    want = _ref(c1, 6, 100, 8, lambda x: 1.0 / x)
    assert _rel(got, want) < 1e-9


def test_affine_identity_ritz_shift(c1):
    ts = [0.05, 0.5, 3.0]
    op = imate_amd.AffineMatrixFunction(c1)
    val, got = _run(imate_amd.logdet, op, 8, 64, 7, parameters=ts)
    want = _ref(c1, 7, 64, 8, numpy.log, params=ts)
    assert _rel(got, want) < 1e-9


def test_square_is_exact(c1):
    val, got = _run(imate_amd.trace, c1, 2, 64, 8, p=2)
    Z = oracle.rademacher(8, c1.shape[0], 64).T
    AZ = c1 @ Z
    want = numpy.einsum('ip,ip->p', AZ, AZ)
    assert _rel(got[:, 0], want) < 1e-10


@pytest.mark.parametrize('how', ['float32', 'mixed'])
def test_fp32_forms(c1, how):
    if how == 'float32':
        A, kw = c1.astype(numpy.float32), {}
    else:
        A, kw = c1, {'precision': 'mixed'}
    val, got = _run(imate_amd.logdet, A, 8, 64, 9, **kw)
    want = _ref(c1, 9, 64, 8, numpy.log)
    assert _rel(got, want) < 1e-4


def test_rejects_what_it_does_not_serve(c1):
    small = sparse_correlation_matrix(300, nnz_per_row=20, seed=1)
    with pytest.raises(ValueError):
        imate_amd.logdet(small, method='slq', gpu=True, krylov='block',
                         lanczos_degree=5)       # 5 x 64 > 300 rows
    with pytest.raises(ValueError):
        imate_amd.logdet(c1, method='slq', gpu=True, krylov='block',
                         gram=True)
    with pytest.raises(ValueError):
        imate_amd.logdet(c1, method='slq', gpu=False, krylov='block')


def test_gemm_profile_counts_the_block_products(c1):
    """slq_profile_gemm: one block of 64 probes, m = 4 block steps: Cholesky
    QR twice of the probe block and of steps 0..2 (2 Q^T W and 2 W R^-1
    products each) and two projection passes per step over j + 1 blocks."""
    import ctypes
    from imate_amd import _lib
    lib = _lib.load()
    M = imate_amd.Matrix(c1)
    h = M.get_linear_operator(gpu=True).handle
    _lib.check(lib.slq_profile_enable(h, 1), 'slq_profile_enable')
    _run(imate_amd.logdet, M, 4, 64, 3)
    ms, nl, fl = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
    _lib.check(lib.slq_profile_gemm(h, ctypes.byref(ms), ctypes.byref(nl),
                                    ctypes.byref(fl)), 'slq_profile_gemm')
    _lib.check(lib.slq_profile_enable(h, 0), 'slq_profile_enable')
    m, n = 4, c1.shape[0]
    units = 2 * (2 + sum(2 * (j + 1) for j in range(m)) + 2 * (m - 1))
    assert nl.value == 2 * (2 + 2 * m + 2 * (m - 1))
    assert fl.value == units * 2.0 * n * 64 * 64
    assert ms.value > 0.0


def test_affine_general_b_each_t(c1):
    """A + tB with a general B: one block recurrence per t (no Ritz shift),
    each in a slot of its own, the two eigensolved in one batch."""
    n = c1.shape[0]
    B = scipy.sparse.diags(numpy.linspace(0.5, 2.0, n)).tocsr()
    ts = [0.1, 1.5]
    op = imate_amd.AffineMatrixFunction(c1, B)
    val, got = _run(imate_amd.logdet, op, 6, 64, 12, parameters=ts)
    Z = oracle.rademacher(12, n, 64).T
    for j, t in enumerate(ts):
        want = ref_samples((c1 + t * B).tocsr(), Z, 6, numpy.log)[:, 0]
        assert _rel(got[:, j], want) < 1e-9


@pytest.mark.parametrize('nt,count', [(9, 128), (3, 192)])
def test_slot_groups(c1, nt, count):
    """More recurrences per block than slots (9 t: eight eigensolved in one
    batch, then one), and groups of two blocks with three t each (six of
    the eight slots) followed by a group of one: every sample equals the
    NumPy statement of its block and t."""
    n = c1.shape[0]
    D = scipy.sparse.diags(numpy.linspace(0.5, 2.0, n)).tocsr()
    ts = list(numpy.linspace(0.05, 2.0, nt))
    op = imate_amd.AffineMatrixFunction(c1, D)
    val, got = _run(imate_amd.traceinv, op, 5, count, 21, parameters=ts)
    Z = oracle.rademacher(21, n, count).T
    for j, t in enumerate(ts):
        At = (c1 + t * D).tocsr()
        want = numpy.concatenate(
            [ref_samples(At, Z[:, k:k + B], 5, lambda x: 1.0 / x)[:, 0]
             for k in range(0, count, B)])
        assert _rel(got[:, j], want) < 1e-9


def test_convergence_stops_early(c1):
    """min < max: the per-sample convergence test stops the loop at the
    first sample after which the inquiry converged (blocks still run all 64
    probes); samples before it are the block form's, later ones NaN."""
    val, info = imate_amd.traceinv(c1, method='slq', gpu=True, krylov='block',
                                   lanczos_degree=6, min_num_samples=64,
                                   max_num_samples=640, error_rtol=0.05,
                                   seed=13, return_info=True)
    used = int(info['convergence']['num_samples_used'])
    assert 64 <= used < 640
    got = numpy.asarray(info['convergence']['samples'], dtype=numpy.float64)
    assert numpy.all(numpy.isnan(got[used:]))
    want = _ref(c1, 13, used, 6, lambda x: 1.0 / x)
    assert _rel(got[:used].reshape(used, -1), want) < 1e-9


def test_grouping_leaves_samples_unchanged(c1):
    """Blocks eigensolved one at a time (min_num_samples = 64: every block
    may stop the loop, so each is its own group) against all four in one
    batched group (min_num_samples = 256); with error_rtol = 0 neither run
    stops early, and the samples agree to rounding of the eigensolver."""
    kw = dict(method='slq', gpu=True, krylov='block', lanczos_degree=6,
              max_num_samples=256, error_rtol=0.0, error_atol=0.0, seed=17,
              return_info=True)
    _, one = imate_amd.logdet(c1, min_num_samples=64, **kw)
    _, grp = imate_amd.logdet(c1, min_num_samples=256, **kw)
    a = numpy.asarray(one['convergence']['samples'], dtype=numpy.float64)
    b = numpy.asarray(grp['convergence']['samples'], dtype=numpy.float64)
    assert int(one['convergence']['num_samples_used']) == 256
    assert _rel(a, b) < 1e-12


@pytest.mark.parametrize('form', ['dense', 'csc'])
def test_dense_and_csc_operators(c1, form):
    """The block form over the dense operator (matrix-core dense SpMM with
    column splits at this size) and over a CSC input (converted to CSR at
    operator creation): the NumPy statement's samples at 1e-9."""
    A = c1.toarray() if form == 'dense' else c1.tocsc()
    val, got = _run(imate_amd.logdet, A, 6, 64, 23)
    want = _ref(c1, 23, 64, 6, numpy.log)
    assert _rel(got, want) < 1e-9
