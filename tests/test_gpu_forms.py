"""GPU tests of the batched and mixed-precision forms."""

import numpy
import pytest
import scipy.sparse

import imate_amd
from imate_amd.batch import slq_batch
from imate_amd.sample_matrices import sparse_correlation_matrix
from oracle import oracle

pytestmark = pytest.mark.gpu


def _mats(sizes, k=12, dtype='float64'):
    return [sparse_correlation_matrix(n, nnz_per_row=k, seed=i, dtype=dtype)
            for i, n in enumerate(sizes)]


@pytest.mark.parametrize('use_graph', [True, False])
def test_batch_matches_oracle_per_problem(use_graph):
    mats = _mats([300, 1500, 4096, 777])
    tr, info = slq_batch(mats, func='log', lanczos_degree=15, num_samples=70,
                         seed=5, use_graph=use_graph)
    for i, A in enumerate(mats):
        ref = oracle.slq(A, func='log', lanczos_degree=15, min_num_samples=70,
                         max_num_samples=70, seed=5)
        got = info['samples'][i]
        want = ref['samples'][:, 0]
        rel = numpy.abs(got - want) / numpy.abs(want)
        assert rel.max() < 1e-10, (i, rel.max())
        assert abs(tr[i] - ref['trace'][0]) <= 1e-10 * abs(ref['trace'][0])
        assert abs(info['error'][i] - ref['error'][0]) <= 1e-8 * ref['error'][0]


def test_batch_graph_equals_loop():
    mats = _mats([2000, 500, 3000])
    a, ia = slq_batch(mats, func='inv', lanczos_degree=12, num_samples=64,
                      seed=2, use_graph=True, autotune=False)
    b, ib = slq_batch(mats, func='inv', lanczos_degree=12, num_samples=64,
                      seed=2, use_graph=False, autotune=False)
    assert numpy.array_equal(ia['samples'], ib['samples'])


def test_batch_float32_mixed():
    mats = _mats([1000, 2500], dtype='float32')
    tr, info = slq_batch(mats, func='log', lanczos_degree=12, num_samples=32,
                         seed=1)
    for i, A in enumerate(mats):
        ref = oracle.slq(A, func='log', lanczos_degree=12, min_num_samples=32,
                         max_num_samples=32, seed=1)
        rel = numpy.abs(info['samples'][i] - ref['samples'][:, 0]) / \
            numpy.abs(ref['samples'][:, 0])
        assert rel.max() < 1e-4


def test_ritz_shift_equals_direct_runs():
    """A + tI through one Lanczos run and Ritz shifts equals separate runs
    on the shifted matrices, per sample (same probes)."""
    A = sparse_correlation_matrix(3000, nnz_per_row=20, seed=3)
    ts = numpy.logspace(-4, 2, 16)
    Aop = imate_amd.AffineMatrixFunction(A)
    val, info = imate_amd.logdet(Aop, parameters=ts, method='slq', gpu=True,
                                 lanczos_degree=20, min_num_samples=32,
                                 max_num_samples=32, seed=4, return_info=True)
    S = info['convergence']['samples']
    for j in (0, 5, 15):
        At = (A + ts[j] * scipy.sparse.identity(A.shape[0])).tocsr()
        v, i2 = imate_amd.logdet(At, method='slq', gpu=True, lanczos_degree=20,
                                 min_num_samples=32, max_num_samples=32,
                                 seed=4, return_info=True)
        # Lanczos on A + t0 I and on A + tj I from the same probe produce the
        # same Krylov basis in exact arithmetic; rounding differs slightly
        rel = numpy.abs(S[:, j] - i2['convergence']['samples']) / \
            numpy.abs(i2['convergence']['samples'])
        assert rel.max() < 1e-9, (j, rel.max())


def test_mixed_precision_with_reorthogonalization():
    A = sparse_correlation_matrix(4096, nnz_per_row=40, seed=0)
    kw = dict(method='slq', gpu=True, lanczos_degree=30, min_num_samples=64,
              max_num_samples=64, seed=5, orthogonalize=-1, return_info=True)
    v64, i64 = imate_amd.logdet(A, **kw)
    v32, i32 = imate_amd.logdet(A, precision='mixed', **kw)
    rel = numpy.abs(i32['convergence']['samples'] -
                    i64['convergence']['samples']) / \
        numpy.abs(i64['convergence']['samples'])
    assert rel.max() < 1e-5


def test_autotune_keeps_results():
    A = sparse_correlation_matrix(5000, nnz_per_row=30, seed=7)
    M = imate_amd.Matrix(A)
    kw = dict(method='slq', gpu=True, lanczos_degree=15, min_num_samples=64,
              max_num_samples=64, seed=1, return_info=True)
    v1, i1 = imate_amd.logdet(M, **kw)
    from imate_amd import _lib
    import ctypes
    sg, ug = ctypes.c_int32(0), ctypes.c_int32(0)
    _lib.check(_lib.load().slq_autotune(M.gpu_Aop.handle, 0, ctypes.byref(sg),
                                        ctypes.byref(ug)), 'autotune')
    assert sg.value >= 8 and ug.value >= 8
    v2, i2 = imate_amd.logdet(M, **kw)
    # the tuned caps apply to batches only: estimates on the handle keep
    # their reduction order, so every sample is the same bit for bit
    assert numpy.array_equal(i1['convergence']['samples'],
                             i2['convergence']['samples'])


def test_batch_autotune_leaves_handles_untouched():
    mats = _mats([3000, 1200])
    Ms = [imate_amd.Matrix(A) for A in mats]
    kw = dict(method='slq', gpu=True, lanczos_degree=12, min_num_samples=64,
              max_num_samples=64, seed=3, return_info=True)
    before = [imate_amd.logdet(M, **kw)[1]['convergence']['samples']
              for M in Ms]
    tr, info = slq_batch(Ms, func='log', lanczos_degree=12, num_samples=64,
                         seed=3, autotune=True)
    assert len(info['grids']) == 2
    after = [imate_amd.logdet(M, **kw)[1]['convergence']['samples']
             for M in Ms]
    for b, a in zip(before, after):
        assert numpy.array_equal(b, a)
    # the batch itself (tuned grids) still matches per sample
    for i in range(2):
        rel = numpy.abs(info['samples'][i] - before[i]) / numpy.abs(before[i])
        assert rel.max() < 1e-10


def test_batch_gram_rectangular_matches_oracle():
    rng = numpy.random.default_rng(11)
    mats = []
    for i, (r, c) in enumerate(((900, 300), (400, 250), (1200, 1200))):
        A = scipy.sparse.random(r, c, density=0.05, random_state=i,
                                data_rvs=rng.standard_normal, format='csr')
        A = (A + scipy.sparse.eye(r, c, format='csr')).tocsr()
        mats.append(A)
    tr, info = slq_batch(mats, func='log', gram=True, lanczos_degree=10,
                         num_samples=40, seed=4)
    assert info['grids'] == []
    for i, A in enumerate(mats):
        ref = oracle.slq(A, func='log', gram=True, lanczos_degree=10,
                         min_num_samples=40, max_num_samples=40, seed=4)
        want = ref['samples'][:, 0]
        rel = numpy.abs(info['samples'][i] - want) / numpy.abs(want)
        assert rel.max() < 1e-10, (i, rel.max())
