"""Learned form on the GPU: in-place value refresh of a fixed-pattern
operator, and what the trained surrogate gains over plain SLQ on a small
matrix family (n = 2048) whose exact values come from dense eigenvalues."""

import time

import numpy
import pytest

import imate_amd
from imate_amd.sample_matrices import CorrelationFamily
from imate_amd.surrogate import LogdetSurrogate, sample_queries
from oracle import oracle

pytestmark = pytest.mark.gpu


def _slq(A, **kw):
    args = dict(method='slq', gpu=True, lanczos_degree=20, min_num_samples=64,
                max_num_samples=64, seed=3)
    args.update(kw)
    return imate_amd.logdet(A, **args)


def test_set_values_matches_new_operator(gpu_lib):
    fam = CorrelationFamily(3000, 0.06, seed=1)
    M = imate_amd.Matrix(fam.matrix(0.02))
    first = _slq(M)
    M.set_values(fam.values(0.05))
    refreshed = _slq(M)
    fresh = _slq(imate_amd.Matrix(fam.matrix(0.05)))
    assert refreshed == fresh and refreshed != first
    # against the oracle on the same probes
    ref = oracle.slq(fam.matrix(0.05), func='log', lanczos_degree=20,
                     min_num_samples=64, max_num_samples=64, seed=3)
    ref = ref['trace'][0]
    assert abs(refreshed - ref) <= 1e-10 * abs(ref)


@pytest.mark.parametrize('kw', [dict(gram=True), dict(precision='mixed'),
                                dict(gram=True, precision='mixed')])
def test_set_values_rebuilds_derived_forms(gpu_lib, kw):
    fam = CorrelationFamily(3000, 0.06, seed=2)
    M = imate_amd.Matrix(fam.matrix(0.02))
    _slq(M, **kw)                      # builds the transpose / FP32 copies
    M.set_values(fam.values(0.07))
    assert _slq(M, **kw) == _slq(imate_amd.Matrix(fam.matrix(0.07)), **kw)


def test_set_values_affine_shares_matrix(gpu_lib):
    fam = CorrelationFamily(3000, 0.06, seed=3)
    op = imate_amd.AffineMatrixFunction(fam.matrix(0.02))
    ts = [1e-3, 0.1, 10.0]
    _slq(op, parameters=ts)
    op.set_values(fam.values(0.04))
    a = _slq(op, parameters=ts)
    b = _slq(imate_amd.AffineMatrixFunction(fam.matrix(0.04)), parameters=ts)
    assert numpy.array_equal(a, b)


def test_set_values_rejects_bad_input(gpu_lib):
    fam = CorrelationFamily(500, 0.1, seed=4)
    M = imate_amd.Matrix(fam.matrix(0.02))
    _slq(M)
    with pytest.raises(ValueError):
        M.set_values(numpy.ones(fam.nnz - 1))


def test_surrogate_gain_over_slq(gpu_lib):
    n = 2048
    fam = CorrelationFamily(n, 0.08, seed=5)
    sur = LogdetSurrogate(fam, lanczos_degree=20, num_samples=64, seed=0)
    train = sur.generate(16, 25, seed=1)
    sur.train(train, steps=(4000, 4000, 4000), lr=(3e-3, 1e-3, 3e-4))
    assert sur.losses[-1] < 1e-3 * sur.losses[0], (sur.losses[0], sur.losses[-1])
    sur.calibrate(sur.generate(6, 20, seed=2))
    ls, ts = sample_queries(4, 8, sur.scale_range, sur.t_range, seed=3)
    hs, ht = numpy.repeat(ls, 8), ts.ravel()
    exact = numpy.empty(len(ht))
    for i, s in enumerate(ls):
        lam = numpy.linalg.eigvalsh(fam.matrix(s).toarray())
        exact[8 * i:8 * i + 8] = [numpy.log(lam + t).sum() for t in ts[i]]
    y, info = sur.query(hs, ht, judge='slq', seed=77)
    pred, slq, ci = info['surrogate'], info['slq'], info['ci']
    err_sur = numpy.abs(pred - exact)
    err_slq = numpy.abs(slq - exact)
    msg = 'surrogate err %s slq err %s ci %s' % (
        numpy.median(err_sur), numpy.median(err_slq), numpy.median(ci))
    # the surrogate alone is as accurate as one SLQ estimate ...
    assert numpy.median(err_sur) <= 2.0 * numpy.median(err_slq) + 1e-9, msg
    # ... what it returns after the SLQ judge lies within the SLQ interval of
    # the SLQ value, so within about two intervals of the exact value
    assert numpy.all(numpy.abs(y - slq) <= ci)
    assert numpy.median(numpy.abs(y - exact) / ci) <= 1.5, msg
    # ... and it answers thousands of queries in less time than SLQ takes
    # for a handful
    Q = 4096
    rng = numpy.random.default_rng(9)
    qs = numpy.exp(rng.uniform(numpy.log(0.01), numpy.log(0.1), Q))
    qt = numpy.exp(rng.uniform(numpy.log(1e-4), numpy.log(1e2), Q))
    sur.predict(qs[:16], qt[:16])
    t0 = time.perf_counter()
    sur.predict(qs, qt)
    t_sur = time.perf_counter() - t0
    _, _, t_slq = sur.solve(qs[:8], qt[:8])
    assert t_sur < t_slq, (t_sur, t_slq)
    # the trust rule keeps the surrogate for most of the box
    assert sur.trusted(qs, qt).mean() >= 0.5
