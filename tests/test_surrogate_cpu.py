"""CPU checks of the learned form's host side: the fixed-pattern family and
the surrogate's target transform."""

import numpy

from imate_amd.sample_matrices import CorrelationFamily, \
    sparse_correlation_matrix
from imate_amd.surrogate import LogdetSurrogate, sample_queries, init_params


def test_family_member_equals_thresholded_construction():
    # with support = l ln(1/tau) the member is the tapered threshold matrix
    n, scale = 3000, 0.02
    fam = CorrelationFamily(n, scale * numpy.log(1 / 0.03), seed=3)
    A = fam.matrix(scale)
    B = sparse_correlation_matrix(n, scale=scale, seed=3)
    assert A.nnz == B.nnz and A.has_sorted_indices
    assert abs(A - B).max() < 1e-15


def test_family_members_share_pattern_and_are_spd():
    fam = CorrelationFamily(1500, 0.1, seed=4)
    prev = None
    for scale in (0.01, 0.03, 0.1):
        A = fam.matrix(scale)
        assert numpy.array_equal(A.indices, fam.indices)
        assert abs(A - A.T).max() == 0
        assert numpy.linalg.eigvalsh(A.toarray())[0] > 0
        assert numpy.allclose(A.diagonal(), 1.001)
        if prev is not None:
            assert abs(A - prev).max() > 0
        prev = A
    assert abs(fam.frobenius_squared(0.05) -
               (fam.matrix(0.05).toarray() ** 2).sum()) < 1e-9 * fam.nnz


def test_transform_round_trip_and_limit():
    sur = LogdetSurrogate.__new__(LogdetSurrogate)
    sur.n = 1000
    t = numpy.logspace(-4, 2, 7)
    y = numpy.linspace(-300.0, 5000.0, 7)
    assert numpy.allclose(sur._from_g(sur._to_g(y, t), t), y, rtol=1e-13)
    # logdet((1 + j) I + t I) = n log(1 + j + t): g -> j as t grows
    j = 1e-3
    g = sur._to_g(sur.n * numpy.log(1 + j + 1e8), 1e8)
    assert abs(g - j) < 1e-6


def test_sample_queries_and_inputs():
    ls, ts = sample_queries(5, 7, (0.01, 0.1), (1e-4, 1e2), seed=1)
    assert ls.shape == (5,) and ts.shape == (5, 7)
    assert ls.min() >= 0.01 and ls.max() <= 0.1
    assert numpy.all(numpy.diff(ts, axis=1) >= 0)
    sur = LogdetSurrogate.__new__(LogdetSurrogate)
    sur.scale_range, sur.t_range = (0.01, 0.1), (1e-4, 1e2)
    x = sur._inputs(numpy.array([0.01, 0.1]), numpy.array([1e-4, 1e2]))
    assert numpy.allclose(x, [[-1, -1], [1, 1]], atol=1e-6)
    assert init_params(0).shape == (17025,)


def test_stratified_queries_cover_every_slice():
    ls, ts = sample_queries(16, 10, (0.01, 0.1), (1e-4, 1e2), seed=3,
                            stratified=True)
    u = (numpy.log(ls) - numpy.log(0.01)) / (numpy.log(0.1) - numpy.log(0.01))
    assert sorted(numpy.floor(u * 16).astype(int)) == list(range(16))
    v = (numpy.log(ts) - numpy.log(1e-4)) / (numpy.log(1e2) - numpy.log(1e-4))
    for row in v:
        assert sorted(numpy.floor(row * 10).astype(int)) == list(range(10))


def test_family_coefficients_reproduce_values():
    """The device family's per-entry (d, w, c) give the host values by the
    kernel's formula w exp(-d / l) + c, diagonal included."""
    import numpy
    from imate_amd.sample_matrices import CorrelationFamily
    fam = CorrelationFamily(1500, support=0.07, seed=4)
    d, w, c = fam.coefficients()
    assert d.shape == w.shape == c.shape == (fam.nnz,)
    for scale in (0.01, 0.033, 0.1):
        v = w * numpy.exp(-d / scale) + c
        assert numpy.array_equal(v, fam.values(scale))
