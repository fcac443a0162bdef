"""Pins of the NumPy block Lanczos checker (tests/block_lanczos_ref.py)
against the oracle and the mathematics, on the CPU."""

import numpy
import pytest

from imate_amd.sample_matrices import sparse_correlation_matrix
from oracle import oracle
from tests.block_lanczos_ref import samples


def _spd(n, seed):
    A = sparse_correlation_matrix(n, nnz_per_row=20, seed=seed).toarray()
    return A


def test_one_probe_block_is_lanczos_with_full_reorthogonalisation():
    n, m = 500, 15
    A = sparse_correlation_matrix(n, nnz_per_row=20, seed=3)
    Z = oracle.rademacher(4, n, 4)
    ref = oracle.slq(A, func='log', lanczos_degree=m, min_num_samples=4,
                     max_num_samples=4, seed=4, orthogonalize=-1)
    Ad = A.toarray()
    got = [samples(Ad, Z[i][:, None], m, numpy.log)[0, 0] for i in range(4)]
    assert numpy.allclose(got, ref['samples'][:, 0], rtol=1e-10)


def test_square_is_exact_from_two_block_steps():
    A = _spd(300, 1)
    Z = oracle.rademacher(2, 300, 16).T
    got = samples(A, Z, 2, lambda x: x, exponent=2.0)[:, 0]
    want = numpy.einsum('ip,ip->p', Z, A @ (A @ Z))
    assert numpy.allclose(got, want, rtol=1e-12)


@pytest.mark.parametrize('f', [numpy.log, lambda x: 1.0 / x])
def test_whole_space_is_exact_for_any_function(f):
    b, m = 8, 5
    A = _spd(b * m, 2)
    Z = oracle.rademacher(5, b * m, b).T
    lam, V = numpy.linalg.eigh(A)
    FA = (V * f(lam)) @ V.T
    want = numpy.einsum('ip,ip->p', Z, FA @ Z)
    got = samples(A, Z, m, f)[:, 0]
    assert numpy.allclose(got, want, rtol=1e-9)


def test_ritz_shift_equals_shifted_run():
    A = _spd(400, 4)
    Z = oracle.rademacher(6, 400, 8).T
    ts = [0.1, 0.7, 2.0]
    got = samples(A, Z, 6, numpy.log, params=ts)
    for j, t in enumerate(ts):
        direct = samples(A + t * numpy.eye(400), Z, 6, numpy.log)[:, 0]
        assert numpy.allclose(got[:, j], direct, rtol=1e-9)
