"""Device value families (slq_op_family_attach / slq_op_family_values): the
values made on the device against the family's host values, and SLQ on the
refreshed handle against the oracle on the same values."""

import numpy
import pytest
import scipy.sparse

import imate_amd
from imate_amd import _lib
from imate_amd.sample_matrices import CorrelationFamily
from oracle import oracle

pytestmark = pytest.mark.gpu


def _device_values(op):
    A = op.A
    indptr = numpy.empty(A.shape[0] + 1, dtype=numpy.int64)
    indices = numpy.empty(A.nnz, dtype=numpy.int32)
    data = numpy.empty(A.nnz, dtype=A.dtype)
    _lib.check(_lib.load().slq_op_get_csr(op._A_dev.handle, _lib.ptr(indptr),
                                          _lib.ptr(indices), _lib.ptr(data)),
               'slq_op_get_csr')
    assert numpy.array_equal(indices, A.indices)
    return data


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_family_values_and_slq(dtype):
    fam = CorrelationFamily(3000, support=0.06, seed=2, dtype=dtype)
    op = imate_amd.AffineMatrixFunction(fam.matrix(0.02))
    op.attach_family(*fam.coefficients())
    ts = [1e-3, 0.3, 4.0]
    for scale in (0.05, 0.012, 0.02):
        op.set_family_scale(scale)
        got = _device_values(op)
        want = fam.values(scale)
        # one exp and one product per entry: a few ulp of the value type
        tol = 8 * numpy.finfo(numpy.dtype(dtype)).eps
        assert numpy.all(numpy.abs(got - want) <= tol * numpy.abs(want))
        vals, info = imate_amd.logdet(op, parameters=ts, method='slq',
                                      gpu=True, lanczos_degree=20,
                                      min_num_samples=64, max_num_samples=64,
                                      seed=3, return_info=True)
        K = scipy.sparse.csr_matrix((got, fam.indices, fam.indptr),
                                    shape=(fam.n, fam.n))
        ref = oracle.slq(K, func='log', parameters=ts, affine=True,
                         lanczos_degree=20, min_num_samples=64,
                         max_num_samples=64, seed=3)
        rel = numpy.abs(info['convergence']['samples'] - ref['samples']) / \
            numpy.abs(ref['samples'])
        assert rel.max() < (1e-10 if dtype == 'float64' else 1e-4)


def test_family_errors(gpu_lib):
    fam = CorrelationFamily(500, support=0.1, seed=1)
    op = imate_amd.AffineMatrixFunction(fam.matrix(0.02))
    op.initialize(True, 0)
    # no family attached yet / non-positive scale
    assert gpu_lib.slq_op_family_values(op._A_dev.handle, 0.05) == \
        _lib.SLQ_E_ARG
    op.attach_family(*fam.coefficients())
    assert gpu_lib.slq_op_family_values(op._A_dev.handle, -1.0) == \
        _lib.SLQ_E_ARG
    # the affine handle itself takes no family
    assert gpu_lib.slq_op_family_values(op.gpu_Aop.handle, 0.05) == \
        _lib.SLQ_E_ARG
