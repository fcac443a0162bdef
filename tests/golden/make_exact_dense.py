"""Exact spectral functions of the precision-form matrix (n = 65536,
kernel scale 0.01, imate_amd.sample_matrices.sparse_correlation_matrix with
seed 0), by a dense FP64 blocked Cholesky factorisation on the GPU (torch is
used as plumbing for rocSOLVER / rocBLAS on 8192-wide panels; the unblocked
library call rejects n = 65536):

    logdet(A)  = 2 sum log diag(L)
    tr(A^-1)   = |L^-1|_F^2
    tr(A^2)    = sum of squared entries (exact, from the sparse matrix)

Writes the values as JSON to stdout; the committed copy is
tests/golden/exact_precision_form.json. Needs ~110 GB of device memory.

    python tests/golden/make_exact_dense.py > tests/golden/exact_precision_form.json
"""

import json
import os
import sys
import time

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from imate_amd.sample_matrices import sparse_correlation_matrix  # noqa: E402
sys.path.insert(0, HERE)
from dense_cholesky import (dense_from_csr, cholesky_in_place,  # noqa: E402
                            logdet_from_factor)


def main():
    import torch
    n = 65536
    t0 = time.time()
    A = sparse_correlation_matrix(n, scale=0.01, seed=0)
    tr2 = float((A.data.astype(numpy.float64)**2).sum())
    print('built %.1fs nnz/row %.1f' % (time.time() - t0, A.nnz / n),
          file=sys.stderr, flush=True)
    dev = torch.device('cuda:0')
    D = dense_from_csr(A, dev)
    t0 = time.time()
    nb = 8192
    cholesky_in_place(D, nb)
    print('cholesky %.1fs' % (time.time() - t0), file=sys.stderr, flush=True)
    logdet = logdet_from_factor(D)
    # tr(A^-1) = |L^-1|_F^2 by blocked forward substitution, one block of
    # identity columns at a time (rows above the block are zero)
    t0 = time.time()
    traceinv = 0.0
    for c in range(0, n, nb):
        ce = min(n, c + nb)
        Y = []
        for i in range(c, n, nb):
            ie = min(n, i + nb)
            rhs = torch.zeros((ie - i, ce - c), dtype=torch.float64, device=dev)
            if i == c:
                rhs += torch.eye(ce - c, dtype=torch.float64, device=dev)
            for jdx, j in enumerate(range(c, i, nb)):
                rhs -= D[i:ie, j:min(n, j + nb)] @ Y[jdx]
            Yi = torch.linalg.solve_triangular(D[i:ie, i:ie].contiguous(), rhs,
                                               upper=False)
            Y.append(Yi)
            traceinv += float((Yi * Yi).sum())
        del Y
        print('  column block %d done %.1fs' % (c // nb, time.time() - t0),
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    print('inverse %.1fs' % (time.time() - t0), file=sys.stderr, flush=True)
    out = {'n': n, 'scale': 0.01, 'threshold': 0.03, 'jitter': 1e-3, 'seed': 0,
           'nnz': int(A.nnz), 'logdet': logdet, 'traceinv': traceinv,
           'trace_A2': tr2, 'schatten2': (tr2 / n) ** 0.5,
           'method': 'dense FP64 Cholesky on the GPU'}
    print(json.dumps(out, indent=1))


if __name__ == '__main__':
    main()
