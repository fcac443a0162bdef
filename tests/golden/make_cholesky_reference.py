"""Writes tests/golden/cholesky_logdet.json: exact log-determinants of the
benchmark matrices (imate_amd.sample_matrices.sparse_correlation_matrix),
from a sparse LDL^T factorisation (SciPy SuperLU in symmetric mode with a
minimum-degree ordering on A^T + A and diagonal pivots), logdet = sum log d_ii.
This is the oracle the reference uses for logdet (imate/logdet/
_cholesky_method.py:262-330 with sparse_cholesky.py:23-66). The n = 2^20 case
takes about 12-34 minutes and 15-23 GB of memory.

    python tests/golden/make_cholesky_reference.py [n ...]
"""

import json
import os
import sys
import time

import numpy
import scipy.sparse.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from imate_amd.sample_matrices import sparse_correlation_matrix  # noqa: E402

PATH = os.path.join(HERE, 'cholesky_logdet.json')


def logdet(A):
    lu = scipy.sparse.linalg.splu(A.tocsc(), permc_spec='MMD_AT_PLUS_A',
                                  diag_pivot_thresh=0.0,
                                  options=dict(SymmetricMode=True))
    d = lu.U.diagonal()
    if (d <= 0).any():
        raise ValueError('matrix is not positive definite')
    return float(numpy.sum(numpy.log(d)))


def main(sizes):
    table = {'entries': []}
    if os.path.exists(PATH):
        with open(PATH) as f:
            table = json.load(f)
    for n in sizes:
        t = time.time()
        A = sparse_correlation_matrix(n, nnz_per_row=50, seed=0)
        ld = logdet(A)
        row = {'n': n, 'nnz_per_row': 50.0, 'seed': 0, 'nnz': int(A.nnz),
               'logdet': ld, 'seconds': round(time.time() - t, 1)}
        table['entries'] = [r for r in table['entries']
                            if not (r['n'] == n and r['seed'] == 0 and
                                    r['nnz_per_row'] == 50.0)] + [row]
        print(row, flush=True)
    with open(PATH, 'w') as f:
        json.dump(table, f, indent=1)


if __name__ == '__main__':
    main([int(a) for a in sys.argv[1:]] or [2**16, 2**18])
