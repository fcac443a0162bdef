"""Exact logdet(K(l) + t I) for 12 held-out queries of the learned form
(n = 65536 CorrelationFamily with support radius 0.01 ln(1/0.03), seed 0;
queries sample_queries(10, 20, (0.01, 0.1), (1e-4, 1e2), seed=303), scale
indices 0..3, t indices 0, 10, 19), by dense FP64 blocked Cholesky on the
GPU. Needs ~70 GB of device memory.

    python tests/golden/make_surrogate_cholesky.py > tests/golden/surrogate_cholesky.json
"""

import json
import os
import sys
import time

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)
from imate_amd.sample_matrices import CorrelationFamily  # noqa: E402
from imate_amd.surrogate import sample_queries  # noqa: E402
from dense_cholesky import (dense_from_csr, cholesky_in_place,  # noqa: E402
                            logdet_from_factor)


def main():
    import torch
    n = 65536
    support = 0.01 * numpy.log(1.0 / 0.03)
    fam = CorrelationFamily(n, support, seed=0)
    ls, ts = sample_queries(10, 20, (0.01, 0.1), (1e-4, 1e2), seed=303)
    dev = torch.device('cuda:0')
    entries = []
    for si in range(4):
        A = fam.matrix(ls[si])
        for ti in (0, 10, 19):
            t0 = time.time()
            D = dense_from_csr(A, dev, shift=float(ts[si, ti]))
            cholesky_in_place(D)
            ld = logdet_from_factor(D)
            del D
            torch.cuda.empty_cache()
            entries.append({'scale_index': si, 't_index': ti,
                            'scale': float(ls[si]), 't': float(ts[si, ti]),
                            'logdet': ld})
            print('scale %.5f t %.3g logdet %.10g (%.1fs)'
                  % (ls[si], ts[si, ti], ld, time.time() - t0),
                  file=sys.stderr, flush=True)
    out = {'n': n, 'support': support, 'jitter': 1e-3, 'seed': 0,
           'queries_seed': 303, 'method': 'dense FP64 Cholesky on the GPU',
           'entries': entries}
    print(json.dumps(out, indent=1))


if __name__ == '__main__':
    main()
