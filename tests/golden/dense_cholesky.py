"""Dense FP64 blocked Cholesky on the GPU for the golden-value scripts.

torch is plumbing here (rocSOLVER / rocBLAS on panels of ``nb`` columns);
the unblocked library call rejects n = 65536. Right-looking: factor the
diagonal panel, solve the panel below it, update the trailing matrix.
"""


def dense_from_csr(A, dev, shift=0.0):
    """Dense device copy of the CSR matrix A + shift I (FP64)."""
    import torch
    n = A.shape[0]
    coo = A.tocoo()
    D = torch.zeros((n, n), dtype=torch.float64, device=dev)
    D[torch.from_numpy(coo.row).to(dev), torch.from_numpy(coo.col).to(dev)] = \
        torch.from_numpy(coo.data.astype('float64')).to(dev)
    if shift:
        D.diagonal().add_(shift)
    return D


def cholesky_in_place(D, nb=8192):
    """Overwrite the lower triangle of the SPD matrix D by its Cholesky
    factor L (D = L L^T); the strict upper triangle is left stale."""
    import torch
    n = D.shape[0]
    for k in range(0, n, nb):
        e = min(n, k + nb)
        Lkk = torch.linalg.cholesky(D[k:e, k:e])
        D[k:e, k:e] = Lkk
        if e < n:
            # L21 = A21 Lkk^-T, then the trailing update A22 -= L21 L21^T
            L21 = torch.linalg.solve_triangular(
                Lkk, D[e:, k:e].T.contiguous(), upper=False).T.contiguous()
            D[e:, k:e] = L21
            D[e:, e:].addmm_(L21, L21.T.contiguous(), alpha=-1.0)
            del L21
        del Lkk
    torch.cuda.synchronize()
    return D


def logdet_from_factor(D):
    import torch
    return float(2.0 * torch.log(torch.diagonal(D)).sum())
