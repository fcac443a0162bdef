"""NumPy statement of the block Lanczos form (``krylov='block'``), the
checker of ``slq_estimate_block_krylov``. Test infrastructure only.

One block of b probes Z (n x b): Z = Q_0 R_0 by Cholesky QR twice; for
j = 0 .. m-1: W = A Q_j, H = Q^T W over all stored blocks, A_j = (H_j +
H_j^T) / 2, W -= Q H, then the same projection again, and W = Q_{j+1} B_j
(Cholesky QR twice) except after the last step. T is block tridiagonal (A_j
on the diagonal, B_j below, B_j^T above); with T = Y diag(theta) Y^T,
sample_p = sum_k f(theta_k^p) (R_0^T Y[:b, k])_p^2, shifted by t_j - t_0 for
the further t of A + tI.

Pinned in tests/test_block_lanczos_ref.py: with b = 1 it equals the
oracle's Lanczos with full re-orthogonalisation sample for sample; f(x) = x^2
is integrated exactly for m >= 2; on a matrix whose size is b m the Krylov
space is the whole space and every f is integrated exactly.
"""

import numpy


def _cholqr2(X):
    R1 = numpy.linalg.cholesky(X.T @ X).T
    Y = numpy.linalg.solve(R1.T, X.T).T
    R2 = numpy.linalg.cholesky(Y.T @ Y).T
    Q = numpy.linalg.solve(R2.T, Y.T).T
    return Q, R2 @ R1


def block_tridiagonal(A, Z, m, t=0.0):
    """(T, R_0) of m block Lanczos steps on A + tI from the probe block Z."""
    n, b = Z.shape
    Q0, R0 = _cholqr2(Z)
    Qs = [Q0]
    As, Bs = [], []
    for j in range(m):
        W = A @ Qs[j] + t * Qs[j]
        Qm = numpy.hstack(Qs)
        H = Qm.T @ W
        Hj = H[j * b:(j + 1) * b]
        As.append(0.5 * (Hj + Hj.T))
        W = W - Qm @ H
        W = W - Qm @ (Qm.T @ W)
        if j + 1 < m:
            Qn, Bj = _cholqr2(W)
            Qs.append(Qn)
            Bs.append(Bj)
    N = m * b
    T = numpy.zeros((N, N))
    for j in range(m):
        T[j * b:(j + 1) * b, j * b:(j + 1) * b] = As[j]
        if j + 1 < m:
            T[(j + 1) * b:(j + 2) * b, j * b:(j + 1) * b] = Bs[j]
            T[j * b:(j + 1) * b, (j + 1) * b:(j + 2) * b] = Bs[j].T
    return T, R0


def samples(A, Z, m, f, exponent=1.0, params=None):
    """Per-probe samples of the block (rows: probes; columns: inquiries).
    ``params``: t values of A + tI (the first one runs the recurrence, the
    others are Ritz shifts), or None for A itself."""
    ts = [0.0] if params is None else list(params)
    T, R0 = block_tridiagonal(A, Z, m, ts[0])
    theta, Y = numpy.linalg.eigh(T)
    b = Z.shape[1]
    U = R0.T @ Y[:b, :]
    out = numpy.empty((b, len(ts)))
    for j, tj in enumerate(ts):
        th = theta - ts[0] + tj
        out[:, j] = (U ** 2) @ f(th ** exponent)
    return out
