"""Test-side FP64 operator of the 9-point matrix as nine diagonals.

M = A^T, where A is the reference's row-major triplet matrix (irow = grid row
k, icol = neighbour column): M x has y[icol] += v x[irow] over the triplets,
i.e. y[r] = sum_d D_d[r] x[r + d] for the nine offsets d = irow - icol.
Applied with elementwise products of whole panels (numpy on the host, torch
on the GPU), so the backward error of every shot of a full-size workload is
checked by arithmetic independent of the library's residual kernels.
"""
import numpy as np


def diagonals(n, irow, icol, ar, ai, transpose=False):
    """{offset: values [n]} of M (transpose=False) or M^T (= A)."""
    r, c = (np.asarray(icol), np.asarray(irow)) if not transpose else (np.asarray(irow), np.asarray(icol))
    d = c - r
    out = {}
    for o in np.unique(d):
        v = np.zeros(n, complex)
        sel = d == o
        v[r[sel]] = np.asarray(ar)[sel] + 1j * np.asarray(ai)[sel]
        out[int(o)] = v
    return out


def to_device(D, torch):
    return {o: torch.as_tensor(v, device="cuda") for o, v in D.items()}


def apply(D, X):
    """Y = op X for a panel X [nrhs, n] (numpy or torch), D from diagonals()."""
    Y = X * 0
    n = X.shape[-1]
    for o, v in D.items():
        if o >= 0:
            Y[..., :n - o] += v[:n - o] * X[..., o:]
        else:
            Y[..., -o:] += v[-o:] * X[..., :n + o]
    return Y


def relres(D, X, B):
    """Per-column ||b - op x||_2 / ||b||_2 (numpy or torch panels [nrhs, n])."""
    R = B - apply(D, X)
    return (abs(R) ** 2).sum(-1) ** 0.5 / (abs(B) ** 2).sum(-1) ** 0.5


def backward_error(D, X, B):
    """Per-column normwise backward error ||r||_inf / (||op||_inf ||x||_inf + ||b||_inf)."""
    R = B - apply(D, X)
    n = X.shape[-1]
    rowsum = None
    for o, v in D.items():
        a = abs(v)
        rowsum = a if rowsum is None else rowsum + a
    mnorm = float(rowsum.max())
    return abs(R).max(-1)[0] / (mnorm * abs(X).max(-1)[0] + abs(B).max(-1)[0]) if hasattr(R, "device") else \
        abs(R).max(-1) / (mnorm * abs(X).max(-1) + abs(B).max(-1))
