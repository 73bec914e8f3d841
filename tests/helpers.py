"""Shared test inputs: CSR helpers, M-matrices with fill, oracle systems."""
import numpy as np
import scipy.sparse as sp


def csr_parts(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.uint32), A.indices.astype(np.uint32), A.data.astype(np.float64)


def grid9_mmatrix(nx, ny, seed, zero_rows=0):
    """Symmetric M-matrix of a 2-D 9-point stencil: its lower neighbours share
    columns, so the factor's fill-matching branch (PCGSolver.h:184-195) runs."""
    rng = np.random.default_rng(seed)
    n = nx * ny
    rows, cols, vals = [], [], []
    for j in range(ny):
        for i in range(nx):
            a = j * nx + i
            for dj, di in ((0, 1), (1, -1), (1, 0), (1, 1)):
                jj, ii = j + dj, i + di
                if 0 <= ii < nx and 0 <= jj < ny:
                    b = jj * nx + ii
                    w = rng.uniform(0.1, 1.0)
                    rows += [a, b]
                    cols += [b, a]
                    vals += [-w, -w]
    W = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    d = -np.asarray(W.sum(axis=1)).ravel() + rng.uniform(0.0, 0.05, n)
    A = (W + sp.diags(d)).tolil()
    for r in rng.choice(n, size=zero_rows, replace=False):
        A[r, :] = 0.0
        A[:, r] = 0.0
    A = A.tocsr()
    A.eliminate_zeros()
    return A


def dam_break_system(oracle, n=16):
    g, x, v = oracle.dam_break(n)
    s = oracle.OracleScene((n, n, n), 1.0 / n, x, v)
    s.pre_solve(1.0 / 480)
    rs, col, val, b, _ = s.system()
    return rs, col, val, b


def reference_bucket_frame(x, dx):
    """The reference's bucket grid around particles x (3 x N): the world cell
    of its first cell and its node counts per axis (buckets * 4).
    updateParticleBoundingBox rounds the bounding box out to multiples of two
    buckets (TwoDScene.cpp:1606-1611); rebucketizeParticles pads it by three
    buckets and counts the buckets (:1660-1690)."""
    B = 4 * dx
    D = 2 * B
    bmin = np.floor(x.min(axis=1) / D) * D
    bmax = np.ceil(x.max(axis=1) / D) * D
    cells = np.maximum(1, np.ceil((bmax - bmin + B * 3.0 * 2.0) / dx).astype(int))
    nbk = np.maximum(1, np.ceil(cells * dx / B).astype(int))
    origin = np.rint((bmin - B * 3.0) / dx).astype(int)
    return origin, tuple(int(4 * b) for b in nbk)
