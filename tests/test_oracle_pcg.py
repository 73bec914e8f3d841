"""Pin the oracle's PCGSolver restatement (PCGSolver.h:97-330).

- tau = 0 on a tridiagonal SPD matrix: IC(0) has no dropped fill, so it is
  the exact Cholesky factor (closed form: L L^T = A, M^-1 r = A^-1 r).
- tau = 1, mdr = 0 (fully modified): L L^T preserves row sums, (L L^T) 1 = A 1.
- A dense brute-force factorization written from the MATLAB recurrence the
  reference quotes (PCGSolver.h:124-138), with the min-diagonal guard, on
  random M-matrices whose pattern has fill (9-point) -- an independent
  formulation, so a wrong index or dropped term in the sparse walk fails.
- The solve against scipy.sparse.linalg.spsolve, and the early-exit cases.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from tests.helpers import csr_parts, dam_break_system, grid9_mmatrix


def dense_mic0(A, tau, mdr):
    """MATLAB recurrence of PCGSolver.h:124-138 on a dense lower triangle,
    plus the null-column skip (:141) and the min-diagonal guard (:143-148)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    pattern = A != 0
    L = np.tril(A).copy()
    adiag = np.diag(A).copy()
    for k in range(n):
        if adiag[k] == 0:
            continue
        if L[k, k] < mdr * adiag[k]:
            L[k, k] = np.sqrt(adiag[k])
        else:
            L[k, k] = np.sqrt(L[k, k])
        L[k + 1:, k] /= L[k, k]
        for j in np.nonzero(L[:, k])[0]:
            if j <= k:
                continue
            full = L[:, k] * L[j, k]
            inc = full * pattern[:, j]
            missing = full.sum() - full[j] - inc.sum() + inc[j]
            L[j:, j] -= inc[j:]
            L[j, j] -= tau * missing
    return L


def split_factor(rs, col, invdiag, lv):
    n = len(rs) - 1
    L = np.zeros((n, n))
    for i in range(n):
        L[i, i] = 1.0 / invdiag[i] if invdiag[i] != 0 else 0.0
        for e in range(rs[i], rs[i + 1]):
            if col[e] > i:
                L[col[e], i] = lv[e]
    return L


def tridiag(n, seed):
    rng = np.random.default_rng(seed)
    off = -rng.uniform(0.2, 1.0, n - 1)
    d = np.zeros(n)
    d[:-1] -= off
    d[1:] -= off
    d += rng.uniform(0.01, 0.1, n)
    return sp.diags([off, d, off], [-1, 0, 1]).tocsr()


def test_ic0_on_tridiagonal_is_exact_cholesky(oracle):
    A = tridiag(40, 0)
    rs, col, val = csr_parts(A)
    inv, lv = oracle.mic0(rs, col, val, tau=0.0, mdr=0.0)
    L = split_factor(rs, col, inv, lv)
    Ad = A.toarray()
    assert np.abs(L @ L.T - Ad).max() < 1e-13 * np.abs(Ad).max()
    assert np.allclose(L, np.linalg.cholesky(Ad), rtol=1e-13, atol=1e-15)
    r = np.random.default_rng(1).standard_normal(40)
    z = oracle.apply_precond(rs, col, inv, lv, r)
    assert np.abs(z - np.linalg.solve(Ad, r)).max() < 1e-10 * np.abs(z).max()


def test_full_mic_preserves_row_sums(oracle):
    A = grid9_mmatrix(9, 7, seed=3)
    rs, col, val = csr_parts(A)
    inv, lv = oracle.mic0(rs, col, val, tau=1.0, mdr=0.0)
    L = split_factor(rs, col, inv, lv)
    one = np.ones(A.shape[0])
    assert np.abs(L @ (L.T @ one) - A @ one).max() < 1e-12 * np.abs(A).max()


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("tau,mdr", [(0.97, 0.25), (0.0, 0.25), (1.0, 0.0), (0.5, 0.9)])
def test_sparse_walk_matches_dense_recurrence(oracle, seed, tau, mdr):
    A = grid9_mmatrix(7, 6, seed=seed, zero_rows=2 if seed == 2 else 0)
    rs, col, val = csr_parts(A)
    inv, lv = oracle.mic0(rs, col, val, tau=tau, mdr=mdr)
    L = split_factor(rs, col, inv, lv)
    Ld = dense_mic0(A.toarray(), tau, mdr)
    assert np.abs(L - Ld).max() < 1e-13 * np.abs(Ld).max()


def test_dam_break_system_properties(oracle):
    """The assembled pressure matrix is a symmetric M-matrix and MIC(0) of a
    7-point pattern has no fill (every off-diagonal of L is the matrix entry
    scaled by the column pivot)."""
    rs, col, val, b = dam_break_system(oracle, 12)
    n = len(rs) - 1
    A = sp.csr_matrix((val, col, rs), shape=(n, n))
    assert abs(A - A.T).max() == 0.0
    off = A - sp.diags(A.diagonal())
    assert off.data.max() <= 0.0 and A.diagonal().min() >= 0.0
    inv, lv = oracle.mic0(rs, col, val)
    for i in range(n):
        for e in range(rs[i], rs[i + 1]):
            if col[e] > i:
                assert lv[e] == val[e] * inv[i]


def test_pcg_matches_spsolve(oracle):
    rs, col, val, b = dam_break_system(oracle, 16)
    n = len(rs) - 1
    A = sp.csr_matrix((val, col, rs), shape=(n, n))
    ok, x, it, res = oracle.pcg(rs, col, val, b, tol=1e-10, maxit=2000)
    assert ok and res <= 1e-10 * np.abs(b).max()
    # rows that are identically zero (wall ghost cells away from liquid) get 0
    nz = np.diff(rs) > 0
    nonnull = np.array([A[i].count_nonzero() > 0 and A[i, i] != 0 for i in range(n)])
    Ar = A[nonnull][:, nonnull]
    xs = spla.spsolve(Ar.tocsc(), b[nonnull])
    assert np.abs(x[nonnull] - xs).max() < 1e-7 * np.abs(xs).max()
    assert nz.all()


def test_pcg_early_exits(oracle):
    A = tridiag(20, 4)
    rs, col, val = csr_parts(A)
    b = np.random.default_rng(5).standard_normal(20)
    ok, x, it, res = oracle.pcg(rs, col, val, np.zeros(20))
    assert ok and it == 0 and not x.any()
    # MIC(0) of a tridiagonal drops no fill for any tau: exact, one iteration
    ok, x, it, res = oracle.pcg(rs, col, val, b, tol=1e-12, maxit=1)
    assert ok and it == 1
    A9 = grid9_mmatrix(8, 8, seed=8)
    rs9, col9, val9 = csr_parts(A9)
    ok, x, it, res = oracle.pcg(rs9, col9, val9, np.ones(64), tol=1e-14, maxit=1)
    assert not ok and it == 1
    bn = b.copy()
    bn[0] = np.nan
    ok, x, it, res = oracle.pcg(rs, col, val, bn)
    assert not ok and it == 0


def test_spmv_matches_scipy(oracle):
    A = grid9_mmatrix(11, 9, seed=6)
    rs, col, val = csr_parts(A)
    x = np.random.default_rng(7).standard_normal(A.shape[0])
    y = oracle.spmv(rs, col, val, x)
    assert np.abs(y - A @ x).max() < 1e-14 * np.abs(A).sum(axis=1).max() * np.abs(x).max()
