"""The oracle's AMG-PCG restatement (oracle/wc_oracle_amg.c), pinned on the
CPU against what the mathematics fixes:
  * the coarse operator equals the dense Galerkin product 0.5 R A P with
    P the 2x2x2 aggregation (1 for a DOF's coarse cell) and R = 0.125 P^T,
    and coarse DOFs are numbered in order of first appearance;
  * many red-black Gauss-Seidel sweeps converge to the direct solution;
  * one V-cycle is a contraction on the error (A-norm), and the PCG with it
    reaches scipy's direct solution in far fewer iterations than MIC(0).
Systems: the oracle's own dam-break pressure systems, DOFs at their dense
node coordinates."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def dam_system(oracle, n, steps=1):
    g, x, v = oracle.dam_break(n, seed=2, sigma=0.5)
    s = oracle.OracleScene((n, n, n), 1.0 / n, x, v)
    for _ in range(steps):
        s.substep(1.0 / 480, 1e-8, 1000)
    s.pre_solve(1.0 / 480)
    rs, col, val, b, _ = s.system()
    dof = s.field("dof")
    nx = tuple(int(t) for t in s.grid.nx)
    ii, jj, kk = np.meshgrid(np.arange(nx[0]), np.arange(nx[1]), np.arange(nx[2]),
                             indexing="ij")
    nodes = oracle.node_index_array(s.grid, ii.ravel(), jj.ravel(), kk.ravel())
    d = dof[nodes]
    m = d >= 0
    ijk = np.zeros((len(rs) - 1, 3), dtype=np.int32)
    ijk[d[m]] = np.stack([ii.ravel()[m], jj.ravel()[m], kk.ravel()[m]], axis=1)
    return rs, col, val, b, ijk, nx


def csr(rs, col, val):
    n = len(rs) - 1
    return sp.csr_matrix((val, col.astype(np.int64), rs.astype(np.int64)), shape=(n, n))


def test_coarse_operator_is_the_galerkin_product(oracle):
    rs, col, val, b, ijk, nx = dam_system(oracle, 24)
    parent, crs, ccol, cval = oracle.amg_coarse(rs, col, val, ijk, nx)
    n = len(rs) - 1
    nc = len(crs) - 1
    # the same coarse cell <=> the same coarse DOF, numbered by first appearance
    cells = {}
    seen = 0
    for f in range(n):
        key = tuple(ijk[f] // 2)
        if key not in cells:
            cells[key] = seen
            seen += 1
        assert parent[f] == cells[key]
    assert seen == nc
    P = sp.csr_matrix((np.ones(n), (np.arange(n), parent.astype(np.int64))), shape=(n, nc))
    R = 0.125 * P.T
    Ac = (0.5 * (R @ csr(rs, col, val) @ P)).toarray()
    got = csr(crs, ccol, cval).toarray()
    scale = np.abs(Ac).max()
    assert np.abs(got - Ac).max() <= 1e-13 * scale
    # rows are sorted and the operator stays symmetric
    for I in range(nc):
        c = ccol[crs[I]:crs[I + 1]]
        assert np.all(np.diff(c.astype(np.int64)) > 0)
    assert np.abs(got - got.T).max() <= 1e-13 * scale


def direct(A, b):
    """spsolve on the rows that are not identically zero (wall ghost cells away
    from the liquid hold all-zero rows and b = 0; the solvers leave them 0)"""
    nonnull = np.abs(A).sum(axis=1).A1 > 0
    assert not b[~nonnull].any()
    xs = np.zeros(len(b))
    xs[nonnull] = spla.spsolve(A[nonnull][:, nonnull].tocsc(), b[nonnull])
    return xs


def test_rbgs_converges_to_the_solution(oracle):
    rs, col, val, b, ijk, nx = dam_system(oracle, 12)
    A = csr(rs, col, val)
    xs = direct(A, b)
    x = oracle.amg_rbgs(rs, col, val, ijk, b, np.zeros(len(b)), 3000)
    assert np.abs(x - xs).max() <= 1e-8 * np.abs(xs).max()
    # one sweep on a planted error reduces it in the A-norm
    e0 = np.random.default_rng(1).standard_normal(len(b))
    e1 = oracle.amg_rbgs(rs, col, val, ijk, np.zeros(len(b)), e0, 1)
    assert e1 @ (A @ e1) < e0 @ (A @ e0)


def test_vcycle_contracts_the_error(oracle):
    rs, col, val, b, ijk, nx = dam_system(oracle, 32)
    sizes = oracle.amg_levels(rs, col, val, ijk, nx)
    assert len(sizes) >= 2 and sizes[0] > 4096 and sizes[-1] <= 4096
    assert all(sizes[l + 1] < sizes[l] for l in range(len(sizes) - 1))
    A = csr(rs, col, val)
    xt = np.random.default_rng(5).standard_normal(len(b))
    z = oracle.amg_vcycle(rs, col, val, ijk, nx, A @ xt)
    e = xt - z
    assert e @ (A @ e) < 0.5 * (xt @ (A @ xt))


def test_amgpcg_matches_the_direct_solve(oracle):
    rs, col, val, b, ijk, nx = dam_system(oracle, 32)
    A = csr(rs, col, val)
    ok, x, it, res = oracle.amgpcg(rs, col, val, ijk, nx, b, tol=1e-10, maxit=200)
    assert ok
    assert res <= 1e-10 * np.abs(b).max()
    assert np.abs(A @ x - b).max() <= 1.0001e-10 * np.abs(b).max()
    xs = direct(A, b)
    assert np.abs(x - xs).max() <= 1e-6 * np.abs(xs).max()
    ok2, x2, it2, res2 = oracle.pcg(rs, col, val, b, tol=1e-10, maxit=2000)
    assert ok2 and it < it2
    # a zero right-hand side returns at once, like the reference
    ok0, x0, it0, res0 = oracle.amgpcg(rs, col, val, ijk, nx, np.zeros(len(b)))
    assert ok0 and it0 == 0 and not x0.any()
