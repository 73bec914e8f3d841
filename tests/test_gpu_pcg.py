"""MIC(0)-PCG on the GPU against the oracle's PCGSolver restatement.

The factor, the preconditioner application and the SpMV are required to be
bitwise equal to the sequential CPU form (PCGSolver.h:97-239,
SparseMatrix.h:298-322); the solve itself uses FP64 block-reduced dot
products, so its solution is compared to a tolerance.
"""
import numpy as np
import pytest

from tests.helpers import csr_parts, dam_break_system, grid9_mmatrix

pytestmark = pytest.mark.gpu

TOL_SOLUTION = 1e-9  # relative, between two 1e-8 relative-residual solves


@pytest.fixture(scope="module")
def systems(oracle):
    out = {"dam16": dam_break_system(oracle, 16)}
    A = grid9_mmatrix(23, 17, seed=5)
    rs, col, val = csr_parts(A)
    out["grid9"] = (rs, col, val, np.random.default_rng(1).standard_normal(len(rs) - 1))
    A = grid9_mmatrix(19, 13, seed=7, zero_rows=6)
    rs, col, val = csr_parts(A)
    b = np.random.default_rng(2).standard_normal(len(rs) - 1)
    # keep the rhs in the range of the matrix
    b[np.diff(rs) == 0] = 0.0
    out["grid9_null_rows"] = (rs, col, val, b)
    return out


@pytest.mark.parametrize("name", ["dam16", "grid9", "grid9_null_rows"])
def test_spmv_bitwise(oracle, wc, systems, name):
    rs, col, val, b = systems[name]
    x = np.random.default_rng(3).standard_normal(len(rs) - 1)
    y_gpu = wc.spmv_csr(rs, col, val, x)
    y_cpu = oracle.spmv(rs, col, val, x)
    assert np.array_equal(y_gpu, y_cpu)


@pytest.mark.parametrize("name", ["dam16", "grid9", "grid9_null_rows"])
@pytest.mark.parametrize("tau,mdr", [(0.97, 0.25), (0.0, 0.25), (1.0, 0.0)])
def test_mic0_factor_bitwise(oracle, wc, systems, name, tau, mdr):
    rs, col, val, _ = systems[name]
    inv_gpu, lv_gpu = wc.mic0_factor_csr(rs, col, val, tau, mdr)
    inv_cpu, lv_cpu = oracle.mic0(rs, col, val, tau, mdr)
    upper = np.zeros(len(col), dtype=bool)
    for i in range(len(rs) - 1):
        upper[rs[i]:rs[i + 1]] = col[rs[i]:rs[i + 1]] > i
    assert np.array_equal(inv_gpu, inv_cpu)
    assert np.array_equal(lv_gpu[upper], lv_cpu[upper])


@pytest.mark.parametrize("name", ["dam16", "grid9", "grid9_null_rows"])
def test_mic0_apply_bitwise(oracle, wc, systems, name):
    rs, col, val, _ = systems[name]
    inv, lv = oracle.mic0(rs, col, val)
    r = np.random.default_rng(4).standard_normal(len(rs) - 1)
    z_gpu = wc.mic0_apply_csr(rs, col, inv, lv, r)
    z_cpu = oracle.apply_precond(rs, col, inv, lv, r)
    assert np.array_equal(z_gpu, z_cpu)


@pytest.mark.parametrize("name", ["dam16", "grid9", "grid9_null_rows"])
def test_pcg_solve_matches_oracle(oracle, wc, systems, name):
    rs, col, val, b = systems[name]
    ok_g, x_g, it_g, res_g = wc.pcg_solve_csr(rs, col, val, b, tol=1e-8, maxit=1000)
    ok_c, x_c, it_c, res_c = oracle.pcg(rs, col, val, b, tol=1e-8, maxit=1000)
    assert ok_g and ok_c
    assert abs(it_g - it_c) <= 1
    rel = np.abs(x_g - x_c).max() / np.abs(x_c).max()
    assert rel < TOL_SOLUTION, rel
    r = b - oracle.spmv(rs, col, val, x_g)
    assert np.abs(r).max() <= 1e-8 * np.abs(b).max() * 1.0000001


def test_pcg_edge_cases(oracle, wc, systems):
    rs, col, val, b = systems["grid9"]
    # zero rhs: success after 0 iterations (PCGSolver.h:271-275)
    ok, x, it, res = wc.pcg_solve_csr(rs, col, val, np.zeros_like(b))
    assert ok and it == 0 and res == 0.0 and not x.any()
    # maxit = 1: not converged, like the oracle
    ok, x, it, res = wc.pcg_solve_csr(rs, col, val, b, tol=1e-12, maxit=1)
    okc, xc, itc, resc = oracle.pcg(rs, col, val, b, tol=1e-12, maxit=1)
    assert not ok and not okc and it == itc == 1
    # NaN rhs: rho is NaN -> failure with 0 iterations (PCGSolver.h:281-284)
    bn = b.copy()
    bn[3] = np.nan
    ok, x, it, res = wc.pcg_solve_csr(rs, col, val, bn)
    assert not ok and it == 0
    # empty system
    ok, x, it, res = wc.pcg_solve_csr(np.zeros(1, np.uint32), np.zeros(0, np.uint32),
                                      np.zeros(0), np.zeros(0))
    assert ok and it == 0


def test_csr_validation_rejects_unsorted(wc):
    rs = np.array([0, 2, 3], np.uint32)
    col = np.array([1, 0, 1], np.uint32)  # row 0 unsorted
    with pytest.raises(wc.WetclothError):
        wc.spmv_csr(rs, col, np.ones(3), np.ones(2))
