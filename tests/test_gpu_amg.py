"""AMGPCGSolveSparse on the GPU (wc_amgpcg_solve_csr, wc_amg_apply_csr)
against the oracle's restatement (oracle/wc_oracle_amg.c, pinned in
tests/test_oracle_amg.py), on the oracle's dam-break pressure systems:
  * one V-cycle is bitwise the oracle's (the levels, Galerkin products,
    smoother, restriction and prolongation run the same operations in the
    same order);
  * the solve converges on the FP64 criterion with the oracle's iteration
    count (+-1: the inner products are summed in another order) and a
    solution within 1e-6 of the oracle's at tol 1e-10;
  * bad DOF cells and CSR are refused."""
import numpy as np
import pytest

from tests.test_oracle_amg import dam_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [24, 40])
def test_vcycle_is_bitwise_the_oracles(oracle, wc, n):
    rs, col, val, b, ijk, nx = dam_system(oracle, n)
    assert len(oracle.amg_levels(rs, col, val, ijk, nx)) >= 2
    r = np.random.default_rng(n).standard_normal(len(b))
    z_gpu = wc.amg_apply_csr(rs, col, val, ijk, nx, r)
    z_ref = oracle.amg_vcycle(rs, col, val, ijk, nx, r)
    assert np.array_equal(z_gpu, z_ref)
    z_gpu_b = wc.amg_apply_csr(rs, col, val, ijk, nx, b)
    assert np.array_equal(z_gpu_b, oracle.amg_vcycle(rs, col, val, ijk, nx, b))


@pytest.mark.parametrize("n", [24, 40])
def test_amgpcg_matches_the_oracle(oracle, wc, n):
    rs, col, val, b, ijk, nx = dam_system(oracle, n)
    ok_r, x_r, it_r, res_r = oracle.amgpcg(rs, col, val, ijk, nx, b, tol=1e-10, maxit=300)
    ok, x, it, res = wc.amgpcg_solve_csr(rs, col, val, ijk, nx, b, tol=1e-10, maxit=300)
    assert ok_r and ok
    assert abs(it - it_r) <= 1
    assert res <= 1e-10 * np.abs(b).max()
    assert np.abs(x - x_r).max() <= 1e-6 * np.abs(x_r).max()
    # fewer iterations than the MIC(0) PCG on the same system
    ok_m, x_m, it_m, res_m = wc.pcg_solve_csr(rs, col, val, b, tol=1e-10, maxit=3000)
    assert ok_m and it < it_m


def test_amg_rejects_bad_input(wc, oracle):
    rs, col, val, b, ijk, nx = dam_system(oracle, 12)
    bad = ijk.copy()
    bad[3, 0] = nx[0]
    with pytest.raises(wc.WetclothError):
        wc.amgpcg_solve_csr(rs, col, val, bad, nx, b)
    with pytest.raises(wc.WetclothError):
        wc.amg_apply_csr(rs, col, val, ijk, (nx[0], nx[1], 0), b)
    # a zero right-hand side returns at once
    ok, x, it, res = wc.amgpcg_solve_csr(rs, col, val, ijk, nx, np.zeros(len(b)))
    assert ok and it == 0 and not x.any()
