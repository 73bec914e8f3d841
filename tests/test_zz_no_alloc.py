"""wc_pcg_solve_csr_noalloc against wc_pcg_solve_csr: the same kernels in
the same order on device memory taken from the library's fixed array, so the
solutions, iteration counts and residuals are bitwise equal.  Two calls in a
row, the second on a smaller system (the array keeps the first call's
contents), then a system larger than the array, which is refused."""
import numpy as np
import pytest

from tests.helpers import csr_parts, dam_break_system, grid9_mmatrix

pytestmark = pytest.mark.gpu


def _same(wc, rs, col, val, b, tol):
    ok0, x0, it0, r0 = wc.pcg_solve_csr(rs, col, val, b, tol=tol, maxit=2000)
    ok1, x1, it1, r1 = wc.pcg_solve_csr(rs, col, val, b, tol=tol, maxit=2000, no_alloc=True)
    assert ok0 and ok1
    assert it0 == it1 and r0 == r1
    assert np.array_equal(x0, x1)
    return it1


def test_noalloc_matches_allocating_solve_twice(oracle, wc):
    # a dam-break pressure system (24^3), then a smaller one (16^3)
    rs, col, val, b = dam_break_system(oracle, 24)
    it_a = _same(wc, rs, col, val, b, 1e-8)
    rs2, col2, val2, b2 = dam_break_system(oracle, 16)
    assert len(rs2) < len(rs)
    it_b = _same(wc, rs2, col2, val2, b2, 1e-8)
    print(f"no-alloc PCG bitwise equal: {len(rs) - 1} rows in {it_a} iterations, "
          f"then {len(rs2) - 1} rows in {it_b}")


def test_noalloc_matrix_with_fill_and_zero_rows(wc):
    A = grid9_mmatrix(40, 30, seed=5, zero_rows=3)
    rs, col, val = csr_parts(A)
    b = np.random.default_rng(6).standard_normal(A.shape[0])
    b[np.asarray(A.getnnz(axis=1)) == 0] = 0.0
    _same(wc, rs, col, val, b, 1e-10)


def test_noalloc_refuses_systems_beyond_its_array(wc):
    # a diagonal system of 5.6 M rows needs more than the fixed 128 MiB
    n = 5_600_000
    rs = np.arange(n + 1, dtype=np.uint32)
    col = np.arange(n, dtype=np.uint32)
    val = np.ones(n)
    with pytest.raises(wc.WetclothError, match="fixed"):
        wc.pcg_solve_csr(rs, col, val, np.ones(n), no_alloc=True)
