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

from tests.helpers import reference_bucket_frame
from tests.test_oracle_amg import dam_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [24, 40, 64])
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
    # two DOFs in one cell are refused (a coarse DOF keeps one child per
    # position of its 2x2x2 block)
    dup = ijk.copy()
    dup[5] = dup[4]
    with pytest.raises(wc.WetclothError, match="no two DOFs"):
        wc.amgpcg_solve_csr(rs, col, val, dup, nx, b)
    # a zero right-hand side returns at once
    ok, x, it, res = wc.amgpcg_solve_csr(rs, col, val, ijk, nx, np.zeros(len(b)))
    assert ok and it == 0 and not x.any()


def test_context_solver_option(oracle, wc):
    """set_option("solver", 1): wc_project solves the context's own pressure
    system with AMG-PCG; the solution is the oracle's AMG-PCG on the
    downloaded system and the DOFs' cells (bucket handle * 4 + node in the
    bucket: the dense node coordinates, as Pressure.cpp:433-436 on the
    reference's bucket grid), and whole substeps converge with it."""
    n, dt = 24, 1.0 / 480
    g, x, v = oracle.dam_break(n, seed=2, sigma=0.5)
    cpu = oracle.OracleScene((n, n, n), 1.0 / n, x, v)
    s = wc.Scene((n, n, n), 1.0 / n)
    s.set_particles(x, v)
    with pytest.raises(wc.WetclothError):
        s.set_option("solver", 2)
    s.set_option("solver", 1)
    s.rebucketizeParticles()
    s.mapParticleNodesAPIC()
    s.stepVelocity(dt)
    ok, it, res = s.projectFine(dt, tol=1e-10, maxit=300)
    assert ok and it > 0
    rs, col, val, b, p = s.system()
    assert res <= 1e-10 * np.abs(b).max()
    nx = tuple(int(t) for t in cpu.grid.nx)
    assert s.grid_info()[0] == nx
    dof = s.field("dof")
    ii, jj, kk = np.meshgrid(np.arange(nx[0]), np.arange(nx[1]), np.arange(nx[2]),
                             indexing="ij")
    d = dof[oracle.node_index_array(cpu.grid, ii.ravel(), jj.ravel(), kk.ravel())]
    m = d >= 0
    # the cells in the frame of the reference's bucket grid, whose origin and
    # size follow the particles' bounding box (TwoDScene.cpp:1588-1690)
    origin, dims = reference_bucket_frame(x, 1.0 / n)
    ijk = np.zeros((len(b), 3), dtype=np.int32)
    ijk[d[m]] = np.stack([ii.ravel()[m], jj.ravel()[m], kk.ravel()[m]], axis=1) \
        - (origin + oracle.GHOST)
    assert ijk.min() >= 0 and np.all(ijk < np.array(dims))
    ok_r, x_r, it_r, _ = oracle.amgpcg(rs, col, val, ijk, dims, b, tol=1e-10, maxit=300)
    assert ok_r and abs(it - it_r) <= 1
    assert np.abs(p - x_r).max() <= 1e-6 * np.abs(x_r).max()
    # the V-cycle on those cells is the oracle's, bit for bit
    r = np.random.default_rng(7).standard_normal(len(b))
    assert np.array_equal(wc.amg_apply_csr(rs, col, val, ijk, dims, r),
                          oracle.amg_vcycle(rs, col, val, ijk, dims, r))
    s.applyPressureDragFluid(dt)
    s.map_node_particles_advect(dt)
    for _ in range(3):
        ok, it, res = s.step(dt, tol=1e-8, maxit=300)
        assert ok
    # and back to MIC(0) on the same context
    s.set_option("solver", 0)
    ok, it, res = s.step(dt, tol=1e-8, maxit=1000)
    assert ok
    s.close()


def test_solver_option_in_a_mixed_context(oracle, wc):
    """A WC_MIXED context with "solver" = 1 solves its FP64 system with
    AMG-PCG to the FP64 criterion; no refinement passes are reported."""
    n, dt = 24, 1.0 / 480
    s = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED)
    s.seed_block((0.0, 0.0, 0.0), (0.4, 0.6, 1.0), seed=3, sigma=0.3)
    ok, it, res = s.step(dt, tol=1e-8, maxit=1000)  # MIC(0) with refinement first
    assert ok and s.refinement_passes() >= 1
    s.set_option("solver", 1)
    for _ in range(3):
        ok, it, res = s.step(dt, tol=1e-8, maxit=300)
        assert ok and 0 < it < 30
    assert s.refinement_passes() == 0
    s.rebucketizeParticles()
    s.mapParticleNodesAPIC()
    s.stepVelocity(dt)
    ok, it, res = s.projectFine(dt, tol=1e-8, maxit=300)
    rs, col, val, b, p = s.system()
    r = b - oracle.spmv(rs, col, val, p)
    assert ok and np.abs(r).max() <= 1e-8 * np.abs(b).max()
    s.close()


def block_coupled(rs, col, val, ijk):
    """The system plus a -0.1 coupling between opposite corners of each
    2x2x2 block (0.1 more on both diagonals): rows of up to 8 entries, coarse
    rows still 7-point.  Opposite corners differ in all three coordinates,
    so they have different colours: the reference smooths each colour in a
    parallel loop (AlgebraicMultigrid.h:123, 150), which is deterministic
    only when rows of one colour do not couple."""
    import scipy.sparse as sp
    n = len(rs) - 1
    A = sp.csr_matrix((val, col.astype(np.int64), rs.astype(np.int64)), shape=(n, n)).tolil()
    where = {tuple(c): d for d, c in enumerate(ijk.tolist())}
    for d, (i, j, k) in enumerate(ijk.tolist()):
        for di, dj, dk in ((1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1)):
            e = where.get((i + di, j + dj, k + dk))
            if e is None or ((i // 2, j // 2, k // 2) != ((i + di) // 2, (j + dj) // 2,
                                                          (k + dk) // 2)):
                continue
            A[d, e] = -0.1
            A[e, d] = -0.1
            A[d, d] += 0.1
            A[e, e] += 0.1
    A = A.tocsr()
    A.sort_indices()
    return A.indptr.astype(np.uint32), A.indices.astype(np.uint32), A.data.copy()


def test_rows_longer_than_seven(oracle, wc):
    """Rows of more than 7 entries take the generic kernels (the
    colour-ordered copies hold 7): still bitwise the oracle's V-cycle, and
    the solve matches it."""
    rs0, col0, val0, b, ijk, nx = dam_system(oracle, 24)
    rs, col, val = block_coupled(rs0, col0, val0, ijk)
    assert np.diff(rs).max() > 7
    r = np.random.default_rng(5).standard_normal(len(b))
    assert np.array_equal(wc.amg_apply_csr(rs, col, val, ijk, nx, r),
                          oracle.amg_vcycle(rs, col, val, ijk, nx, r))
    ok_r, x_r, it_r, _ = oracle.amgpcg(rs, col, val, ijk, nx, b, tol=1e-10, maxit=300)
    ok, x, it, res = wc.amgpcg_solve_csr(rs, col, val, ijk, nx, b, tol=1e-10, maxit=300)
    assert ok_r and ok and abs(it - it_r) <= 1
    assert np.abs(x - x_r).max() <= 1e-6 * np.abs(x_r).max()


def test_amgpcg_not_converged_after_maxit(oracle, wc):
    """maxit = 1: the solve returns false (WC_NOT_CONVERGED, not an error)
    after one iteration, like the reference (AlgebraicMultigrid.h:352-360),
    with the oracle's iterate."""
    rs, col, val, b, ijk, nx = dam_system(oracle, 24)
    ok_r, x_r, it_r, res_r = oracle.amgpcg(rs, col, val, ijk, nx, b, tol=1e-10, maxit=1)
    ok, x, it, res = wc.amgpcg_solve_csr(rs, col, val, ijk, nx, b, tol=1e-10, maxit=1)
    assert not ok_r and not ok and it == it_r == 1
    assert np.abs(x - x_r).max() <= 1e-10 * np.abs(x_r).max()
    assert abs(res - res_r) <= 1e-8 * res_r
