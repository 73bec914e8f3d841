"""The pipelined MIC(0) sweeps and the box-resident PCG of a grid system.

The sweeps and the box SpMV (windowed or one row per thread) apply the
reference's operations in the reference's order, so on the context's own system they are bitwise equal to
the oracle's apply_preconditioner and multiply.  The box PCG sums its dot
products in box order rather than DOF order, so whole solves agree with the
level-scheduled (DOF-ordered) solve to rounding: same convergence, iteration
counts within one, pressures to 1e-9 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def prepared(wc, oracle, n, seed=1, sigma=0.5, block_hi=(0.4, 0.6, 1.0)):
    """A scene up to the pressure solve (P2G sums with atomics, so two scenes
    differ in the last bits: compare solves on one assembled system)."""
    g, x, v = oracle.dam_break(n, block_hi=block_hi, seed=seed, sigma=sigma)
    s = wc.Scene((n, n, n), 1.0 / n)
    s.set_particles(x, v)
    s.rebucketizeParticles()
    s.mapParticleNodesAPIC()
    s.stepVelocity(1.0 / 480)
    return s


def solve(s, sweeps):
    s.set_option("pipelined_sweeps", 1 if sweeps else 0)
    ok, it, res = s.projectFine(1.0 / 480, 1e-8, 2000)
    rs, col, val, b, p = s.system()
    return ok, it, res, p, val, b


def solve_with(wc, oracle, n, sweeps, seed=1, sigma=0.5, block_hi=(0.4, 0.6, 1.0)):
    return solve(prepared(wc, oracle, n, seed, sigma, block_hi), sweeps)


@pytest.mark.parametrize("n,block_hi", [(16, (0.4, 0.6, 1.0)), (32, (0.4, 0.6, 1.0)),
                                         (40, (0.9, 0.3, 0.55)), (64, (0.4, 0.6, 1.0))])
def test_pipelined_matches_level_scheduled(oracle, wc, n, block_hi):
    s = prepared(wc, oracle, n, block_hi=block_hi)
    a = solve(s, True)
    b = solve(s, False)
    assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])  # same system
    assert a[0] and b[0]
    assert abs(a[1] - b[1]) <= 1
    assert np.abs(a[3] - b[3]).max() <= 1e-9 * np.abs(b[3]).max()


@pytest.mark.parametrize("n,block_hi", [(16, (0.4, 0.6, 1.0)), (40, (0.9, 0.3, 0.55)),
                                         (64, (0.4, 0.6, 1.0)),
                                         (24, (1.0, 0.15, 0.2))])  # one tile row, few layers
def test_box_apply_and_multiply_bitwise(oracle, wc, n, block_hi):
    """apply_preconditioner and multiply on the context's system, through the
    pipelined sweeps (one wave per 8x8 quarter tile, and one workgroup per
    16x16 tile) and the box stencil, equal the oracle's bit for bit."""
    s = prepared(wc, oracle, n, block_hi=block_hi)
    ok, it, res, p, val, b = solve(s, True)
    assert ok
    rs, col, val, b, _ = s.system()
    rng = np.random.default_rng(n)
    r = rng.standard_normal(len(b))
    invdiag, lv = oracle.mic0(rs, col, val)
    want = oracle.apply_precond(rs, col, invdiag, lv, r)
    for sub in (1, 0, 1):
        s.set_option("sweep_subtiles", sub)
        assert np.array_equal(s.apply_preconditioner(r), want), sub
    y = oracle.spmv(rs, col, val, r)
    # the windowed product (runs of 1, 7 and 64 steps) and the slot-parallel one
    for win, run in ((1, 1), (1, 7), (0, 64), (1, 64)):
        s.set_option("spmv_window", win)
        s.set_option("spmv_run", run)
        assert np.array_equal(s.multiply(r), y), (win, run)
    # and the level-scheduled kernels agree with both
    solve(s, False)
    assert np.array_equal(s.apply_preconditioner(r), oracle.apply_precond(rs, col, invdiag, lv, r))
    assert np.array_equal(s.multiply(r), oracle.spmv(rs, col, val, r))


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("n,block_hi", [(32, (0.4, 0.6, 1.0)), (48, (0.9, 0.3, 0.55))])
def test_subtile_and_tile_solves_agree(oracle, wc, mixed, n, block_hi):
    """Whole solves with either sweep kernel: the applies are the same
    values, and only the z . r partials are summed over other units (a
    quarter tile instead of a tile), so the iterations agree to within one
    and the pressures to rounding."""
    g, x, v = oracle.dam_break(n, block_hi=block_hi, seed=3, sigma=0.5)
    res = {}
    for sub in (1, 0):
        s = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED if mixed else wc.WC_FP64)
        s.set_option("sweep_subtiles", sub)
        s.set_particles(x, v)
        s.rebucketizeParticles()
        s.mapParticleNodesAPIC()
        s.stepVelocity(1.0 / 480)
        ok, it, r = s.projectFine(1.0 / 480, 1e-8, 2000)
        assert ok, sub
        res[sub] = (it, s.system()[4])
        s.close()
    (i1, p1), (i0, p0) = res[1], res[0]
    print(f"{n}^3 {'mixed' if mixed else 'FP64'}: iterations {i1} (quarter tiles) / {i0} (tiles), "
          f"pressure difference {np.abs(p1 - p0).max() / np.abs(p0).max():.2e}")
    assert abs(i1 - i0) <= 1
    assert np.abs(p1 - p0).max() <= (1e-6 if mixed else 1e-9) * np.abs(p0).max()


@pytest.mark.parametrize("mixed", [False, True])
def test_windowed_and_slot_spmv_solves_agree(oracle, wc, mixed):
    """Whole solves with either SpMV: the products are the same values, and
    only the s . A s partials are summed over other blocks, so the
    iterations agree to within one and the pressures to rounding."""
    n = 40
    g, x, v = oracle.dam_break(n, block_hi=(0.9, 0.3, 0.55), seed=5, sigma=0.5)
    res = {}
    for win in (1, 0):
        s = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED if mixed else wc.WC_FP64)
        s.set_option("spmv_window", win)
        s.set_particles(x, v)
        s.rebucketizeParticles()
        s.mapParticleNodesAPIC()
        s.stepVelocity(1.0 / 480)
        ok, it, r = s.projectFine(1.0 / 480, 1e-8, 2000)
        assert ok, win
        res[win] = (it, s.system()[4])
        s.close()
    (i1, p1), (i0, p0) = res[1], res[0]
    assert abs(i1 - i0) <= 1
    assert np.abs(p1 - p0).max() <= (1e-6 if mixed else 1e-9) * np.abs(p0).max()


@pytest.mark.parametrize("mixed", [False, True])
def test_sweep_abort_reaches_the_host_and_the_next_solve_recovers(oracle, wc, mixed):
    """The sweeps' abort flag (set by a sweep whose bounded wait ran out;
    here by the "debug_abort_sweeps" hook before the loop) reaches the host
    through the update kernel's first residual partial: the solve fails with
    the sweep-wait error instead of waiting, and the next solve without the
    hook converges as usual."""
    n = 24
    g, x, v = oracle.dam_break(n, seed=4, sigma=0.4)
    s = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED if mixed else wc.WC_FP64)
    s.set_particles(x, v)
    s.rebucketizeParticles()
    s.mapParticleNodesAPIC()
    s.stepVelocity(1.0 / 480)
    s.set_option("debug_abort_sweeps", 1)
    with pytest.raises(wc.WetclothError, match="tile wait"):
        s.projectFine(1.0 / 480, 1e-8, 2000)
    s.set_option("debug_abort_sweeps", 0)
    ok, it, res = s.projectFine(1.0 / 480, 1e-8, 2000)
    assert ok and it > 0
    s.close()


def test_spmv_run_option_rejects_bad_values(wc):
    s = wc.Scene((16, 16, 16), 1.0 / 16)
    for bad in (0, -3, 5000):
        with pytest.raises(Exception):
            s.set_option("spmv_run", bad)
    s.close()


def test_pipelined_solve_matches_oracle(oracle, wc):
    n = 24
    g, x, v = oracle.dam_break(n, seed=2, sigma=0.4)
    cpu = oracle.OracleScene((n, n, n), 1.0 / n, x, v)
    cpu.pre_solve(1.0 / 480)
    cpu.solve_pressure(1e-8, 2000)
    _, _, _, _, pc = cpu.system()
    ok, it, res, pg, _, _ = solve_with(wc, oracle, n, True, seed=2, sigma=0.4)
    assert ok
    assert np.abs(pg - pc).max() < 1e-7 * np.abs(pc).max()


def test_kernel_timing_counts_launches(oracle, wc):
    """With kernel_timing on, every sweep and SpMV launch of a solve is
    counted once: one MIC(0) apply before the loop and one after every
    iteration (the apply after an iteration is enqueued before the host has
    read that iteration's residual, so the converging iteration has one
    too), one SpMV per iteration; units are DOFs."""
    s = prepared(wc, oracle, 24)
    s.set_option("kernel_timing", 1)
    ok, it, res = s.projectFine(1.0 / 480, 1e-8, 2000)
    st = s.kernel_stats()
    ndof = len(s.system()[3])
    assert ok and it > 0
    for k in ("sweep_fwd", "sweep_bwd"):
        ms, nl, units = st[k]
        assert nl == it + 1 and units == (it + 1) * ndof and ms > 0
    ms, nl, units = st["spmv"]
    assert nl == it and units == it * ndof and ms > 0
    s.set_option("kernel_timing", 0)  # clears the counters
    assert all(v == (0.0, 0, 0) for v in s.kernel_stats().values())
    # kernel_timing n > 1: the apply before the loop, then the launches of
    # every n-th iteration
    s.set_option("kernel_timing", 4)
    ok, it, res = s.projectFine(1.0 / 480, 1e-8, 2000)
    st = s.kernel_stats()
    assert ok
    for k in ("sweep_fwd", "sweep_bwd"):
        ms, nl, units = st[k]
        assert nl == 1 + it // 4 and units == nl * ndof and ms > 0
    ms, nl, units = st["spmv"]
    assert nl == it // 4 and units == nl * ndof
    # the transfers: every launch, units = particles
    np_ = s.num_particles()
    s.set_option("kernel_timing", 4)
    s.step(1.0 / 480, 1e-8, 2000)
    st = s.kernel_stats()
    for k in ("p2g", "g2p"):
        ms, nl, units = st[k]
        assert nl == 1 and units == np_ and ms > 0
