"""The pipelined MIC(0) sweeps against the level-scheduled ones: both apply
the same operations in the same order, so every PCG iterate, the iteration
count and the pressure are bitwise identical."""
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
def test_pipelined_equals_level_scheduled(oracle, wc, n, block_hi):
    s = prepared(wc, oracle, n, block_hi=block_hi)
    a = solve(s, True)
    b = solve(s, False)
    assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])  # same system
    assert a[0] and b[0]
    assert a[1] == b[1]
    assert a[2] == b[2]
    assert np.array_equal(a[3], b[3])


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
