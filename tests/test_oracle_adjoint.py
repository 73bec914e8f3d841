"""The oracle's adjoint of the substep (wco_substep_adjoint) against central
differences of the oracle's own forward run, with every discrete decision
of the run (activated buckets, the liquid phi's particle per cell, liquid
cells and DOFs, theta clamps, faces with mass, particles the solid
projection moves) recorded on the first pass and replayed in the perturbed
runs, so the map is smooth and the two must agree to truncation error.

Scene: a 12^3 dam break with a seeded velocity field, 3 substeps at dt =
1/480, PCG to 1e-13; loss sum |x - x_target|^2 with the target the same run
from a +0.1 m/s x-velocity.  The directional derivatives along random
directions of the initial velocities, and the derivative by the gravity
magnitude, agree to 1e-7 relative (measured ~1e-10).  Each part of the
adjoint is exercised: the oracle built without its theta terms, its solid
projection, its liquid-phi particle term or P2G's weight derivative misses
by 1e-4 or more.
"""
import numpy as np
import pytest

N = 12
DT = 1.0 / 480
STEPS = 3
EPS = 1e-4


@pytest.fixture(scope="module")
def scene(oracle):
    g, x, v = oracle.dam_break(N, block_hi=(0.5, 0.5, 0.5), seed=3, sigma=0.5)
    return x, v, np.zeros((9, x.shape[1]))


def check(oracle, scene, L, ndir=4):
    x, v, C0 = scene
    run = oracle.AdjointRun((N, N, N), 1.0 / N, x, v, C0, STEPS, DT, L=L)
    # the recorded decisions include projected particles and clamped faces
    assert sum(int(r["proj"].sum()) for r in run.records) > 0
    assert sum(int(c.sum()) for r in run.records for c in r["clamp"]) > 0
    tgt = run.forward(v + np.array([[0.1], [0.0], [0.0]]))
    loss, gv, gg = run.gradient(v, tgt)

    def loss_at(vv, g=None):
        return float(((run.forward(vv, g) - tgt) ** 2).sum())
    rng = np.random.default_rng(0)
    for _ in range(ndir):
        d = rng.standard_normal(v.shape)
        fd = (loss_at(v + EPS * d) - loss_at(v - EPS * d)) / (2 * EPS)
        assert abs(fd - (gv * d).sum()) <= 1e-7 * abs(fd), (fd, (gv * d).sum())
    fdg = (loss_at(v, 9.81 + EPS) - loss_at(v, 9.81 - EPS)) / (2 * EPS)
    assert abs(fdg - gg) <= 1e-7 * abs(fdg), (fdg, gg)


def test_adjoint_matches_central_differences(oracle, scene):
    check(oracle, scene, oracle.lib())


@pytest.mark.parametrize("plant,what", [
    (13, "no theta terms"), (14, "no solid projection"), (15, "no liquid-phi particle term"),
    (16, "no P2G weight derivative")])
def test_adjoint_check_fails_on_planted_omission(oracle, scene, tmp_path_factory, plant, what):
    L = oracle.load_variant(oracle.build_variant(str(tmp_path_factory.mktemp("adj")), plant))
    with pytest.raises(AssertionError):
        check(oracle, scene, L, ndir=2)
