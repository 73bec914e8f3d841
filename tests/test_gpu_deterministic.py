"""The deterministic option (wc_set_option "deterministic"): P2G's sums in
64-bit fixed point, so that a run repeats bit for bit.

* Two runs of the same seeded dam break (24^3, 15 substeps, FP64 and the
  mixed form) give the same particles bit for bit; the mixed form does so
  by default.
* The fixed-point sums against the oracle's FP64 sums at 16^3: P2G's mass
  and volume within 1e-12 relative; the velocity u = P / M node by node
  within the format's bound, |u_gpu - u_cpu| * W <= 300 * 2^-49 +
  1e-13 |u_cpu| W, W = the node's weight sum (each of a node's at most
  27 * 8 momentum contributions is rounded by at most 2^-49, and a light
  node's division by its small W magnifies that); and a 20-substep
  trajectory at 32^3 within 1e-9 dx.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 1.0 / 480


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def run(wc, n, precision, steps, seed=3):
    s = wc.Scene((n, n, n), 1.0 / n, precision=precision)
    s.set_option("deterministic", 1)
    s.seed_block((0.0, 0.0, 0.0), (0.5, 0.6, 1.0), seed=seed, sigma=0.5)
    for _ in range(steps):
        assert s.substep(DT, 1e-8, 2000)[0]
    out = s.particles()
    s.close()
    return out


@pytest.mark.parametrize("mixed", [False, True])
def test_deterministic_runs_repeat_bitwise(wc, mixed):
    prec = wc.WC_MIXED if mixed else wc.WC_FP64
    a = run(wc, 24, prec, 15)
    b = run(wc, 24, prec, 15)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_mixed_form_is_deterministic_by_default(wc):
    """WC_MIXED contexts sum P2G in fixed point unless told otherwise."""
    def plain(n=24, steps=10):
        s = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED)
        s.seed_block((0.0, 0.0, 0.0), (0.5, 0.6, 1.0), seed=5, sigma=0.5)
        for _ in range(steps):
            assert s.substep(DT, 1e-8, 2000)[0]
        out = s.particles()
        s.close()
        return out
    a, b = plain(), plain()
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_deterministic_sums_match_oracle(oracle, wc):
    n = 16
    _, x, v = oracle.dam_break(n, seed=1, sigma=0.3)
    cpu = oracle.OracleScene((n, n, n), 1.0 / n, x, v)
    gpu = wc.Scene((n, n, n), 1.0 / n)
    gpu.set_option("deterministic", 1)
    gpu.set_particles(x, v)
    cpu.activate()
    cpu.update_liquid_phi()
    cpu.update_solid_phi()
    cpu.update_solid_weights()
    cpu.p2g()
    gpu.rebucketizeParticles()
    gpu.mapParticleNodesAPIC()
    fm = 1000.0 * (1.0 / n) ** 3 / 8.0  # the particle mass
    for d in range(3):
        assert rel(gpu.field("mass", d), cpu.field("mass", d)) < 1e-12
        assert rel(gpu.field("vol", d), cpu.field("vol", d)) < 1e-12
        W = cpu.field("mass", d) / fm
        uc, ug = cpu.field("vel", d), gpu.field("vel", d)
        assert np.all(np.abs(ug - uc) * W <= 300 * 2.0 ** -49 + 1e-13 * np.abs(uc) * W)
    n = 32
    _, x, v = oracle.dam_break(n)
    cpu = oracle.OracleScene((n, n, n), 1.0 / n, x, v)
    gpu = wc.Scene((n, n, n), 1.0 / n)
    gpu.set_option("deterministic", 1)
    gpu.set_particles(x, v)
    for _ in range(20):
        cpu.substep(DT, 1e-8, 1000)
        assert gpu.substep(DT, 1e-8, 1000)[0]
    err = np.abs(gpu.particles()[0] - cpu.particles()[0]).max() * n
    print(f"deterministic P2G: 20 substeps at 32^3 within {err:.2e} dx of the oracle")
    assert err < 1e-9
