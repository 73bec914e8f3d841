"""The mixed-precision form (WC_MIXED) against the FP64 oracle.

The mixed form keeps particle state and the transfers in FP32 and runs the
PCG on FP32 vectors inside FP64 iterative refinement, so nothing here is
bitwise.  What is exact: seeding (FP64 seeds rounded once to FP32) and the
host round trip.  What is bounded, with the tolerances in each test:
  * the pressure solve meets the FP64 criterion |b - A p|_inf <= tol |b|_inf
    on the FP64 system (checked on the CPU with the oracle's multiply), and
    the pressure lies within 20 tol (relative, max norm) of the oracle's FP64
    PCG solution of the same system (measured: 1.2-1.6 tol);
  * P2G grid fields agree with the oracle run on the FP32-rounded particles
    to 1e-5 relative (FP32 weights, FP64 sums; measured ~1e-7);
  * 20 substeps at 32^3 keep positions within 1e-4 dx of the oracle
    (measured 1.5e-5 dx against the FP64 GPU path).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 1.0 / 480


def rel(a, b):
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def test_mixed_seeding_rounds_fp64_seeds(oracle, wc):
    n = 32
    _, x, v = oracle.dam_break(n, seed=3, sigma=0.5)
    s = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED)
    cnt = s.seed_block((0.0, 0.0, 0.0), (0.4, 0.6, 1.0), seed=3, sigma=0.5)
    assert cnt == x.shape[1]
    xg, vg, Cg = s.particles()
    assert np.array_equal(xg, f32(x))
    assert np.array_equal(vg, f32(v))
    assert not Cg.any()


def test_mixed_host_roundtrip_rounds_to_fp32(wc):
    n = 8
    rng = np.random.default_rng(5)
    npart = 200
    x = rng.uniform(0.2, 0.8, (3, npart))
    v = rng.standard_normal((3, npart))
    Cm = rng.standard_normal((9, npart))
    s = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED)
    s.set_particles(x, v, Cm)
    s.rebucketizeParticles()
    xg, vg, Cg = s.particles()
    assert np.array_equal(xg, f32(x)) and np.array_equal(vg, f32(v))
    assert np.array_equal(Cg, f32(Cm))


def test_mixed_p2g_matches_oracle(oracle, wc):
    """FP32 transfers with FP64 atomic sums within 1e-5 of the oracle; the
    default fixed-point sums ("deterministic", units of 2^-56 for weights
    and 2^-48 for momentum) differ from those node by node only by the
    rounding of each of a node's at most 27 * 8 contributions:
    |u_fixed - u_atomic| W <= 300 * 2^-49 + 1e-12 |u| W, W the weight sum."""
    n = 24
    _, x, v = oracle.dam_break(n, seed=1, sigma=0.3)
    x, v = f32(x), f32(v)
    cpu = oracle.OracleScene((n, n, n), 1.0 / n, x, v)
    cpu.activate()
    cpu.update_liquid_phi()
    cpu.update_solid_phi()
    cpu.update_solid_weights()
    cpu.p2g()
    fields = {}
    for det in (0, 1):
        gpu = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED)
        gpu.set_option("deterministic", det)
        gpu.set_particles(x, v)
        gpu.rebucketizeParticles()
        gpu.mapParticleNodesAPIC()
        assert np.array_equal(gpu.field("act"), cpu.field("act"))
        assert rel(gpu.field("lphi"), cpu.field("lphi")) < 1e-5
        for d in range(3):
            assert np.array_equal(gpu.field("w", d), cpu.field("w", d))
            assert rel(gpu.field("mass", d), cpu.field("mass", d)) < 1e-5
            assert rel(gpu.field("vol", d), cpu.field("vol", d)) < 1e-5
            if det == 0:
                assert rel(gpu.field("vel", d), cpu.field("vel", d)) < 1e-5
            fields[det, d] = gpu.field("vel", d), gpu.field("mass", d)
        gpu.close()
    fm = np.float32(1000.0 * (1.0 / n) ** 3 / 8.0)  # the mixed form's FP32 particle mass
    for d in range(3):
        (ua, Ma), (uf, _) = fields[0, d], fields[1, d]
        W = Ma / float(fm)
        assert np.all(np.abs(uf - ua) * W <= 300 * 2.0 ** -49 + 1e-12 * np.abs(ua) * W)


@pytest.mark.parametrize("n,tol", [(16, 1e-6), (32, 1e-6), (32, 1e-8)])
def test_mixed_solve_meets_fp64_criterion(oracle, wc, n, tol):
    s = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED)
    s.seed_block((0.0, 0.0, 0.0), (0.4, 0.6, 1.0), seed=1, sigma=0.3)
    s.rebucketizeParticles()
    s.mapParticleNodesAPIC()
    s.stepVelocity(DT)
    ok, it, res = s.projectFine(DT, tol, 2000)
    assert ok and s.refinement_passes() >= 1
    rs, col, val, b, p = s.system()
    r = b - oracle.spmv(rs, col, val, p)
    assert np.abs(r).max() <= tol * np.abs(b).max()
    assert res <= tol * np.abs(b).max()
    okc, p64, itc, _ = oracle.pcg(rs, col, val, b, tol=tol, maxit=2000)
    assert okc
    print(f"n={n} tol={tol:g}: mixed {it} iterations in {s.refinement_passes()} passes, "
          f"FP64 {itc}; pressure rel diff {rel(p, p64):.3e}")
    assert rel(p, p64) < 20 * tol


def test_mixed_options_and_fp64_passes(wc):
    s = wc.Scene((8, 8, 8), 1.0 / 8)
    assert s.refinement_passes() == 0
    m = wc.Scene((8, 8, 8), 1.0 / 8, precision=wc.WC_MIXED)
    m.set_option("mixed_digits", 6)
    m.set_option("mixed_refinements", 3)
    with pytest.raises(wc.WetclothError):
        m.set_option("mixed_digits", 0)


def test_mixed_trajectory_32(oracle, wc):
    """20 substeps at 32^3: positions within 1e-4 dx of the FP64 oracle."""
    n = 32
    _, x, v = oracle.dam_break(n)
    cpu = oracle.OracleScene((n, n, n), 1.0 / n, x, v)
    gpu = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED)
    gpu.set_particles(x, v)
    for _ in range(20):
        cpu.substep(DT, 1e-8, 1000)
        ok, it, res = gpu.step(DT, 1e-8, 1000)
        assert ok
    xg, vg, _ = gpu.particles()
    xc, vc, _ = cpu.particles()
    err = np.abs(xg - xc).max() * n
    print(f"mixed: max position error {err:.3e} dx, velocity rel {rel(vg, vc):.3e}")
    assert err < 1e-4
