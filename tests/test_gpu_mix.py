"""The mixture terms on the GPU (a resting porous solid in the liquid path,
wc_set_elasto_particles) against the FP64 oracle, through the C ABI.

* Phase by phase at 16^3 with the wet sheet (seeded orientations and shape
  factors, a fifth of its pores filled, 2 cm thick so that psi reaches
  ~0.15) under a liquid block with a seeded velocity: the solid's face and
  cell fields, the saturations and the pore pressure, the grid update with
  its drag ratio Q and hdV, the assembled system, the pressure and the
  projected velocities, the particles after G2P.  The solid's and the
  cell-centred sums are atomic sums on the GPU (another order than the
  oracle's), so FP64 fields agree to 1e-12 relative.
* Trajectories of 20 substeps at 24^3 in FP64 and in the mixed form, and
  5 substeps solved with AMG-PCG.
* The input checks, and that the adjoint refuses a scene with a solid.
* The 128^3 wet-cloth scene: the 128^3 liquid column (4.0 M particles,
  block 0.4 x 0.6 x 1 m) starting just above a 257 x 257-vertex cloth at
  y = 0.3 (1 m square, 2 mm, 0.2 kg/m^2, porosity U[0.3, 0.7]), 150
  substeps of dt = 1/960 in the mixed form: every solve converges, the
  particles stay finite and inside the container, and the cloth holds
  liquid back (fewer particles below it than in the same run without it);
  and one substep of each form from the same FP32 state once the column
  has reached the cloth agrees within 1e-4 dx.
"""
import numpy as np
import pytest

from tests import mixscene

pytestmark = pytest.mark.gpu

DT = 1.0 / 480
RT = 1e-12


def rel(a, b):
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


def pair(oracle, wc, n, solid, lo, hi, sigma=0.5, seed=1, precision=None, **prm):
    g = oracle.grid((n, n, n), 1.0 / n)
    x = oracle.seed_block(g, lo, hi, seed)
    v = oracle.seed_velocity(g, seed, sigma, x) if sigma > 0 else np.zeros_like(x)
    cpu = oracle.OracleScene((n, n, n), 1.0 / n, x, v)
    cpu.set_elasto(*mixscene.args(solid), prm=oracle.mix_params(**prm))
    kw = {} if precision is None else {"precision": precision}
    gpu = wc.Scene((n, n, n), 1.0 / n, **kw)
    gpu.set_particles(x, v)
    gpu.set_elasto(*mixscene.args(solid), prm=wc.mix_params(**prm))
    return cpu, gpu


def test_mix_phases_match_oracle(oracle, wc):
    n = 16
    cpu, gpu = pair(oracle, wc, n, mixscene.wet_sheet(), (0.2, 0.32, 0.2), (0.8, 0.6, 0.8))
    cpu.activate()
    cpu.update_liquid_phi()
    cpu.update_solid_phi()
    cpu.update_solid_weights()
    cpu.p2g()
    gpu.rebucketizeParticles()
    gpu.mapParticleNodesAPIC()
    assert np.array_equal(gpu.field("act"), cpu.field("act"))
    errs = {}
    for d in range(3):
        for name in ("psi", "sat", "smass", "svol", "sf"):
            errs[f"{name}{d}"] = rel(gpu.field(name, d), cpu.mix_field(name, d))
        for c, nm in enumerate(("ori_x", "ori_y", "ori_z")):
            errs[f"{nm}{d}"] = rel(gpu.field(nm, d), cpu.mix_field("ori", 3 * d + c))
        errs[f"vol{d}"] = rel(gpu.field("vol", d), cpu.field("vol", d))
    for name in ("psi_p", "sat_p", "pp"):
        errs[name] = rel(gpu.field(name), cpu.mix_field(name))
    worst = max(errs, key=errs.get)
    print(f"P2G fields: worst {worst} {errs[worst]:.2e}")
    assert errs[worst] < RT
    assert cpu.mix_field("psi", 1).max() > 0.1 and cpu.mix_field("pp").max() > 0
    cpu.step_velocity(DT)
    gpu.stepVelocity(DT)
    for d in range(3):
        for name in ("rhs", "ustar"):
            assert rel(gpu.field(name, d), cpu.field(name, d)) < 1e-11, (name, d)
        assert rel(gpu.field("Q", d), cpu.mix_field("Q", d)) < 1e-11
        assert rel(gpu.field("hdv", d), cpu.mix_field("hdv", d)) < 1e-11
    assert max(cpu.mix_field("hdv", d).max() for d in range(3)) > 0  # the drag acts
    cpu.assemble(DT)
    ok, it, _ = gpu.projectFine(DT, tol=1e-12, maxit=2000)
    assert ok
    assert np.array_equal(gpu.field("dof"), cpu.field("dof"))
    rs_g, col_g, val_g, b_g, p_g = gpu.system()
    rs_c, col_c, val_c, b_c, _ = cpu.system()
    assert np.array_equal(rs_g, rs_c) and np.array_equal(col_g, col_c)
    assert rel(val_g, val_c) < 1e-11 and rel(b_g, b_c) < 1e-11
    assert cpu.solve_pressure(1e-12, 2000)
    p_c = cpu.system()[4]
    assert rel(p_g, p_c) < 1e-9
    cpu.apply_pressure(DT)
    gpu.applyPressureDragFluid(DT)
    for d in range(3):
        assert np.array_equal(gpu.field("valid", d), cpu.field("valid", d))
        assert rel(gpu.field("unew", d), cpu.field("unew", d)) < 1e-9
    cpu.g2p_advect(DT)
    gpu.map_node_particles_advect(DT)
    xg, vg, _ = gpu.particles()
    xc, vc, _ = cpu.particles()
    assert np.abs(xg - xc).max() * n < 1e-10 and rel(vg, vc) < 1e-9


@pytest.mark.parametrize("mixed", [False, True])
def test_mix_trajectory_24(oracle, wc, mixed):
    n, steps = 24, 20
    cpu, gpu = pair(oracle, wc, n, mixscene.wet_sheet(), (0.2, 0.32, 0.2), (0.8, 0.7, 0.8),
                    precision=wc.WC_MIXED if mixed else None)
    for _ in range(steps):
        assert cpu.substep(DT, 1e-10, 2000)
        ok, _, _ = gpu.substep(DT, 1e-10, 2000)
        assert ok
    xg, vg, _ = gpu.particles()
    xc, vc, _ = cpu.particles()
    err = np.abs(xg - xc).max() * n
    print(f"{'mixed' if mixed else 'FP64'}: {steps} substeps at {n}^3 with the solid: position "
          f"error {err:.2e} dx, velocity {rel(vg, vc):.2e} relative")
    assert err < (1e-4 if mixed else 1e-8)


def test_mix_amg_solver_matches_oracle(oracle, wc):
    """The mixture's system solved with AMG-PCG ("solver" 1) instead of
    MIC(0)-PCG: both converge to the same pressure, so 5 substeps at 16^3
    stay within 1e-8 dx of the oracle at tol 1e-12."""
    n = 16
    cpu, gpu = pair(oracle, wc, n, mixscene.wet_sheet(), (0.2, 0.32, 0.2), (0.8, 0.6, 0.8))
    gpu.set_option("solver", 1)
    for _ in range(5):
        assert cpu.substep(DT, 1e-12, 2000)
        ok, it, _ = gpu.substep(DT, 1e-12, 2000)
        assert ok and 0 < it < 60
    err = np.abs(gpu.particles()[0] - cpu.particles()[0]).max() * n
    print(f"AMG-PCG with the solid: 5 substeps within {err:.2e} dx of the oracle")
    assert err < 1e-8


def test_mix_input_checks(wc):
    s = wc.Scene((8, 8, 8), 1.0 / 8)
    s.seed_block((0, 0, 0), (0.5, 0.5, 0.5))
    sh = mixscene.sheet(5, 0.3)
    bad = dict(sh)
    bad["x"] = sh["x"].copy()
    bad["x"][0, 0] = -0.01
    with pytest.raises(wc.WetclothError, match="outside the container"):
        s.set_elasto(*mixscene.args(bad))
    bad = dict(sh)
    bad["m"] = -sh["m"]
    with pytest.raises(wc.WetclothError, match="non-negative"):
        s.set_elasto(*mixscene.args(bad))
    with pytest.raises(wc.WetclothError, match="finite"):
        s.set_elasto(*mixscene.args(sh), prm=wc.mix_params(viscosity=float("nan")))
    s.set_elasto(*mixscene.args(sh))
    with pytest.raises(wc.WetclothError, match="liquid-only"):
        s.set_frozen(1, 0)
    assert s.substep(DT)[0]
    assert s.field("psi", 1).max() > 0
    s.set_elasto(None, None, None, None, None, None, None, None)
    with pytest.raises(wc.WetclothError, match="no solid"):
        s.field("psi", 1)
    s.set_frozen(1, 0)  # liquid only again: decisions can be recorded
    with pytest.raises(wc.WetclothError, match="adjoint"):
        s.set_elasto(*mixscene.args(sh))
    s.set_frozen(0, 0)


def test_config2_mixed_one_substep_against_fp64(wc):
    """The 128^3 wet-cloth scene after 110 mixed substeps (the column has
    reached the cloth), rounded to FP32, then one substep in each form from
    that state: positions within 1e-4 dx, and both pressures meet the
    solve's FP64 residual criterion."""
    n, dt, tol = 128, 1.0 / 960, 1e-6
    s = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED)
    s.seed_block((0.0, 0.35, 0.0), (0.4, 0.95, 1.0), seed=2)
    cloth = mixscene.config2_sheet()
    s.set_elasto(*mixscene.args(cloth))
    for _ in range(110):
        assert s.substep(dt, tol, 2000)[0]
    x, v, Cm = s.particles()
    s.close()
    wet = float((x[1] < 0.3 + 2.0 / n).mean())
    out = {}
    for prec in (wc.WC_FP64, wc.WC_MIXED):
        t = wc.Scene((n, n, n), 1.0 / n, precision=prec)
        t.set_particles(x, v, Cm)  # FP32 values, so both forms start alike
        t.set_elasto(*mixscene.args(cloth))
        ok, it, res = t.substep(dt, tol, 2000)
        assert ok
        rs, col, val, b, p = t.system()
        import scipy.sparse as sp
        A = sp.csr_matrix((val, col, rs), shape=(len(b), len(b)))
        assert np.abs(b - A @ p).max() <= tol * np.abs(b).max()
        out[prec] = t.particles()[0], it
        t.close()
    (xf, itf), (xm, itm) = out[wc.WC_FP64], out[wc.WC_MIXED]
    dxe = float(np.abs(xm - xf).max() * n)
    print(f"128^3 wet cloth, one substep after 110: {wet:.1%} of the liquid at or below "
          f"the cloth; |x_mixed - x_fp64| {dxe:.2e} dx; iterations {itm} mixed, {itf} FP64")
    assert wet > 0.01 and dxe < 1e-4


def below_sheet(wc, cloth, steps, n=128, dt=1.0 / 960):
    s = wc.Scene((n, n, n), 1.0 / n, precision=wc.WC_MIXED)
    cnt = s.seed_block((0.0, 0.35, 0.0), (0.4, 0.95, 1.0), seed=2)
    if cloth:
        s.set_elasto(*mixscene.args(mixscene.config2_sheet()))
    for k in range(steps):
        ok, it, _ = s.substep(dt, 1e-6, 2000)
        assert ok, f"substep {k}"
    x = s.particles()[0]
    assert np.isfinite(x).all() and x.min() >= -1.0 / n and x.max() <= 1.0 + 1.0 / n
    return cnt, int((x[1] < 0.3 - 2.0 / n).sum()), s


def test_config2_wet_cloth_128(wc):
    cnt, below, s = below_sheet(wc, True, 150)
    _, below_free, _ = below_sheet(wc, False, 150)
    psi = s.field("psi", 1)
    print(f"128^3 wet cloth: {cnt} liquid particles, 66049 cloth vertices, 150 substeps; "
          f"max psi {psi.max():.3f}; {below} particles below the cloth ({below_free} without it)")
    assert cnt > 4_000_000 and psi.max() > 0.01
    assert below_free > 0 and below < below_free
