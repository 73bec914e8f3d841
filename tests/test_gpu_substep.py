"""The GPU substep against the FP64 oracle, phase by phase, through the C ABI.

Integer and index results (bucket activation, the DOF map, the matrix
structure) must be identical.  Floating-point grid quantities are sums in a
different order (P2G scatters with atomics), so they are compared with a
relative tolerance of 1e-12; after the pressure solve (1e-8 relative
residual) velocities and positions are compared with the tolerances stated
in each test.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 1.0 / 480
RT = 1e-12  # FP64 grid quantities summed in a different order


def rel(a, b):
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


def make_pair(oracle, wc, n, sigma=0.0, seed=0, block_hi=(0.4, 0.6, 1.0)):
    g, x, v = oracle.dam_break(n, block_hi=block_hi, seed=seed, sigma=sigma)
    cpu = oracle.OracleScene((n, n, n), 1.0 / n, x, v)
    gpu = wc.Scene((n, n, n), 1.0 / n)
    gpu.set_particles(x, v)
    return cpu, gpu, x, v


@pytest.mark.parametrize("n", [16, 24])
def test_phases_match_oracle(oracle, wc, n):
    cpu, gpu, x, v = make_pair(oracle, wc, n, sigma=0.3, seed=1)
    # --- P2G and grid setup ---
    cpu.activate()
    cpu.update_liquid_phi()
    cpu.update_solid_phi()
    cpu.update_solid_weights()
    cpu.p2g()
    gpu.rebucketizeParticles()
    gpu.mapParticleNodesAPIC()
    assert np.array_equal(gpu.field("act"), cpu.field("act"))
    assert np.array_equal(gpu.field("lphi"), cpu.field("lphi"))
    for d in range(3):
        assert np.array_equal(gpu.field("w", d), cpu.field("w", d))
        assert rel(gpu.field("mass", d), cpu.field("mass", d)) < RT
        assert rel(gpu.field("vol", d), cpu.field("vol", d)) < RT
        assert rel(gpu.field("vel", d), cpu.field("vel", d)) < 1e-11
    # --- grid update ---
    cpu.step_velocity(DT)
    gpu.stepVelocity(DT)
    for d in range(3):
        assert rel(gpu.field("rhs", d), cpu.field("rhs", d)) < 1e-11
        assert rel(gpu.field("ustar", d), cpu.field("ustar", d)) < 1e-11
    # --- assembly ---
    cpu.assemble(DT)
    ok, it, res = gpu.projectFine(DT, tol=1e-8, maxit=1000)
    assert ok
    assert np.array_equal(gpu.field("dof"), cpu.field("dof"))
    rs_g, col_g, val_g, b_g, p_g = gpu.system()
    rs_c, col_c, val_c, b_c, _ = cpu.system()
    assert np.array_equal(rs_g, rs_c) and np.array_equal(col_g, col_c)
    assert rel(val_g, val_c) < 1e-11
    assert np.abs(b_g - b_c).max() <= 1e-11 * np.abs(b_c).max()
    # --- solve ---
    cpu.solve_pressure(1e-8, 1000)
    okc, itc, resc = cpu.last_solve
    _, _, _, _, p_c = cpu.system()
    assert okc and abs(it - itc) <= 1
    assert rel(p_g, p_c) < 1e-7
    # --- pressure application, extrapolation, constraint ---
    cpu.apply_pressure(DT)
    gpu.applyPressureDragFluid(DT)
    for d in range(3):
        assert np.array_equal(gpu.field("valid", d), cpu.field("valid", d))
        assert rel(gpu.field("unew", d), cpu.field("unew", d)) < 1e-7
    # --- G2P, advection, solid projection: the reference's three calls ---
    cpu.g2p_advect(DT)
    gpu.mapNodeParticlesAPIC()
    gpu.advectScene(DT)
    gpu.solidProjection(DT)
    xg, vg, Cg = gpu.particles()
    xc, vc, Cc = cpu.particles()
    dx = 1.0 / n
    assert np.abs(xg - xc).max() < 1e-9 * dx
    assert rel(vg, vc) < 1e-7
    assert rel(Cg, Cc) < 1e-6


def test_trajectory_32(oracle, wc):
    """20 substeps at 32^3: positions stay within 1e-4 dx of the oracle."""
    n = 32
    cpu, gpu, x, v = make_pair(oracle, wc, n)
    dt = 1.0 / 480
    for _ in range(20):
        cpu.substep(dt, 1e-8, 1000)
        ok, it, res = gpu.step(dt, 1e-8, 1000)
        assert ok
    xg, vg, _ = gpu.particles()
    xc, vc, _ = cpu.particles()
    dx = 1.0 / n
    err = np.abs(xg - xc).max() / dx
    print(f"max position error {err:.3e} dx, velocity rel {rel(vg, vc):.3e}")
    assert err < 1e-4


def test_seeding_matches_oracle(oracle, wc):
    n = 32
    g, x, v = oracle.dam_break(n, seed=3, sigma=0.5)
    gpu = wc.Scene((n, n, n), 1.0 / n)
    cnt = gpu.seed_block((0.0, 0.0, 0.0), (0.4, 0.6, 1.0), seed=3, sigma=0.5)
    assert cnt == x.shape[1]
    xg, vg, Cg = gpu.particles()
    assert np.array_equal(xg, x)
    assert np.array_equal(vg, v)
    assert not Cg.any()


def test_reference_layout_roundtrip(wc):
    n = 8
    rng = np.random.default_rng(0)
    npart = 100
    x4 = rng.uniform(0.2, 0.8, 4 * npart)
    v4 = rng.standard_normal(4 * npart)
    fB = rng.standard_normal(9 * npart)  # Eigen 3N x 3, column-major
    s = wc.Scene((n, n, n), 1.0 / n)
    s.set_particles_reference_layout(x4, v4, fB)
    s.rebucketizeParticles()  # reorders on the device; the ABI restores the order
    x4b, v4b, fBb = s.particles_reference_layout()
    m = np.arange(4 * npart) % 4 != 3
    assert np.array_equal(x4b[m], x4[m]) and np.array_equal(v4b[m], v4[m])
    assert np.array_equal(fBb, fB)
    xs, vs, Cs = s.particles()
    # SoA view: fB(r, c) of particle p sits at column-major (c*3N + 3p + r)
    assert np.array_equal(Cs[1 * 3 + 2], fB[2 * 3 * npart + 3 * np.arange(npart) + 1])


def test_matches_committed_golden(wc):
    """The GPU path against the committed oracle fixture (tests/golden)."""
    import os
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "dam_break_12.npz"))
    n = int(ref["n"])
    dt = float(ref["dt"])
    tol = float(ref["tol"])
    s = wc.Scene((n, n, n), 1.0 / n)
    s.set_particles(ref["x0"], ref["v0"])
    s.rebucketizeParticles()
    s.mapParticleNodesAPIC()
    s.stepVelocity(dt)
    ok, it, res = s.projectFine(dt, tol, 2000)
    assert ok
    rs, col, val, b, p = s.system()
    assert np.array_equal(s.field("dof"), ref["dof"])
    assert np.array_equal(rs, ref["rowstart"]) and np.array_equal(col, ref["col"])
    assert rel(val, ref["val"]) < 1e-11
    assert np.abs(b - ref["rhs"]).max() <= 1e-11 * np.abs(ref["rhs"]).max()
    assert rel(p, ref["pressure"]) < 1e-8
    s.applyPressureDragFluid(dt)
    s.map_node_particles_advect(dt)
    for _ in range(2):
        ok, it, res = s.step(dt, tol, 2000)
        assert ok
    x3, v3, _ = s.particles()
    assert np.abs(x3 - ref["x3"]).max() < 1e-8 / n
    assert rel(v3, ref["v3"]) < 1e-6


def test_cfl_substep_split(oracle, wc):
    """WetClothCore::stepSystem's split (WetClothCore.cpp:52-62): with v_max =
    getMaxFluidVelocity, dt runs as max(1, ceil(dt / min(3 dx / v_max,
    1/30))) equal substeps; the result is the oracle's over that many
    substeps of dt / n."""
    n = 16
    dx = 1.0 / n
    g, x, v = oracle.dam_break(n, block_hi=(0.4, 0.6, 1.0), seed=3, sigma=0.3)
    v[0] += 2.5  # fast enough that 3 dx / v_max < 1/30
    gpu = wc.Scene((n, n, n), dx)
    gpu.set_particles(x, v)
    vmax = gpu.getMaxFluidVelocity()
    assert vmax == np.sqrt((v * v).sum(axis=0)).max()
    dt = 0.05
    expect = max(1, int(np.ceil(dt / min(min(dx / 1e-63 / 3.0, 1.0 / 30.0),
                                         min(dx / max(1e-63, vmax) * 3.0, 1.0 / 30.0)))))
    assert expect >= 2
    ok, it, res = gpu.step(dt, 1e-8, 2000)
    assert ok and gpu.step_substeps() == expect
    cpu = oracle.OracleScene((n, n, n), dx, x, v)
    for _ in range(expect):
        assert cpu.substep(dt / expect, 1e-8, 2000)
    xg, vg, _ = gpu.particles()
    xc, vc, _ = cpu.particles()
    assert np.abs(xg - xc).max() < 1e-8 * dx
    assert rel(vg, vc) < 1e-6
    # a scene at rest takes the 1/30 s bound: 0.05 s is two substeps
    gpu2 = wc.Scene((n, n, n), dx)
    gpu2.set_particles(x)
    assert gpu2.getMaxFluidVelocity() == 0.0
    assert gpu2.step(0.05, 1e-8, 2000)[0] and gpu2.step_substeps() == 2
