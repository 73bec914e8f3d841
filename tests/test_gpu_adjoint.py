"""The adjoint of the substep on the GPU.

* Against the oracle's adjoint (wco_substep_adjoint, FP64 on the host): the
  same particles stepped by both in record mode, then random gradients of
  the substep's outputs taken back through both; three substeps of a 16^3
  dam break with a seeded velocity field.
* Checkpoints: a rollout's gradient does not depend on the checkpoint
  spacing (every 4 substeps against one segment).
* The differentiation workload: the 64^3 dam break (block [0,0.4] x [0,0.6]
  x [0,1], 8 particles per cell, zero initial velocity) over 50 substeps of
  dt = 1/480 with checkpoints every 10; loss sum |x - x_target|^2 with the
  target the same run from a +0.1 m/s uniform x-velocity.  Gradients with
  respect to the initial velocities (64 seeded random unit directions) and
  the gravity magnitude against central differences of the FP64 GPU forward
  with eps = 1e-4, the decisions of the recorded run replayed (and no
  replay departing from them: a uniform change of |g| by 1e-4 moves every
  particle, and the difference is taken again at 1e-5 or 1e-6 where that
  departs), within 1e-4 relative.  The difference of the two losses is
  formed per particle, (x+ - x-) . (x+ + x- - 2 x_t), so that it does not
  cancel between two sums of ~50.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_substep_adjoint_matches_oracle(oracle, wc):
    n, dt = 16, 1.0 / 480
    _, x, v = oracle.dam_break(n, block_hi=(0.5, 0.5, 0.5), seed=3, sigma=0.5)
    C0 = np.zeros((9, x.shape[1]))
    cpu = oracle.OracleScene((n, n, n), 1.0 / n, x, v, C0)
    gpu = wc.Scene((n, n, n), 1.0 / n)
    gpu.set_particles(x, v, C0)
    cpu.frozen_mode(1)
    rng = np.random.default_rng(7)
    worst = 0.0
    for step in range(3):
        xs, vs, Cs = cpu.particles()
        assert cpu.substep(dt, TOL, 5000)
        gpu.set_frozen(1, step)
        ok, _, _ = gpu.substep(dt, TOL, 5000)
        assert ok
        xg, _, _ = gpu.particles()
        assert np.abs(xg - cpu.particles()[0]).max() * n < 1e-9
        xb = rng.standard_normal(xs.shape)
        vb = rng.standard_normal(vs.shape)
        Cb = rng.standard_normal(Cs.shape)
        okc, xc, vc, Cc, gc = cpu.substep_adjoint(dt, TOL, 5000, xs, vs, Cs, xb, vb, Cb)
        okg, xa, va, Ca, ga = gpu.substep_adjoint(dt, xb, vb, Cb, tol=TOL)
        assert okc and okg
        errs = (rel(xa, xc), rel(va, vc), rel(Ca, Cc), abs(ga - gc) / abs(gc))
        print(f"substep {step}: adjoint vs oracle x {errs[0]:.2e} v {errs[1]:.2e} "
              f"C {errs[2]:.2e} |g| {errs[3]:.2e}")
        worst = max(worst, *errs)
    assert worst < 1e-8


def test_adjoint_needs_a_recorded_substep(wc):
    s = wc.Scene((8, 8, 8), 1.0 / 8)
    s.seed_block((0, 0, 0), (0.5, 0.5, 0.5))
    s.substep(1.0 / 480)
    with pytest.raises(wc.WetclothError, match="recorded or replayed"):
        s.substep_adjoint(1.0 / 480, np.zeros((3, s.num_particles())))
    with pytest.raises(wc.WetclothError, match="no recorded substep"):
        s.set_frozen(2, 5)
    m = wc.Scene((8, 8, 8), 1.0 / 8, precision=wc.WC_MIXED)
    with pytest.raises(wc.WetclothError, match="WC_FP64"):
        m.set_frozen(1, 0)


def rollout_grad(wc, n, x0, v0, steps, dt, every, xt):
    s = wc.Scene((n, n, n), 1.0 / n)
    s.set_particles(x0, v0)
    ok, _ = s.rollout_forward(steps, dt, tol=TOL, every=every)
    assert ok
    xf = s.particles()[0]
    ok, gx, gv, gC, gg = s.rollout_backward(dt, 2.0 * (xf - xt), tol=TOL)
    assert ok
    return s, xf, gx, gv, gg


def test_rollout_gradient_independent_of_checkpoint_spacing(oracle, wc):
    n, dt, steps = 24, 1.0 / 480, 12
    _, x0, v0 = oracle.dam_break(n, seed=5, sigma=0.5)
    xt = x0 + 0.01
    _, _, gx1, gv1, gg1 = rollout_grad(wc, n, x0, v0, steps, dt, 4, xt)
    _, _, gx2, gv2, gg2 = rollout_grad(wc, n, x0, v0, steps, dt, steps, xt)
    print(f"every 4 vs one segment: dL/dx0 {rel(gx1, gx2):.2e}, dL/dv0 {rel(gv1, gv2):.2e}, "
          f"dL/d|g| {abs(gg1 - gg2) / abs(gg2):.2e}")
    assert rel(gx1, gx2) < 1e-10 and rel(gv1, gv2) < 1e-10
    assert abs(gg1 - gg2) <= 1e-10 * abs(gg2)


def test_config3_gradients_against_central_differences(wc):
    n, dt, steps, every, eps, ndir = 64, 1.0 / 480, 50, 10, 1e-4, 64
    seed = wc.Scene((n, n, n), 1.0 / n)
    seed.seed_block((0.0, 0.0, 0.0), (0.4, 0.6, 1.0), seed=3)
    x0, v0, _ = seed.particles()
    seed.close()
    v0 = v0.astype(np.float32).astype(np.float64)  # FP32 initial velocities (zero here)
    # the target: the same run from a +0.1 m/s x-velocity, its own decisions
    tg = wc.Scene((n, n, n), 1.0 / n)
    tg.set_particles(x0, v0 + np.array([[0.1], [0.0], [0.0]]))
    for _ in range(steps):
        assert tg.substep(dt, TOL, 5000)[0]
    xt = tg.particles()[0]
    tg.close()
    s, xf, _, gv, gg = rollout_grad(wc, n, x0, v0, steps, dt, every, xt)
    loss0 = float(((xf - xt) ** 2).sum())
    print(f"  forward/backward done: loss {loss0:.6e}", flush=True)

    def final(v, g=9.81):
        """The replayed run's final positions and its departures from the
        recorded decisions."""
        s.set_gravity((0.0, -g, 0.0))
        s.set_particles(x0, v)
        ok, _ = s.rollout_forward(steps, dt, tol=TOL, replay=True)
        dep = s.frozen_departures()
        assert ok or dep > 0
        return s.particles()[0], dep

    def central(xp, xm, h):
        # (L(+) - L(-)) / 2h with the difference formed per particle,
        # (x+ - x-) . (x+ + x- - 2 x_t), instead of from the two sums of ~50
        return float(((xp - xm) * (xp + xm - 2.0 * xt)).sum()) / (2 * h)

    x1, dep1 = final(v0)
    x2, _ = final(v0)
    l1 = float(((x1 - xt) ** 2).sum())
    assert dep1 == 0 and abs(l1 - loss0) <= 1e-12 * loss0  # the replay is the recorded run
    noise = float(np.abs(x1 - x2).max())
    # unit directions: the perturbation eps d moves each particle by ~1e-8 m
    # over the run, so the replays stay in the record's smooth neighbourhood
    # (a direction whose replay departs from a recorded decision -- a face
    # losing its last particle at the edge of its support -- is not a valid
    # central difference and is drawn again)
    rng = np.random.default_rng(64)
    errs, skipped = [], 0
    while len(errs) < ndir and skipped < ndir // 2:
        d = rng.standard_normal(v0.shape).astype(np.float32).astype(np.float64)
        d /= np.linalg.norm(d)
        (xp, dp), (xm, dm) = final(v0 + eps * d), final(v0 - eps * d)
        if dp or dm:
            skipped += 1
            continue
        fd = central(xp, xm, eps)
        errs.append(abs(fd - float((gv * d).sum())) / abs(fd))
        if len(errs) % 8 == 0:
            print(f"  direction {len(errs)}/{ndir}: |dL/dv0 . d| {abs(fd):.3e}, worst relative "
                  f"error so far {max(errs):.2e}", flush=True)
    # |g| moves every particle by ~0.5 eps t^2; where that departs from a
    # recorded decision, the difference is taken again with a smaller eps
    for eps_g in (eps, eps / 10, eps / 100):
        (xp, dp), (xm, dm) = final(v0, 9.81 + eps_g), final(v0, 9.81 - eps_g)
        fdg = central(xp, xm, eps_g)
        eg = abs(fdg - gg) / abs(fdg)
        print(f"  dL/d|g| with eps {eps_g:g}: {fdg:.8e} ({eg:.2e}), {dp + dm} departures",
              flush=True)
        if dp + dm == 0:
            break
    print(f"64^3, {x0.shape[1]} particles, {steps} substeps (checkpoints every {every}): "
          f"replay noise {noise:.1e} m in the positions; dL/dv0 along {len(errs)} unit "
          f"directions vs "
          f"central differences (eps {eps:g}): median {np.median(errs):.2e}, worst "
          f"{max(errs):.2e}, {skipped} directions drawn again; dL/d|g| {gg:.8e} vs "
          f"{fdg:.8e} ({eg:.2e}, eps {eps_g:g})")
    assert len(errs) == ndir and max(errs) < 1e-4
    assert dp + dm == 0 and eg < 1e-4
