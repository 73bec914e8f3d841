// Drives the C++ host mirror the way a reference user would: a
// robertbridson::PCGSolver<double> solve on a matrix assembled with
// SparseMatrix::add_to_element, AMGPCGSolveSparse on a 3-D one with the
// DOFs' cells, then liquid substeps through WetClothCore (MIC(0), then two
// with the AMG solver), then a SubstepRollout forward/backward pair whose
// gradient is compared with a central difference of replayed runs.
// Prints one JSON line.  Usage: dam_break <n> <steps> [outdir]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <array>
#include <random>
#include <vector>

#include "wetcloth_amd.hpp"

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 32;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 5;
  const char* outdir = argc > 3 ? argv[3] : nullptr;  // optional: write the particle files
  try {
    // 1. PCGSolver<double> on a 2-D Poisson matrix with a known solution
    const int m = 40, N = m * m;
    robertbridson::SparseMatrix<double> A(N);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        const int r = j * m + i;
        A.add_to_element(r, r, 4.0 + 0.01);
        if (i > 0) A.add_to_element(r, r - 1, -1.0);
        if (i + 1 < m) A.add_to_element(r, r + 1, -1.0);
        if (j > 0) A.add_to_element(r, r - m, -1.0);
        if (j + 1 < m) A.add_to_element(r, r + m, -1.0);
      }
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<double> xt(N), b(N, 0.0), x;
    for (auto& v : xt) v = nd(rng);
    for (int r = 0; r < N; ++r)
      for (size_t k = 0; k < A.index[r].size(); ++k) b[r] += A.value[r][k] * xt[A.index[r][k]];
    robertbridson::PCGSolver<double> solver;
    solver.set_solver_parameters(1e-12, 500);
    double residual = 0.0;
    int iterations = 0;
    const bool ok = solver.solve(A, b, x, residual, iterations);
    double err = 0.0, nrm = 0.0;
    for (int r = 0; r < N; ++r) {
      err = std::fmax(err, std::fabs(x[r] - xt[r]));
      nrm = std::fmax(nrm, std::fabs(xt[r]));
    }

    // 1b. AMGPCGSolveSparse on a 3-D 7-point Poisson matrix (20^3 cells, two
    // AMG levels) with a known solution
    const int q = 20, Q = q * q * q;
    robertbridson::SparseMatrix<double> A3(Q);
    std::vector<std::array<int, 3>> ijk(Q);
    for (int k = 0; k < q; ++k)
      for (int j = 0; j < q; ++j)
        for (int i = 0; i < q; ++i) {
          const int r = (k * q + j) * q + i;
          ijk[r] = {i, j, k};
          A3.add_to_element(r, r, 6.0 + 0.01);
          if (i > 0) A3.add_to_element(r, r - 1, -1.0);
          if (i + 1 < q) A3.add_to_element(r, r + 1, -1.0);
          if (j > 0) A3.add_to_element(r, r - q, -1.0);
          if (j + 1 < q) A3.add_to_element(r, r + q, -1.0);
          if (k > 0) A3.add_to_element(r, r - q * q, -1.0);
          if (k + 1 < q) A3.add_to_element(r, r + q * q, -1.0);
        }
    std::vector<double> xt3(Q), b3(Q, 0.0), x3;
    for (auto& v : xt3) v = nd(rng);
    for (int r = 0; r < Q; ++r)
      for (size_t k = 0; k < A3.index[r].size(); ++k)
        b3[r] += A3.value[r][k] * xt3[A3.index[r][k]];
    double amg_res = 0.0;
    int amg_it = 0;
    const bool amg_ok =
        AMGPCGSolveSparse(A3, b3, x3, ijk, 1e-12, 100, amg_res, amg_it, q, q, q);
    double amg_err = 0.0, amg_nrm = 0.0;
    for (int r = 0; r < Q; ++r) {
      amg_err = std::fmax(amg_err, std::fabs(x3[r] - xt3[r]));
      amg_nrm = std::fmax(amg_nrm, std::fabs(xt3[r]));
    }

    // 2. liquid substeps
    wc_desc d{};
    d.ni = d.nj = d.nk = n;
    d.dx = 1.0 / n;
    d.liquid_density = 1000.0;
    d.flip_coeff = 0.99;
    d.liquid_boundary_friction = 1.0;
    d.gravity[1] = -9.81;
    d.particle_volume = d.dx * d.dx * d.dx / 8.0;
    d.particle_mass = d.liquid_density * d.particle_volume;
    d.particle_radius = std::cbrt(3.0 * d.particle_volume / (4.0 * M_PI));
    d.solid_projection = 1;
    d.precision = WC_FP64;
    wcamd::TwoDScene scene(d);
    const double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.4, 0.6, 1.0};
    const int64_t np = scene.seedBlock(lo, hi, 0, 0.0);
    wcamd::LinearizedImplicitEuler stepper(1e-8, 1000);
    wcamd::WetClothCore core(scene, stepper);
    const auto t0 = std::chrono::steady_clock::now();
    bool all_ok = true;
    for (int s = 0; s < steps; ++s) all_ok &= core.stepSystem(1.0 / 480);
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const int last_iterations = stepper.lastIterations();
    if (outdir) {
      // the reference's output files of the frame (TwoDSceneSerializer)
      wcamd::TwoDSceneSerializer ser;
      ser.serializeFluid(scene, std::string(outdir) + "/fluid.obj");
      ser.serializePositionOnly(scene, std::string(outdir) + "/pos.bin");
    }
    std::vector<double> x4, v4, fB;
    scene.getParticles(x4, v4, fB);
    double vy = 0.0;
    for (int64_t p = 0; p < np; ++p) vy += v4[4 * p + 1];
    // 2b. the reference's three calls after the projection against the fused
    // pass, on the same grid: positions, velocities and affine matrices are
    // bit for bit the same
    bool split_fused_bitwise = false;
    {
      const double dt = 1.0 / 480;
      scene.rebucketizeParticles();
      scene.mapParticleNodesAPIC();
      stepper.stepVelocity(scene, dt);
      stepper.projectFine(scene, dt);
      stepper.applyPressureDragFluid(scene, dt);
      std::vector<double> x0, v0, B0, xs, vs, Bs, xf, vf, Bf;
      scene.getParticles(x0, v0, B0);
      scene.mapNodeParticlesAPIC();
      stepper.advectScene(scene, dt);
      scene.solidProjection(dt);
      scene.getParticles(xs, vs, Bs);
      scene.setParticles(x0, v0, B0);
      scene.mapNodeParticlesAdvect(dt);
      scene.getParticles(xf, vf, Bf);
      split_fused_bitwise = xs == xf && vs == vf && Bs == Bf;
    }
    // 2c. a long step splits into CFL substeps (WetClothCore.cpp:52-62)
    const double vmax = scene.getMaxFluidVelocity();
    const bool big_ok = core.stepSystem(0.1);
    const int big_substeps = core.lastSubsteps();
    // 3. two more substeps with the reference's own solver, AMG-PCG
    scene.setOption("solver", 1);
    bool amg_steps_ok = true;
    for (int s = 0; s < 2; ++s) amg_steps_ok &= core.stepSystem(1.0 / 480);
    const int amg_step_iterations = stepper.lastIterations();
    // 4. the rollout as a differentiable function: dL/dv0 along a random
    // direction against a central difference of two replays, and dL/d|g|
    double rollout_dir_err = 1.0, rollout_g_err = 1.0;
    bool rollout_ok = false;
    {
      wcamd::TwoDScene rs(d);
      const double rlo[3] = {0.0, 0.0, 0.0}, rhi[3] = {0.5, 0.5, 0.5};
      rs.seedBlock(rlo, rhi, 3, 0.5);
      std::vector<double> x0, v0, C0, x, v, C;
      rs.getParticlesSoA(x0, v0, C0);
      const int rsteps = 6;
      const double rdt = 1.0 / 480;
      wcamd::SubstepRollout::Context rc;
      rollout_ok = wcamd::SubstepRollout::forward(rc, rs, rsteps, rdt, x0, v0, C0, x, v, C,
                                                  1e-12, 5000, 4);
      std::vector<double> gx(x.size());
      for (size_t i = 0; i < x.size(); ++i) gx[i] = 2.0 * (x[i] - x0[i]);  // L = |x - x0|^2
      const wcamd::SubstepRollout::Grads gr = wcamd::SubstepRollout::backward(rc, gx, {}, {});
      rollout_ok = rollout_ok && gr.converged;
      auto loss = [&](const std::vector<double>& vv) {
        std::vector<double> xe, ve, Ce;
        wcamd::SubstepRollout::replay(rc, x0, vv, C0, xe, ve, Ce);
        double l = 0.0;
        for (size_t i = 0; i < xe.size(); ++i) l += (xe[i] - x0[i]) * (xe[i] - x0[i]);
        return l;
      };
      const double eps = 1e-4;
      std::vector<double> dir(v0.size()), vp(v0), vm(v0);
      double ad = 0.0;
      for (size_t i = 0; i < dir.size(); ++i) {
        dir[i] = nd(rng);
        vp[i] += eps * dir[i];
        vm[i] -= eps * dir[i];
        ad += gr.v[i] * dir[i];
      }
      const double fd = (loss(vp) - loss(vm)) / (2 * eps);
      rollout_dir_err = std::fabs(fd - ad) / std::fabs(fd);
      const double gp[3] = {0.0, -(9.81 + eps), 0.0}, gm[3] = {0.0, -(9.81 - eps), 0.0},
                   g0[3] = {0.0, -9.81, 0.0};
      rs.setGravity(gp);
      const double lp = loss(v0);
      rs.setGravity(gm);
      const double lm = loss(v0);
      rs.setGravity(g0);
      const double fdg = (lp - lm) / (2 * eps);
      rollout_g_err = std::fabs(fdg - gr.gravity) / std::fabs(fdg);
    }
    // 5. the block over a resting porous sheet (wc_set_elasto_particles):
    // 25 x 25 vertices at y = 0.3, 2 mm thick, 0.2 kg/m^2, rest volume
    // fraction 0.5, orientation along the normal, shape factor 0.25; three
    // substeps, the same again without the sheet
    bool mix_ok = false;
    double mix_vy = 0.0, free_vy = 0.0;
    {
      const int nv = 25;
      const int64_t ne = (int64_t)nv * nv;
      std::vector<double> ex(3 * ne), em(ne), efm(ne, 0.0), efv(ne, 0.0), erv(ne), erf(ne, 0.5),
          eo(3 * ne, 0.0), esf(ne, 0.25);
      const double h = 1.0 / (nv - 1);
      for (int i = 0; i < nv; ++i)
        for (int k = 0; k < nv; ++k) {
          const int64_t p = (int64_t)i * nv + k;
          ex[p] = i * h;
          ex[ne + p] = 0.3;
          ex[2 * ne + p] = k * h;
          const double share = (i == 0 || i == nv - 1 ? 0.5 : 1.0) * (k == 0 || k == nv - 1 ? 0.5 : 1.0);
          em[p] = 0.2 * share * h * h;
          erv[p] = 2e-3 * share * h * h;
          eo[ne + p] = 1.0;
        }
      const wc_mix_params prm = {1, 1, 1e-4, 8.9e-4, 1.6, 5e-5, 0.0728, 40.8 / 180.0 * M_PI};
      auto run = [&](bool sheet, double& mean_vy) {
        wcamd::TwoDScene ms(d);
        const int64_t mn = ms.seedBlock(lo, hi, 0, 0.0);
        if (sheet) ms.setElastoParticles(ex, em, efm, efv, erv, erf, eo, esf, prm);
        wcamd::LinearizedImplicitEuler mst(1e-8, 1000);
        wcamd::WetClothCore mc(ms, mst);
        bool okm = true;
        for (int s = 0; s < 3; ++s) okm &= mc.stepSystem(1.0 / 480);
        std::vector<double> mx, mv, mC;
        ms.getParticlesSoA(mx, mv, mC);
        mean_vy = 0.0;
        for (int64_t p = 0; p < mn; ++p) mean_vy += mv[mn + p];
        mean_vy /= (double)mn;
        return okm;
      };
      mix_ok = run(true, mix_vy) && run(false, free_vy);
    }
    std::printf(
        "{\"pcg_ok\": %s, \"pcg_iterations\": %d, \"pcg_residual\": %.17g, "
        "\"pcg_max_error\": %.17g, \"amg_ok\": %s, \"amg_iterations\": %d, "
        "\"amg_max_error\": %.17g, \"particles\": %lld, \"steps_ok\": %s, "
        "\"last_iterations\": %d, \"mean_vy\": %.17g, \"seconds\": %.6f, "
        "\"amg_steps_ok\": %s, \"amg_step_iterations\": %d, "
        "\"split_fused_bitwise\": %s, \"vmax\": %.17g, \"big_step_ok\": %s, "
        "\"big_step_substeps\": %d, \"rollout_ok\": %s, \"rollout_dir_err\": %.6g, "
        "\"rollout_g_err\": %.6g, \"mix_ok\": %s, \"mix_vy\": %.17g, \"free_vy\": %.17g}\n",
        ok ? "true" : "false", iterations, residual, err / nrm, amg_ok ? "true" : "false",
        amg_it, amg_err / amg_nrm, (long long)np,
        all_ok ? "true" : "false", last_iterations, vy / (double)np, secs,
        amg_steps_ok ? "true" : "false", amg_step_iterations,
        split_fused_bitwise ? "true" : "false", vmax, big_ok ? "true" : "false", big_substeps,
        rollout_ok ? "true" : "false", rollout_dir_err, rollout_g_err, mix_ok ? "true" : "false",
        mix_vy, free_vy);
  } catch (const wcamd::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
