// C++ host mirror of the reference interfaces on the liquid substep path,
// implemented on the C ABI of libwcgpu.so (include/wetcloth_amd.h).  Header
// only; link with -lwcgpu.
//
//  * robertbridson::SparseMatrix<double> / PCGSolver<double>
//    (Core/PCGSolver/SparseMatrix.h:24-167, PCGSolver.h:245-330): same names,
//    same argument meaning, same return convention (solve() returns false on
//    non-convergence, on rho == 0 or NaN, and the residual/iteration outputs
//    are filled as the reference fills them).
//  * wcamd::TwoDScene / LinearizedImplicitEuler / WetClothCore: the phase calls
//    of one liquid substep (Core/WetClothCore.cpp:77-300) on a scene whose
//    particles and grids stay resident on the GPU.
//  * wcamd::TwoDSceneSerializer: the reference's particle files
//    (App/TwoDSceneSerializer.cpp) written from that scene.
//  * wcamd::SubstepRollout: substeps as a differentiable function with a
//    forward/backward pair and checkpoints (no reference counterpart).
//
// Errors of the library (bad arguments, HIP failures) throw wcamd::Error; the
// reference has no exceptions because it has no failure modes of this kind.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/wetcloth_amd.h"

namespace wcamd {

struct Error : std::runtime_error {
  wc_status status;
  Error(wc_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(wc_status st, const wc_ctx* ctx) {
  if (st == WC_OK || st == WC_NOT_CONVERGED) return;
  throw Error(st, std::string("wetcloth_amd: ") + wc_last_error(ctx));
}

// ---------------------------------------------------------------- scene

class TwoDScene {
 public:
  explicit TwoDScene(const wc_desc& d) : desc_(d) {
    wc_ctx* c = nullptr;
    check(wc_create(&d, &c), nullptr);
    ctx_ = c;
  }
  ~TwoDScene() { wc_destroy(ctx_); }
  TwoDScene(const TwoDScene&) = delete;
  TwoDScene& operator=(const TwoDScene&) = delete;

  // m_x, m_fluid_v (4 doubles per particle) and m_fB (3N x 3 column-major)
  void setParticles(const std::vector<double>& x4, const std::vector<double>& fluid_v4,
                    const std::vector<double>& fB) {
    const int64_t n = (int64_t)x4.size() / 4;
    check(wc_set_particles(ctx_, n, x4.data(), fluid_v4.empty() ? nullptr : fluid_v4.data(),
                           fB.empty() ? nullptr : fB.data()),
          ctx_);
  }
  void getParticles(std::vector<double>& x4, std::vector<double>& fluid_v4,
                    std::vector<double>& fB) const {
    int64_t n = 0;
    check(wc_num_particles(ctx_, &n), ctx_);
    x4.assign(4 * n, 0.0);
    fluid_v4.assign(4 * n, 0.0);
    fB.assign(9 * n, 0.0);
    check(wc_get_particles(ctx_, x4.data(), fluid_v4.data(), fB.data()), ctx_);
  }
  // structure of arrays in the caller's order: x[d*n+p], v[d*n+p],
  // C[(r*3+c)*n+p] (C may be empty: zero)
  void setParticlesSoA(const std::vector<double>& x, const std::vector<double>& v,
                       const std::vector<double>& C) {
    const int64_t n = (int64_t)x.size() / 3;
    check(wc_set_particles_soa(ctx_, n, x.data(), v.empty() ? nullptr : v.data(),
                               C.empty() ? nullptr : C.data()),
          ctx_);
  }
  void getParticlesSoA(std::vector<double>& x, std::vector<double>& v,
                       std::vector<double>& C) const {
    const int64_t n = getNumParticles();
    x.assign(3 * n, 0.0);
    v.assign(3 * n, 0.0);
    C.assign(9 * n, 0.0);
    check(wc_get_particles_soa(ctx_, x.data(), v.data(), C.data()), ctx_);
  }
  void setGravity(const double g[3]) { check(wc_set_gravity(ctx_, g), ctx_); }
  // the elasto particles of a resting, unsolved solid (the cloth's
  // vertices, LiquidInfo::solve_solid = 0): positions and orientations
  // 3*ne (x[d*ne+p]), the rest ne each; an empty x removes the solid
  void setElastoParticles(const std::vector<double>& x, const std::vector<double>& m,
                          const std::vector<double>& fluid_m,
                          const std::vector<double>& fluid_vol,
                          const std::vector<double>& rest_vol,
                          const std::vector<double>& rest_volume_fraction,
                          const std::vector<double>& orientation,
                          const std::vector<double>& shape_factor, const wc_mix_params& prm) {
    const int64_t ne = (int64_t)x.size() / 3;
    const size_t sizes[7] = {m.size(), fluid_m.size(), fluid_vol.size(), rest_vol.size(),
                             rest_volume_fraction.size(), shape_factor.size(),
                             orientation.size() / 3};
    for (size_t k : sizes)
      if ((int64_t)k != ne) throw Error(WC_ERR_ARG, "wetcloth_amd: solid array sizes differ");
    check(wc_set_elasto_particles(ctx_, ne, x.data(), m.data(), fluid_m.data(), fluid_vol.data(),
                                  rest_vol.data(), rest_volume_fraction.data(),
                                  orientation.data(), shape_factor.data(), &prm),
          ctx_);
  }
  int64_t getNumParticles() const {
    int64_t n = 0;
    check(wc_num_particles(ctx_, &n), ctx_);
    return n;
  }
  int64_t seedBlock(const double lo[3], const double hi[3], uint64_t seed, double sigma) {
    int64_t n = 0;
    check(wc_seed_block(ctx_, lo, hi, seed, sigma, &n), ctx_);
    return n;
  }

  // TwoDScene.cpp:1588-1708
  void rebucketizeParticles() { check(wc_rebucketize(ctx_), ctx_); }
  // resampleNodes activation, updateLiquidPhi, updateSolidPhi, updateSolidWeights
  // and TwoDScene::mapParticleNodesAPIC (TwoDScene.h:487)
  void mapParticleNodesAPIC() { check(wc_map_particle_nodes(ctx_), ctx_); }
  // TwoDScene::mapNodeParticlesAPIC() (TwoDScene.h:493): v and C from the grid
  void mapNodeParticlesAPIC() { check(wc_map_node_particles(ctx_), ctx_); }
  // TwoDScene::solidProjection(dt) (TwoDScene.cpp:1776-1822)
  void solidProjection(double dt) { check(wc_solid_projection(ctx_, dt), ctx_); }
  // mapNodeParticlesAPIC, SceneStepper::advectScene and solidProjection in one
  // pass over the particles (bitwise the same as the three calls)
  void mapNodeParticlesAdvect(double dt) { check(wc_map_node_particles_advect(ctx_, dt), ctx_); }
  // TwoDScene::getMaxFluidVelocity (TwoDScene.cpp:4180-4188)
  double getMaxFluidVelocity() const {
    double v = 0.0;
    check(wc_max_fluid_velocity(ctx_, &v), ctx_);
    return v;
  }
  double getCellSize() const { return desc_.dx; }

  // wc_set_option: "solver" 1 makes projectFine solve with AMGPCGSolveSparse
  // (the reference's call at Pressure.cpp:797), 0 with the MIC(0) PCG
  void setOption(const char* name, int64_t value) { check(wc_set_option(ctx_, name, value), ctx_); }

  wc_ctx* handle() const { return ctx_; }
  const wc_desc& desc() const { return desc_; }

 private:
  wc_desc desc_;
  wc_ctx* ctx_ = nullptr;
};

// LinearizedImplicitEuler's liquid phases (LinearizedImplicitEuler.cpp)
class LinearizedImplicitEuler {
 public:
  LinearizedImplicitEuler(double pressure_criterion = 1e-6, int maxiters = 1000)
      : criterion_(pressure_criterion), maxiters_(maxiters) {}
  bool stepVelocity(TwoDScene& s, double dt) {
    check(wc_step_velocity(s.handle(), dt), s.handle());
    return true;
  }
  // projectFine -> pressure::solveNodePressure with the MIC(0) PCG solve, or
  // AMG-PCG after setOption("solver", 1); like the reference, a failed solve
  // is reported and the substep goes on
  bool projectFine(TwoDScene& s, double dt) {
    const wc_status st = wc_project(s.handle(), dt, criterion_, maxiters_, &residual_, &iterations_);
    check(st, s.handle());
    return st == WC_OK;
  }
  bool applyPressureDragFluid(TwoDScene& s, double dt) {
    check(wc_apply_pressure(s.handle(), dt), s.handle());
    return true;
  }
  // SceneStepper::advectScene (SceneStepper.cpp:17-41)
  bool advectScene(TwoDScene& s, double dt) {
    check(wc_advect_scene(s.handle(), dt), s.handle());
    return true;
  }
  double lastResidual() const { return residual_; }
  int lastIterations() const { return iterations_; }

 private:
  double criterion_;
  int maxiters_;
  double residual_ = 0.0;
  int32_t iterations_ = 0;
};

// TwoDSceneSerializer's particle files (App/TwoDSceneSerializer.cpp) for a
// liquid scene: serializePositionOnly (:161-170), loadPosOnly (:172-183, here
// from a file name rather than an open stream) and the fluid file of
// serializeScene (:135-159; the cloth, hair, boundary and spring files of a
// liquid-only scene are not written).  Files are complete when the call
// returns (the reference writes from a detached thread).
class TwoDSceneSerializer {
 public:
  void serializePositionOnly(TwoDScene& s, const std::string& fn_pos) {
    check(wc_serialize_position_only(s.handle(), fn_pos.c_str()), s.handle());
  }
  void loadPosOnly(TwoDScene& s, const std::string& fn_pos) {
    check(wc_load_position_only(s.handle(), fn_pos.c_str()), s.handle());
  }
  void serializeFluid(TwoDScene& s, const std::string& fn_fluid) {
    check(wc_serialize_fluid(s.handle(), fn_fluid.c_str()), s.handle());
  }
};

// WetClothCore::stepSystem (WetClothCore.cpp:45-330) for liquid scenes:
// the CFL split of dt (:52-62) and, per substep, the liquid path's calls in
// the reference's order (:115-262).  With fused set, the transfer back,
// advection and projection run as one pass (the same result, bit for bit).
class WetClothCore {
 public:
  WetClothCore(TwoDScene& scene, LinearizedImplicitEuler& stepper, bool fused = false)
      : scene_(scene), stepper_(stepper), fused_(fused) {}
  bool stepSystem(double dt) {
    const double max_elasto_vel = 0.0;  // no elastic particles, or a resting solid
    const double max_fluid_vel = scene_.getMaxFluidVelocity();
    const double dx = scene_.getCellSize();
    const double max_elasto_dt = std::min(dx / std::max(1e-63, max_elasto_vel) / 3.0, 1.0 / 30.0);
    const double max_fluid_dt = std::min(dx / std::max(1e-63, max_fluid_vel) * 3.0, 1.0 / 30.0);
    const double max_dt = std::min(max_elasto_dt, max_fluid_dt);
    const int num_substeps = std::max(1, (int)std::ceil(dt / max_dt));
    const double sub_dt = dt / (double)num_substeps;
    last_substeps_ = num_substeps;
    bool ok = true;
    for (int k = 0; k < num_substeps; ++k) {
      scene_.rebucketizeParticles();
      scene_.mapParticleNodesAPIC();
      stepper_.stepVelocity(scene_, sub_dt);
      ok &= stepper_.projectFine(scene_, sub_dt);
      stepper_.applyPressureDragFluid(scene_, sub_dt);
      if (fused_) {
        scene_.mapNodeParticlesAdvect(sub_dt);
      } else {
        scene_.mapNodeParticlesAPIC();
        stepper_.advectScene(scene_, sub_dt);
        if (scene_.desc().solid_projection) scene_.solidProjection(sub_dt);
      }
    }
    return ok;
  }
  int lastSubsteps() const { return last_substeps_; }

 private:
  TwoDScene& scene_;
  LinearizedImplicitEuler& stepper_;
  bool fused_;
  int last_substeps_ = 0;
};

// ------------------------------------------------------------- the adjoint

// `steps` liquid substeps of dt as a differentiable function of the initial
// particle state and the gravity magnitude (no reference counterpart: the
// reference has no differentiation), shaped like a torch.autograd.Function:
// forward(ctx, ...) runs the substeps on the GPU, recording each one's
// discrete decisions and checkpointing the particles before every
// `every`-th; backward(ctx, ...) takes dL/d(final x, v, C) to dL/d(initial
// x, v, C) and dL/d|g| (gravity (0, -|g|, 0)), recomputing each segment from
// its checkpoint.  replay(ctx, ...) runs the recorded decisions again from
// other initial particles (the same smooth map, for central differences).
// FP64 scenes only; vectors are structure-of-arrays in the caller's order.
struct SubstepRollout {
  struct Context {
    TwoDScene* scene = nullptr;
    int steps = 0;
    double dt = 0.0, tol = 1e-12;
    int maxit = 5000, every = 10;
  };
  struct Grads {
    std::vector<double> x, v, C;
    double gravity = 0.0;
    bool converged = true;
  };

  static bool forward(Context& ctx, TwoDScene& scene, int steps, double dt,
                      const std::vector<double>& x0, const std::vector<double>& v0,
                      const std::vector<double>& C0, std::vector<double>& x,
                      std::vector<double>& v, std::vector<double>& C, double tol = 1e-12,
                      int maxit = 5000, int every = 10) {
    ctx.scene = &scene;
    ctx.steps = steps;
    ctx.dt = dt;
    ctx.tol = tol;
    ctx.maxit = maxit;
    ctx.every = every;
    scene.setParticlesSoA(x0, v0, C0);
    int32_t it = 0;
    const wc_status st = wc_rollout_forward(scene.handle(), steps, dt, tol, maxit, every, 0, &it);
    check(st, scene.handle());
    scene.getParticlesSoA(x, v, C);
    return st == WC_OK;
  }
  static bool replay(const Context& ctx, const std::vector<double>& x0,
                     const std::vector<double>& v0, const std::vector<double>& C0,
                     std::vector<double>& x, std::vector<double>& v, std::vector<double>& C) {
    TwoDScene& scene = *ctx.scene;
    scene.setParticlesSoA(x0, v0, C0);
    int32_t it = 0;
    const wc_status st =
        wc_rollout_forward(scene.handle(), ctx.steps, ctx.dt, ctx.tol, ctx.maxit, ctx.every, 1, &it);
    check(st, scene.handle());
    scene.getParticlesSoA(x, v, C);
    return st == WC_OK;
  }
  static Grads backward(const Context& ctx, const std::vector<double>& grad_x,
                        const std::vector<double>& grad_v, const std::vector<double>& grad_C) {
    TwoDScene& scene = *ctx.scene;
    const int64_t n = scene.getNumParticles();
    Grads g;
    g.x = grad_x;
    g.v = grad_v;
    g.C = grad_C;
    g.x.resize(3 * n, 0.0);
    g.v.resize(3 * n, 0.0);
    g.C.resize(9 * n, 0.0);
    int32_t it = 0;
    const wc_status st = wc_rollout_backward(scene.handle(), ctx.dt, ctx.tol, ctx.maxit,
                                             g.x.data(), g.v.data(), g.C.data(), &g.gravity, &it);
    check(st, scene.handle());
    g.converged = st == WC_OK;
    return g;
  }
};

}  // namespace wcamd

// ------------------------------------------------- PCGSolver<double> drop-in

namespace robertbridson {

// Dynamic compressed sparse row matrix with sorted rows
// (Core/PCGSolver/SparseMatrix.h:24-167).
template <class T>
struct SparseMatrix {
  unsigned int n;
  std::vector<std::vector<unsigned int> > index;
  std::vector<std::vector<T> > value;

  explicit SparseMatrix(unsigned int n_ = 0, unsigned int expected_nonzeros_per_row = 7)
      : n(n_), index(n_), value(n_) {
    for (unsigned int i = 0; i < n; ++i) {
      index[i].reserve(expected_nonzeros_per_row);
      value[i].reserve(expected_nonzeros_per_row);
    }
  }
  void clear() {
    n = 0;
    index.clear();
    value.clear();
  }
  void zero() {
    for (unsigned int i = 0; i < n; ++i) {
      index[i].resize(0);
      value[i].resize(0);
    }
  }
  void resize(int n_) {
    n = n_;
    index.resize(n);
    value.resize(n);
  }
  T operator()(unsigned int i, unsigned int j) const {
    for (unsigned int k = 0; k < index[i].size(); ++k) {
      if (index[i][k] == j) return value[i][k];
      if (index[i][k] > j) return 0;
    }
    return 0;
  }
  void set_element(unsigned int i, unsigned int j, T v) { place(i, j, v, false); }
  void add_to_element(unsigned int i, unsigned int j, T v) { place(i, j, v, true); }

 private:
  void place(unsigned int i, unsigned int j, T v, bool add) {
    auto& ix = index[i];
    auto& vx = value[i];
    unsigned int k = 0;
    for (; k < ix.size(); ++k) {
      if (ix[k] == j) {
        vx[k] = add ? vx[k] + v : v;
        return;
      }
      if (ix[k] > j) break;
    }
    ix.insert(ix.begin() + k, j);
    vx.insert(vx.begin() + k, v);
  }
};

// PCGSolver<double> whose solve runs on the GPU (PCGSolver.h:245-330).
template <class T>
struct PCGSolver {
  static_assert(sizeof(T) == 8, "the reference's BLAS wrappers exist only for double");
  PCGSolver() { set_solver_parameters(1e-8, 500, 0.97, 0.25); }
  void set_solver_parameters(T tolerance_factor_, int max_iterations_,
                             T modified_incomplete_cholesky_parameter_ = 0.97,
                             T min_diagonal_ratio_ = 0.25) {
    tolerance_factor = tolerance_factor_;
    if (tolerance_factor < 1e-30) tolerance_factor = 1e-30;
    max_iterations = max_iterations_;
    modified_incomplete_cholesky_parameter = modified_incomplete_cholesky_parameter_;
    min_diagonal_ratio = min_diagonal_ratio_;
  }
  bool solve(const SparseMatrix<T>& matrix, const std::vector<T>& rhs, std::vector<T>& result,
             T& residual_out, int& iterations_out) {
    const unsigned int n = matrix.n;
    rowstart.assign(n + 1, 0);
    colindex.clear();
    values.clear();
    for (unsigned int i = 0; i < n; ++i) {
      rowstart[i + 1] = rowstart[i] + (uint32_t)matrix.index[i].size();
      colindex.insert(colindex.end(), matrix.index[i].begin(), matrix.index[i].end());
      values.insert(values.end(), matrix.value[i].begin(), matrix.value[i].end());
    }
    result.assign(n, 0);
    double res = 0.0;
    int32_t it = 0;
    // like form_preconditioner (PCGSolver.h:322-324) the factor always uses
    // tau = 0.97 and min-diagonal ratio 0.25
    const wc_status st = wc_pcg_solve_csr(n, rowstart.data(), colindex.data(), values.data(),
                                          rhs.data(), result.data(), tolerance_factor,
                                          max_iterations, &res, &it);
    wcamd::check(st, nullptr);
    residual_out = res;
    iterations_out = it;
    return st == WC_OK;
  }

  T tolerance_factor;
  int max_iterations;
  T modified_incomplete_cholesky_parameter;
  T min_diagonal_ratio;

 private:
  std::vector<uint32_t> rowstart, colindex;
  std::vector<double> values;
};

}  // namespace robertbridson

// AMGPCGSolveSparse (AlgebraicMultigrid.h:247-251), the solve Pressure.cpp:797
// calls, with its signature: PCG preconditioned by the aggregation AMG
// V-cycle, run on the GPU.  Dof_ijk[d] is DOF d's cell (any type with
// operator[] over three ints, such as Eigen's Vector3i); 0 <= i < ni etc.
// Returns false when the solve does not converge; throws wcamd::Error on a
// bad argument or a device error.
template <class T, class IJK>
bool AMGPCGSolveSparse(const robertbridson::SparseMatrix<T>& matrix, const std::vector<T>& rhs,
                       std::vector<T>& result, std::vector<IJK>& Dof_ijk, T tolerance_factor,
                       int max_iterations, T& residual_out, int& iterations_out, int ni, int nj,
                       int nk) {
  static_assert(sizeof(T) == 8, "the GPU solve runs in double");
  const unsigned int n = matrix.n;
  std::vector<uint32_t> rowstart(n + 1, 0), colindex;
  std::vector<double> values;
  std::vector<int32_t> ijk(3 * (size_t)n);
  for (unsigned int i = 0; i < n; ++i) {
    rowstart[i + 1] = rowstart[i] + (uint32_t)matrix.index[i].size();
    colindex.insert(colindex.end(), matrix.index[i].begin(), matrix.index[i].end());
    values.insert(values.end(), matrix.value[i].begin(), matrix.value[i].end());
    for (int d = 0; d < 3; ++d) ijk[3 * (size_t)i + d] = (int32_t)Dof_ijk[i][d];
  }
  result.assign(n, 0);
  double res = 0.0;
  int32_t it = 0;
  const wc_status st =
      wc_amgpcg_solve_csr(n, rowstart.data(), colindex.data(), values.data(), ijk.data(), ni, nj,
                          nk, rhs.data(), result.data(), tolerance_factor, max_iterations, &res,
                          &it);
  wcamd::check(st, nullptr);
  residual_out = res;
  iterations_out = it;
  return st == WC_OK;
}
