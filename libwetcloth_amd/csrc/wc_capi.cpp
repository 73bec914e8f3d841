// extern "C" entry points of libwcgpu.so (declared in include/wetcloth_amd.h).
// Each call validates its arguments, runs its phase on the context's stream
// and converts HIP failures into a status code plus wc_last_error().
#include <hip/hip_runtime.h>
#include <string.h>

#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <charconv>
#include <climits>
#include <cmath>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "wc_ctx.h"

using wc::HipError;

namespace {

thread_local std::string g_free_err;  // errors of context-free calls

wc_status fail(wc_ctx* ctx, wc_status code, const std::string& msg) {
  if (ctx)
    ctx->err = msg;
  else
    g_free_err = msg;
  return code;
}

wc_status hip_fail(wc_ctx* ctx, const HipError& e) {
  std::string m = std::string("HIP error ") + hipGetErrorString(e.code) +
                  " in " + e.what + " (" + e.file + ":" + std::to_string(e.line) + ")";
  return fail(ctx, WC_ERR_HIP, m);
}

#define WC_TRY(ctx, ...)                               \
  try {                                                \
    __VA_ARGS__                                        \
  } catch (const HipError& e) {                        \
    return hip_fail(ctx, e);                           \
  } catch (const std::bad_alloc&) {                    \
    return fail(ctx, WC_ERR_NOMEM, "host allocation failed"); \
  }

void alloc_grid(wc_ctx* c) {
  const int64_t nn = c->g.nnode;
  for (int d = 0; d < 3; ++d) {
    c->mass[d].ensure(nn);
    c->mom[d].ensure(nn);
    c->vol[d].ensure(nn);
    c->w[d].ensure(nn);
    c->rhs[d].ensure(nn);
    c->invC[d].ensure(nn);
    c->ustar[d].ensure(nn);
    c->unew[d].ensure(nn);
    c->utmp[d].ensure(nn);
    c->valid[d].ensure(nn);
    c->vtmp[d].ensure(nn);
  }
  c->phikey.ensure(nn);
  c->lphi.ensure(nn);
  c->pressure.ensure(nn);
  c->isdof.ensure(nn + 1);
  c->dofscan.ensure(nn + 1);
  c->touched.ensure(c->g.nbucket);
  c->act.ensure(c->g.nbucket);
  c->bstart.ensure(c->g.nbucket + 1);
  hipStream_t st = c->stream;
  for (int d = 0; d < 3; ++d) {
    WC_HIP(hipMemsetAsync(c->mass[d].p, 0, sizeof(double) * nn, st));
    WC_HIP(hipMemsetAsync(c->mom[d].p, 0, sizeof(double) * nn, st));
    WC_HIP(hipMemsetAsync(c->vol[d].p, 0, sizeof(double) * nn, st));
    WC_HIP(hipMemsetAsync(c->w[d].p, 0, sizeof(double) * nn, st));
    WC_HIP(hipMemsetAsync(c->unew[d].p, 0, sizeof(double) * nn, st));
    WC_HIP(hipMemsetAsync(c->valid[d].p, 0, nn, st));
  }
  WC_HIP(hipMemsetAsync(c->lphi.p, 0, sizeof(double) * nn, st));
  WC_HIP(hipMemsetAsync(c->pressure.p, 0, sizeof(double) * nn, st));
  WC_HIP(hipMemsetAsync(c->isdof.p, 0, sizeof(uint32_t) * (nn + 1), st));
  WC_HIP(hipMemsetAsync(c->act.p, 0, c->g.nbucket, st));
  WC_HIP(hipMemsetAsync(c->bstart.p, 0, sizeof(uint32_t) * (c->g.nbucket + 1), st));
  WC_HIP(hipStreamSynchronize(st));
}

void release_all(wc_ctx* c) {
  c->x.release();
  c->v.release();
  c->C.release();
  c->x2.release();
  c->v2.release();
  c->C2.release();
  c->fx.release();
  c->fv.release();
  c->fC.release();
  c->fx2.release();
  c->fv2.release();
  c->fC2.release();
  c->pid.release();
  c->pid2.release();
  c->pkey.release();
  c->xw.release();
  c->fxw.release();
  c->vmax_bits.release();
  c->bcount.release();
  c->bstart.release();
  c->bcursor.release();
  c->touched.release();
  c->act.release();
  for (int d = 0; d < 3; ++d) {
    c->mass[d].release();
    c->mom[d].release();
    c->vol[d].release();
    c->w[d].release();
    c->rhs[d].release();
    c->invC[d].release();
    c->ustar[d].release();
    c->unew[d].release();
    c->utmp[d].release();
    c->valid[d].release();
    c->vtmp[d].release();
  }
  c->phikey.release();
  c->lphi.release();
  c->isdof.release();
  c->dofscan.release();
  c->pressure.release();
  c->rowlen.release();
  c->col.release();
  c->val.release();
  c->b.release();
  c->p.release();
  c->dof_level.release();
  c->pcg.release();
  c->amg.release();
  c->dof_ijk.release();
  c->scan_tmp.release();
  wc::adjoint_release(c);
  c->mix.release();
}

wc::DevCsr system_csr(const wc_ctx* c) {
  wc::DevCsr A;
  A.n = c->ndof;
  A.stride = 7;
  A.len = c->rowlen.p;
  A.col = c->col.p;
  A.val = c->val.p;
  A.nnz_slots = 7 * c->ndof;
  return A;
}

void record(wc_ctx* c, int i) {
  if (c->events.have) WC_HIP(hipEventRecord(c->events.ev[i], c->stream));
}

// The reference's bucket grid around the particles: updateParticleBoundingBox
// rounds the bounding box out to multiples of two buckets (TwoDScene.cpp:
// 1606-1611) and rebucketizeParticles pads it by three buckets and counts
// its buckets (:1660-1690).  origin: the dense node coordinates of the
// grid's first cell; dims: its node counts (buckets * 4), as
// solveNodePressure passes them to the AMG (Pressure.cpp:316-318).
void reference_bucket_frame(wc_ctx* c, int origin[3], int dims[3]) {
  double lo[3], hi[3];
  wc::particles_bbox(c, lo, hi);
  const double dx = c->g.dx;
  const double bucket = wc::kBucket * dx;
  const double two = bucket * 2.0;
  for (int d = 0; d < 3; ++d) {
    const double bmin = floor(lo[d] / two) * two, bmax = ceil(hi[d] / two) * two;
    const double content = bmax - bmin + bucket * 3.0 * 2.0;
    const int cells = std::max(1, (int)ceil(content / dx));
    const double size = (double)cells * dx;
    const int nbk = std::max(1, (int)ceil(size / bucket));
    const double corner = bmin - bucket * 3.0;
    origin[d] = (int)llround(corner / dx) + wc::kGhost;
    dims[d] = nbk * wc::kBucket;
  }
}

// solve the assembled system into c->p and scatter it onto the grid
bool solve_system(wc_ctx* c, double tol, int maxit, double* res, int* it) {
  const wc::DevCsr A = system_csr(c);
  wc::GridSystem& gs = c->pcg.grid;
  gs.g = &c->g;
  gs.act = c->act.p;
  gs.isdof = c->isdof.p;
  gs.dofscan = c->dofscan.p;
  gs.rowlen = c->rowlen.p;
  gs.col = c->col.p;
  bool ok;
  if (c->solver == 1 && A.n > 0) {
    wc::pcg_prepare_vectors(c->stream, c->pcg, A.n);
    // AMGPCGSolveSparse (Pressure.cpp:797) on the DOFs' cells, bucket
    // handle * 4 + node in the bucket, and the bucket grid's node counts
    // (Pressure.cpp:316-318, 433-436), in the frame of the reference's
    // bucket grid: the aggregation and colouring of its levels depend on it
    c->pcg.sweep.valid = false;
    c->pcg.factored = false;
    c->pcg.passes = 0;  // FP64 throughout: no refinement passes
    int origin[3], dims[3];
    reference_bucket_frame(c, origin, dims);
    c->dof_ijk.ensure(3 * (size_t)A.n);
    wc::grid_dof_cells(c, c->dof_ijk.p, origin);
    wc::amg_build(c->stream, c->amg, A, c->dof_ijk.p, dims[0], dims[1], dims[2]);
    ok = wc::amgpcg_solve(c->stream, c->pcg, c->amg, A, c->b.p, c->p.p, tol, maxit, res, it);
  } else {
    wc::pcg_prepare(c->stream, c->pcg, A, c->dof_level.p, c->level_max - c->level_min + 1);
    ok = wc::pcg_solve(c->stream, c->pcg, A, c->b.p, c->p.p, tol, maxit, 0.97, 0.25, res, it);
  }
  wc::grid_scatter_pressure(c);
  c->last_residual = *res;
  c->last_iters = *it;
  return ok;
}

bool ready(wc_ctx* c) { return c && c->stream; }

}  // namespace

extern "C" {

wc_status wc_create(const wc_desc* desc, wc_ctx** out) {
  if (!desc || !out) return fail(nullptr, WC_ERR_ARG, "null argument");
  *out = nullptr;
  if (desc->ni <= 0 || desc->nj <= 0 || desc->nk <= 0 || !(desc->dx > 0))
    return fail(nullptr, WC_ERR_ARG, "grid size and dx must be positive");
  if (desc->precision != WC_FP64 && desc->precision != WC_MIXED)
    return fail(nullptr, WC_ERR_ARG, "unknown precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, WC_ERR_HIP, "no HIP device");
  wc_ctx* c = new wc_ctx();
  c->desc = *desc;
  c->mixed = desc->precision == WC_MIXED;
  // the mixed form sums P2G in fixed point by default (order independent,
  // and measured faster there: 5.6 against 6.0 ms at 256^3); the FP64 form
  // keeps FP64 sums, the closest to the sequential sums
  c->deterministic = c->mixed;
  c->pcg.mixed = c->mixed;
  c->g = wc::make_grid(desc->ni, desc->nj, desc->nk, desc->dx);
  try {
    WC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int i = 0; i < 8; ++i) WC_HIP(hipEventCreate(&c->events.ev[i]));
    c->events.have = true;
    alloc_grid(c);
  } catch (const HipError& e) {
    g_free_err = std::string("wc_create: ") + hipGetErrorString(e.code) + " in " + e.what;
    release_all(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return WC_ERR_HIP;
  }
  *out = c;
  return WC_OK;
}

void wc_destroy(wc_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  release_all(c);
  if (c->events.have)
    for (int i = 0; i < 8; ++i) (void)hipEventDestroy(c->events.ev[i]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* wc_last_error(const wc_ctx* c) {
  return c ? c->err.c_str() : g_free_err.c_str();
}

wc_status wc_set_particles(wc_ctx* c, int64_t n, const double* x4,
                           const double* v4, const double* fB) {
  if (!ready(c) || n < 0) return fail(c, WC_ERR_ARG, "bad context or count");
  if (n >= 0xFFFFFFFFll) return fail(c, WC_ERR_ARG, "too many particles");
  WC_TRY(c, { wc::particles_set_soa(c, n, x4, v4, fB, true); })
  return WC_OK;
}

wc_status wc_get_particles(wc_ctx* c, double* x4, double* v4, double* fB) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  WC_TRY(c, { wc::particles_get_soa(c, x4, v4, fB, true); })
  return WC_OK;
}

wc_status wc_set_particles_soa(wc_ctx* c, int64_t n, const double* x,
                               const double* v, const double* C) {
  if (!ready(c) || n < 0) return fail(c, WC_ERR_ARG, "bad context or count");
  if (n >= 0xFFFFFFFFll) return fail(c, WC_ERR_ARG, "too many particles");
  WC_TRY(c, { wc::particles_set_soa(c, n, x, v, C, false); })
  return WC_OK;
}

wc_status wc_get_particles_soa(wc_ctx* c, double* x, double* v, double* C) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  WC_TRY(c, { wc::particles_get_soa(c, x, v, C, false); })
  return WC_OK;
}

wc_status wc_num_particles(const wc_ctx* c, int64_t* n) {
  if (!c || !n) return WC_ERR_ARG;
  *n = c->np;
  return WC_OK;
}

wc_status wc_seed_block(wc_ctx* c, const double lo[3], const double hi[3],
                        uint64_t seed, double sigma, int64_t* n_out) {
  if (!ready(c) || !lo || !hi) return fail(c, WC_ERR_ARG, "bad argument");
  WC_TRY(c, {
    const int64_t n = wc::particles_seed_block(c, lo, hi, seed, sigma);
    if (n_out) *n_out = n;
  })
  return WC_OK;
}

wc_status wc_set_elasto_particles(wc_ctx* c, int64_t ne, const double* x, const double* m,
                                  const double* fluid_m, const double* fluid_vol,
                                  const double* rest_vol, const double* rest_volume_fraction,
                                  const double* orientation, const double* shape_factor,
                                  const wc_mix_params* prm) {
  if (!ready(c) || ne < 0) return fail(c, WC_ERR_ARG, "bad argument");
  if (c->frz_mode != 0)
    return fail(c, WC_ERR_STATE,
                "decisions are being recorded or replayed: the adjoint covers the liquid-only "
                "substep");
  if (ne > 0) {
    if (!x || !m || !fluid_m || !fluid_vol || !rest_vol || !rest_volume_fraction ||
        !orientation || !shape_factor || !prm)
      return fail(c, WC_ERR_ARG, "every solid array and the parameters are needed");
    const double vals[6] = {prm->yarn_diameter, prm->viscosity, prm->yazdchi_power,
                            prm->pore_radius, prm->surf_tension_coeff, prm->rest_contact_angle};
    for (double v : vals)
      if (!std::isfinite(v)) return fail(c, WC_ERR_ARG, "the mixture parameters must be finite");
    if (prm->use_drag && (prm->yarn_diameter <= 0 || prm->viscosity <= 0))
      return fail(c, WC_ERR_ARG, "the drag needs a positive yarn diameter and viscosity");
    for (int d = 0; d < 3; ++d) {
      const double hi = c->g.n[d] * c->g.dx;
      for (int64_t p = 0; p < ne; ++p) {
        const double v = x[d * ne + p];
        if (!(v >= 0.0 && v <= hi))
          return fail(c, WC_ERR_ARG, "solid particle " + std::to_string(p) +
                                         " lies outside the container box");
      }
    }
    const double* per[6] = {m, fluid_m, fluid_vol, rest_vol, rest_volume_fraction, shape_factor};
    for (const double* a : per)
      for (int64_t p = 0; p < ne; ++p)
        if (!std::isfinite(a[p]) || a[p] < 0.0)
          return fail(c, WC_ERR_ARG, "solid particle " + std::to_string(p) +
                                         ": masses, volumes, fractions and shape factors must "
                                         "be finite and non-negative");
    for (int64_t p = 0; p < 3 * ne; ++p)
      if (!std::isfinite(orientation[p]))
        return fail(c, WC_ERR_ARG, "the orientations must be finite");
  }
  WC_TRY(c, {
    wc::mix_set(c, ne, x, m, fluid_m, fluid_vol, rest_vol, rest_volume_fraction, orientation,
                shape_factor, ne > 0 ? *prm : wc_mix_params{});
  })
  return WC_OK;
}

wc_status wc_rebucketize(wc_ctx* c) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  WC_TRY(c, {
    wc::particles_sort(c);
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  return WC_OK;
}

wc_status wc_map_particle_nodes(wc_ctx* c) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  WC_TRY(c, {
    wc::particles_p2g(c);
    if (c->mix.ne > 0) wc::mix_touch(c);
    wc::grid_activate(c);
    wc::grid_nodes(c);
    if (c->mix.ne > 0) wc::mix_grid(c);
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  return WC_OK;
}

wc_status wc_step_velocity(wc_ctx* c, double dt) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  WC_TRY(c, {
    wc::grid_step_velocity(c, dt);
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  return WC_OK;
}

wc_status wc_project(wc_ctx* c, double dt, double tol, int32_t maxit,
                     double* residual_out, int32_t* iters_out) {
  if (!ready(c) || maxit < 0) return fail(c, WC_ERR_ARG, "bad argument");
  bool ok = true;
  double res = 0;
  int it = 0;
  WC_TRY(c, {
    wc::grid_assemble(c, dt);
    ok = solve_system(c, tol, maxit, &res, &it);
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  if (residual_out) *residual_out = res;
  if (iters_out) *iters_out = it;
  return ok ? WC_OK : WC_NOT_CONVERGED;
}

wc_status wc_apply_pressure(wc_ctx* c, double dt) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  WC_TRY(c, {
    wc::grid_apply_pressure(c, dt);
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  return WC_OK;
}

wc_status wc_map_node_particles(wc_ctx* c) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  WC_TRY(c, {
    wc::particles_g2p(c);
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  return WC_OK;
}

wc_status wc_advect_scene(wc_ctx* c, double dt) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  WC_TRY(c, {
    wc::particles_advect(c, dt);
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  return WC_OK;
}

wc_status wc_solid_projection(wc_ctx* c, double dt) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  WC_TRY(c, {
    wc::particles_solid_projection(c, dt);
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  return WC_OK;
}

wc_status wc_map_node_particles_advect(wc_ctx* c, double dt) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  WC_TRY(c, {
    wc::particles_g2p_advect(c, dt);
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  return WC_OK;
}

wc_status wc_max_fluid_velocity(wc_ctx* c, double* vmax) {
  if (!ready(c) || !vmax) return fail(c, WC_ERR_ARG, "bad argument");
  WC_TRY(c, { *vmax = wc::particles_max_velocity(c); })
  return WC_OK;
}

}  // extern "C"

namespace {

// WetClothCore::stepSystem's substep count (WetClothCore.cpp:52-62): liquid
// particles may move 3 cells per substep, solids 1/3 (none here: their
// bound is then 1/30 s), and a substep is at most 1/30 s.
int cfl_substeps(double dt, double dx, double max_fluid_vel) {
  const double max_elasto_vel = 0.0;
  const double max_elasto_dt = std::min(dx / std::max(1e-63, max_elasto_vel) / 3.0, 1.0 / 30.0);
  const double max_fluid_dt = std::min(dx / std::max(1e-63, max_fluid_vel) * 3.0, 1.0 / 30.0);
  const double max_dt = std::min(max_elasto_dt, max_fluid_dt);
  return std::max(1, (int)ceil(dt / max_dt));
}

// one substep in the order of WetClothCore.cpp:109-262, fused where the
// reference's calls touch the same data (G2P, advection and projection)
bool substep(wc_ctx* c, double dt, double tol, int maxit, double* res, int* it) {
  record(c, 0);
  wc::particles_sort(c);
  record(c, 1);
  wc::particles_p2g(c);
  if (c->mix.ne > 0) wc::mix_touch(c);
  wc::grid_activate(c);
  wc::grid_nodes(c);
  if (c->mix.ne > 0) wc::mix_grid(c);
  record(c, 2);
  wc::grid_step_velocity(c, dt);
  record(c, 3);
  wc::grid_assemble(c, dt);
  record(c, 4);
  const bool ok = solve_system(c, tol, maxit, res, it);
  record(c, 5);
  wc::grid_apply_pressure(c, dt);
  record(c, 6);
  wc::particles_g2p_advect(c, dt);
  record(c, 7);
  WC_HIP(hipStreamSynchronize(c->stream));
  // the adjoint of this substep may follow (its decisions are recorded or
  // replayed and its start state kept)
  c->adj.ready = c->frz_mode != 0 && c->adj.have_start;
  for (int i = 0; i < 7; ++i) {
    float ms = 0.f;
    WC_HIP(hipEventElapsedTime(&ms, c->events.ev[i], c->events.ev[i + 1]));
    c->phase_ms[i] = ms;
  }
  return ok;
}

}  // namespace

extern "C" {

wc_status wc_substep(wc_ctx* c, double dt, double tol, int32_t maxit, double* residual_out,
                     int32_t* iters_out) {
  if (!ready(c) || maxit < 0) return fail(c, WC_ERR_ARG, "bad argument");
  bool ok = true;
  double res = 0;
  int it = 0;
  WC_TRY(c, { ok = substep(c, dt, tol, maxit, &res, &it); })
  if (residual_out) *residual_out = res;
  if (iters_out) *iters_out = it;
  return ok ? WC_OK : WC_NOT_CONVERGED;
}

wc_status wc_step(wc_ctx* c, double dt, double tol, int32_t maxit,
                  double* residual_out, int32_t* iters_out) {
  if (!ready(c) || maxit < 0 || !(dt >= 0)) return fail(c, WC_ERR_ARG, "bad argument");
  bool ok = true;
  double res = 0;
  int it = 0;
  WC_TRY(c, {
    const int n = cfl_substeps(dt, c->g.dx, wc::particles_max_velocity(c));
    const double sub_dt = dt / (double)n;
    c->last_substeps = n;
    // like the reference, every substep runs even when a solve does not
    // converge; the step reports it
    for (int k = 0; k < n; ++k) {
      double r = 0;
      int i = 0;
      if (!substep(c, sub_dt, tol, maxit, &r, &i)) ok = false;
      res = r;
      it += i;
    }
  })
  if (residual_out) *residual_out = res;
  if (iters_out) *iters_out = it;
  return ok ? WC_OK : WC_NOT_CONVERGED;
}

wc_status wc_step_substeps(const wc_ctx* c, int32_t* n) {
  if (!c || !n) return WC_ERR_ARG;
  *n = c->last_substeps;
  return WC_OK;
}

}  // extern "C"

// ---- the adjoint ----

namespace {

constexpr int kMaxSlots = 1 << 20;

// record (mode 1) into or replay (mode 2) from slot; mode 0 turns both off
wc_status set_frozen(wc_ctx* c, int mode, int slot) {
  if (mode < 0 || mode > 2 || slot < 0 || slot >= kMaxSlots)
    return fail(c, WC_ERR_ARG, "mode must be 0, 1 or 2 and slot in [0, 2^20)");
  if (mode != 0 && c->mixed)
    return fail(c, WC_ERR_STATE, "recorded decisions and the adjoint need a WC_FP64 context");
  if (mode != 0 && c->mix.ne > 0)
    return fail(c, WC_ERR_STATE,
                "the adjoint covers the liquid-only substep: remove the solid first "
                "(wc_set_elasto_particles with ne = 0)");
  if (mode == 2 && ((size_t)slot >= c->frz.size() || !c->frz[(size_t)slot].recorded ||
                    c->frz[(size_t)slot].np != c->np))
    return fail(c, WC_ERR_STATE,
                "slot " + std::to_string(slot) + " holds no recorded substep of these particles");
  if (mode != 2 && c->frz_dep.p)
    WC_HIP(hipMemsetAsync(c->frz_dep.p, 0, sizeof(unsigned long long), c->stream));
  c->frz_mode = mode;
  c->frz_slot = slot;
  if (mode == 1) {
    wc::FrozenRecord& r = wc::frozen_record(c);
    WC_HIP(hipMemsetAsync(r.proj.p, 0, std::max<int64_t>(1, c->np), c->stream));
    r.recorded = true;
    r.np = c->np;
  }
  return WC_OK;
}

void upload_grad(wc_ctx* c, const double* xb, const double* vb, const double* Cb) {
  wc::AdjointWork& a = c->adj;
  const int64_t n = c->np;
  a.gx.ensure(std::max<int64_t>(1, 3 * n));
  a.gv.ensure(std::max<int64_t>(1, 3 * n));
  a.gC.ensure(std::max<int64_t>(1, 9 * n));
  a.n = n;
  hipStream_t st = c->stream;
  auto up = [&](double* dst, const double* src, int64_t cnt) {
    if (src)
      WC_HIP(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyHostToDevice, st));
    else
      WC_HIP(hipMemsetAsync(dst, 0, sizeof(double) * cnt, st));
  };
  if (n > 0) {
    up(a.gx.p, xb, 3 * n);
    up(a.gv.p, vb, 3 * n);
    up(a.gC.p, Cb, 9 * n);
  }
}

void download_grad(wc_ctx* c, double* xb, double* vb, double* Cb) {
  const wc::AdjointWork& a = c->adj;
  const int64_t n = a.n;
  hipStream_t st = c->stream;
  if (n > 0) {
    if (xb) WC_HIP(hipMemcpyAsync(xb, a.gx.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
    if (vb) WC_HIP(hipMemcpyAsync(vb, a.gv.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
    if (Cb) WC_HIP(hipMemcpyAsync(Cb, a.gC.p, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, st));
  }
  WC_HIP(hipStreamSynchronize(st));
}

}  // namespace

extern "C" {

wc_status wc_set_gravity(wc_ctx* c, const double g[3]) {
  if (!ready(c) || !g) return fail(c, WC_ERR_ARG, "bad argument");
  for (int d = 0; d < 3; ++d) c->desc.gravity[d] = g[d];
  return WC_OK;
}

wc_status wc_set_frozen(wc_ctx* c, int32_t mode, int32_t slot) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  wc_status st = WC_OK;
  WC_TRY(c, { st = set_frozen(c, mode, slot); })
  return st;
}

wc_status wc_frozen_departures(wc_ctx* c, int64_t* count) {
  if (!ready(c) || !count) return fail(c, WC_ERR_ARG, "bad argument");
  unsigned long long n = 0;
  WC_TRY(c, {
    if (c->frz_dep.p)
      WC_HIP(hipMemcpyAsync(&n, c->frz_dep.p, sizeof(n), hipMemcpyDeviceToHost, c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  *count = (int64_t)n;
  return WC_OK;
}

wc_status wc_checkpoint(wc_ctx* c, int32_t slot, int32_t restore) {
  if (!ready(c) || slot < 0 || slot >= kMaxSlots) return fail(c, WC_ERR_ARG, "bad argument");
  if (c->mixed) return fail(c, WC_ERR_STATE, "checkpoints need a WC_FP64 context");
  if (restore && (size_t)slot >= c->ckpt.size())
    return fail(c, WC_ERR_STATE, "no checkpoint in slot " + std::to_string(slot));
  WC_TRY(c, {
    if (restore)
      wc::checkpoint_restore(c, slot);
    else
      wc::checkpoint_save(c, slot);
    WC_HIP(hipStreamSynchronize(c->stream));
  })
  return WC_OK;
}

wc_status wc_substep_adjoint(wc_ctx* c, double dt, double tol, int32_t maxit, double* xb,
                             double* vb, double* Cb, double* gb, double* residual_out,
                             int32_t* iters_out) {
  if (!ready(c) || maxit < 0) return fail(c, WC_ERR_ARG, "bad argument");
  if (!c->adj.ready)
    return fail(c, WC_ERR_STATE,
                "wc_substep_adjoint must follow wc_substep run with its decisions recorded or "
                "replayed (wc_set_frozen mode 1 or 2)");
  bool ok = true;
  double res = 0, g = 0;
  int it = 0;
  WC_TRY(c, {
    upload_grad(c, xb, vb, Cb);
    ok = wc::adjoint_substep(c, dt, tol, maxit, &g, &res, &it);
    download_grad(c, xb, vb, Cb);
  })
  if (gb) *gb += g;
  if (residual_out) *residual_out = res;
  if (iters_out) *iters_out = it;
  return ok ? WC_OK : WC_NOT_CONVERGED;
}

wc_status wc_rollout_forward(wc_ctx* c, int32_t steps, double dt, double tol, int32_t maxit,
                             int32_t every, int32_t replay, int32_t* iters_out) {
  if (!ready(c) || steps < 0 || steps >= kMaxSlots || every <= 0 || maxit < 0)
    return fail(c, WC_ERR_ARG, "bad argument");
  if (c->mixed) return fail(c, WC_ERR_STATE, "the rollout needs a WC_FP64 context");
  if (replay && steps != c->roll_steps)
    return fail(c, WC_ERR_STATE, "a replay must have the recorded rollout's length");
  bool ok = true;
  int total = 0;
  wc_status st = WC_OK;
  WC_TRY(c, {
    if (!replay) {
      c->roll_steps = 0;  // valid again once the run is recorded
      c->roll_every = every;
    }
    if (c->frz_dep.p)
      WC_HIP(hipMemsetAsync(c->frz_dep.p, 0, sizeof(unsigned long long), c->stream));
    for (int k = 0; k < steps && st == WC_OK; ++k) {
      if (!replay && k % every == 0) wc::checkpoint_save(c, k / every);
      st = set_frozen(c, replay ? 2 : 1, k);
      if (st != WC_OK) break;
      double r = 0;
      int i = 0;
      if (!substep(c, dt, tol, maxit, &r, &i)) ok = false;
      total += i;
    }
    c->frz_mode = 0;
    if (!replay && st == WC_OK) c->roll_steps = steps;
  })
  if (st != WC_OK) return st;
  if (iters_out) *iters_out = total;
  return ok ? WC_OK : WC_NOT_CONVERGED;
}

wc_status wc_rollout_backward(wc_ctx* c, double dt, double tol, int32_t maxit, double* xb,
                              double* vb, double* Cb, double* gb, int32_t* iters_out) {
  if (!ready(c) || maxit < 0) return fail(c, WC_ERR_ARG, "bad argument");
  const int steps = c->roll_steps, every = c->roll_every;
  if (steps <= 0) return fail(c, WC_ERR_STATE, "no recorded rollout: call wc_rollout_forward");
  bool ok = true;
  int total = 0;
  double g = 0;
  wc_status st = WC_OK;
  WC_TRY(c, {
    const int nseg = (steps + every - 1) / every;
    wc::checkpoint_restore(c, 0);
    upload_grad(c, xb, vb, Cb);
    for (int seg = nseg - 1; seg >= 0 && st == WC_OK; --seg) {
      const int c0 = seg * every, c1 = std::min(steps, c0 + every);
      // the segment's start states, from its checkpoint
      wc::checkpoint_restore(c, seg);
      for (int k = c0; k < c1 && st == WC_OK; ++k) {
        wc::checkpoint_save(c, nseg + (k - c0));
        if (k + 1 == c1) break;
        st = set_frozen(c, 2, k);
        double r = 0;
        int i = 0;
        if (st == WC_OK) substep(c, dt, tol, maxit, &r, &i);
      }
      // each substep again from its start state, then its adjoint
      for (int k = c1 - 1; k >= c0 && st == WC_OK; --k) {
        wc::checkpoint_restore(c, nseg + (k - c0));
        st = set_frozen(c, 2, k);
        if (st != WC_OK) break;
        double r = 0;
        int i = 0;
        substep(c, dt, tol, maxit, &r, &i);
        if (!wc::adjoint_substep(c, dt, tol, maxit, &g, &r, &i)) ok = false;
        total += i;
      }
    }
    c->frz_mode = 0;
    if (st == WC_OK) {
      download_grad(c, xb, vb, Cb);
      wc::checkpoint_restore(c, 0);
      WC_HIP(hipStreamSynchronize(c->stream));
    }
  })
  if (st != WC_OK) return st;
  if (gb) *gb += g;
  if (iters_out) *iters_out = total;
  return ok ? WC_OK : WC_NOT_CONVERGED;
}

wc_status wc_grid_info(const wc_ctx* c, int32_t nx[3], int64_t* nnode,
                       int64_t* nbucket) {
  if (!c) return WC_ERR_ARG;
  if (nx)
    for (int d = 0; d < 3; ++d) nx[d] = c->g.nx[d];
  if (nnode) *nnode = c->g.nnode;
  if (nbucket) *nbucket = c->g.nbucket;
  return WC_OK;
}

wc_status wc_download_field(wc_ctx* c, int32_t field, int32_t axis, void* host) {
  if (!ready(c) || !host) return fail(c, WC_ERR_ARG, "bad argument");
  if (axis < 0 || axis > 2) return fail(c, WC_ERR_ARG, "axis must be 0..2");
  const int64_t nn = c->g.nnode;
  hipStream_t st = c->stream;
  WC_TRY(c, {
    WC_HIP(hipStreamSynchronize(st));
    switch (field) {
      case WC_FIELD_ACT:
        WC_HIP(hipMemcpy(host, c->act.p, c->g.nbucket, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_LPHI:
        WC_HIP(hipMemcpy(host, c->lphi.p, sizeof(double) * nn, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_W:
        WC_HIP(hipMemcpy(host, c->w[axis].p, sizeof(double) * nn, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_MASS:
        WC_HIP(hipMemcpy(host, c->mass[axis].p, sizeof(double) * nn, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_VEL:
        WC_HIP(hipMemcpy(host, c->mom[axis].p, sizeof(double) * nn, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_VOL:
        WC_HIP(hipMemcpy(host, c->vol[axis].p, sizeof(double) * nn, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_RHS:
        WC_HIP(hipMemcpy(host, c->rhs[axis].p, sizeof(double) * nn, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_USTAR:
        WC_HIP(hipMemcpy(host, c->ustar[axis].p, sizeof(double) * nn, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_UNEW:
        WC_HIP(hipMemcpy(host, c->unew[axis].p, sizeof(double) * nn, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_PRESSURE:
        WC_HIP(hipMemcpy(host, c->pressure.p, sizeof(double) * nn, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_VALID:
        WC_HIP(hipMemcpy(host, c->valid[axis].p, nn, hipMemcpyDeviceToHost));
        break;
      case WC_FIELD_PSI:
      case WC_FIELD_SAT:
      case WC_FIELD_SMASS:
      case WC_FIELD_SVOL:
      case WC_FIELD_SF:
      case WC_FIELD_ORI_X:
      case WC_FIELD_ORI_Y:
      case WC_FIELD_ORI_Z:
      case WC_FIELD_Q:
      case WC_FIELD_HDV:
      case WC_FIELD_PSI_P:
      case WC_FIELD_SAT_P:
      case WC_FIELD_PORE_P: {
        const wc::MixState& M = c->mix;
        if (M.ne == 0) return fail(c, WC_ERR_STATE, "no solid (wc_set_elasto_particles)");
        const double* src = nullptr;
        switch (field) {
          case WC_FIELD_PSI: src = M.psi[axis].p; break;
          case WC_FIELD_SAT: src = M.sat[axis].p; break;
          case WC_FIELD_SMASS: src = M.plane(c->g, axis, wc::kMixPlaneMass); break;
          case WC_FIELD_SVOL: src = M.plane(c->g, axis, wc::kMixPlaneSvol); break;
          case WC_FIELD_SF: src = M.plane(c->g, axis, wc::kMixPlaneSf); break;
          case WC_FIELD_ORI_X:
          case WC_FIELD_ORI_Y:
          case WC_FIELD_ORI_Z:
            src = M.plane(c->g, axis, wc::kMixPlaneOri + (field - WC_FIELD_ORI_X));
            break;
          case WC_FIELD_Q: src = M.Q[axis].p; break;
          case WC_FIELD_HDV: src = M.hdv[axis].p; break;
          case WC_FIELD_PSI_P: src = M.psi_p.p; break;
          case WC_FIELD_SAT_P: src = M.sat_p.p; break;
          default: src = M.pp.p; break;
        }
        WC_HIP(hipMemcpy(host, src, sizeof(double) * nn, hipMemcpyDeviceToHost));
        break;
      }
      case WC_FIELD_DOF: {
        std::vector<uint32_t> f(nn), s(nn);
        WC_HIP(hipMemcpy(f.data(), c->isdof.p, sizeof(uint32_t) * nn, hipMemcpyDeviceToHost));
        WC_HIP(hipMemcpy(s.data(), c->dofscan.p, sizeof(uint32_t) * nn, hipMemcpyDeviceToHost));
        int32_t* out = (int32_t*)host;
        for (int64_t i = 0; i < nn; ++i) out[i] = f[i] ? (int32_t)s[i] : -1;
        break;
      }
      default:
        return fail(c, WC_ERR_ARG, "unknown field");
    }
  })
  return WC_OK;
}

wc_status wc_download_system(wc_ctx* c, int64_t* ndof, int64_t* nnz,
                             uint32_t* rowstart, uint32_t* col, double* val,
                             double* rhs, double* solution) {
  if (!ready(c)) return fail(c, WC_ERR_ARG, "bad context");
  const int64_t n = c->ndof;
  WC_TRY(c, {
    WC_HIP(hipStreamSynchronize(c->stream));
    std::vector<uint32_t> len(n), hc(7 * n);
    std::vector<double> hv(7 * n);
    if (n > 0) {
      WC_HIP(hipMemcpy(len.data(), c->rowlen.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    }
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) total += len[i];
    if (ndof) *ndof = n;
    if (nnz) *nnz = total;
    if (rowstart || col || val) {
      if (n > 0) {
        WC_HIP(hipMemcpy(hc.data(), c->col.p, sizeof(uint32_t) * 7 * n, hipMemcpyDeviceToHost));
        WC_HIP(hipMemcpy(hv.data(), c->val.p, sizeof(double) * 7 * n, hipMemcpyDeviceToHost));
      }
      int64_t e = 0;
      for (int64_t i = 0; i < n; ++i) {
        if (rowstart) rowstart[i] = (uint32_t)e;
        for (uint32_t t = 0; t < len[i]; ++t, ++e) {
          if (col) col[e] = hc[7 * i + t];
          if (val) val[e] = hv[7 * i + t];
        }
      }
      if (rowstart) rowstart[n] = (uint32_t)e;
    }
    if (rhs && n > 0)
      WC_HIP(hipMemcpy(rhs, c->b.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (solution && n > 0)
      WC_HIP(hipMemcpy(solution, c->p.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  })
  return WC_OK;
}

namespace {
// the context's last factored system, or an error message
const char* no_system(const wc_ctx* c) {
  if (c->ndof == 0) return "no pressure system: call wc_project first";
  if (!c->pcg.factored)
    return "the pressure system has not been factored: call wc_project first";
  return nullptr;
}
}  // namespace

wc_status wc_apply_preconditioner(wc_ctx* c, const double* r, double* z) {
  if (!ready(c) || !r || !z) return fail(c, WC_ERR_ARG, "bad argument");
  if (const char* m = no_system(c)) return fail(c, WC_ERR_STATE, m);
  WC_TRY(c, {
    const int64_t n = c->ndof;
    wc::DevBuf<double> dr, dz;
    dr.ensure(n);
    dz.ensure(n);
    WC_HIP(hipMemcpyAsync(dr.p, r, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    wc::SweepPlan& P = c->pcg.sweep;
    if (P.valid) {
      wc::box_gather(c->stream, P, dr.p, P.rb.p);
      wc::sweep_apply(c->stream, P, P.rb.p, nullptr, n);
      wc::box_scatter(c->stream, P, P.zb.p, dz.p);
    } else {
      wc::pcg_apply(c->stream, c->pcg, system_csr(c), dr.p, dz.p);
    }
    WC_HIP(hipMemcpyAsync(z, dz.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    const bool aborted = P.valid && wc::sweep_aborted(c->stream, P);
    dr.release();
    dz.release();
    if (aborted)
      throw wc::HipError{hipErrorLaunchTimeOut,
                         "pipelined MIC(0) sweep: a tile wait exceeded its bound",
                         __FILE__, __LINE__};
  })
  return WC_OK;
}

wc_status wc_system_multiply(wc_ctx* c, const double* x, double* y) {
  if (!ready(c) || !x || !y) return fail(c, WC_ERR_ARG, "bad argument");
  if (const char* m = no_system(c)) return fail(c, WC_ERR_STATE, m);
  WC_TRY(c, {
    const int64_t n = c->ndof;
    wc::DevBuf<double> dx, dy;
    dx.ensure(n);
    dy.ensure(n);
    WC_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    wc::SweepPlan& P = c->pcg.sweep;
    if (P.valid) {
      wc::box_gather(c->stream, P, dx.p, P.sb.p);
      wc::box_spmv(c->stream, P, P.sb.p, P.zb.p);
      wc::box_scatter(c->stream, P, P.zb.p, dy.p);
    } else {
      wc::pcg_spmv(c->stream, system_csr(c), dx.p, dy.p);
    }
    WC_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    WC_HIP(hipStreamSynchronize(c->stream));
    dx.release();
    dy.release();
  })
  return WC_OK;
}

wc_status wc_set_option(wc_ctx* c, const char* name, int64_t value) {
  if (!c || !name) return fail(c, WC_ERR_ARG, "bad argument");
  const std::string n(name);
  if (n == "solver") {
    if (value != 0 && value != 1) return fail(c, WC_ERR_ARG, "solver must be 0 (MIC(0)) or 1 (AMG)");
    c->solver = (int)value;
    return WC_OK;
  }
  if (n == "deterministic") {
    c->deterministic = value != 0;
    return WC_OK;
  }
  if (n == "pipelined_sweeps") {
    c->pcg.use_sweeps = value != 0;
    return WC_OK;
  }
  if (n == "sweep_subtiles") {
    c->pcg.sweep.sub = value != 0;
    return WC_OK;
  }
  if (n == "debug_abort_sweeps") {
    c->pcg.sweep.debug_abort = value != 0;
    return WC_OK;
  }
  if (n == "spmv_window") {
    c->pcg.sweep.spmv_win = value != 0;
    return WC_OK;
  }
  if (n == "spmv_run") {
    if (value < 1 || value > 4096) return fail(c, WC_ERR_ARG, "spmv_run must be in 1..4096");
    c->pcg.sweep.spmv_run = (int)value;
    return WC_OK;
  }
  if (n == "mixed_digits") {
    if (value < 1 || value > 15) return fail(c, WC_ERR_ARG, "mixed_digits must be 1..15");
    c->pcg.mixed_digits = (int)value;
    return WC_OK;
  }
  if (n == "mixed_refinements") {
    if (value < 0 || value > 100)
      return fail(c, WC_ERR_ARG, "mixed_refinements must be 0..100");
    c->pcg.mixed_refinements = (int)value;
    return WC_OK;
  }
  if (n == "kernel_timing") {
    WC_TRY(c, {
      WC_HIP(hipStreamSynchronize(c->stream));
      c->pcg.timer.reset();
      c->pcg.timer.on = value != 0;
      c->pcg.timer.stride = value > 1 ? (int)std::min<int64_t>(value, 1 << 20) : 1;
      c->pcg.timer.tick = 0;
    })
    return WC_OK;
  }
  return fail(c, WC_ERR_ARG, "unknown option " + n);
}

}  // extern "C"

namespace {

// ---- the reference serializer's particle files (App/TwoDSceneSerializer.cpp) ----

// "v x y z r\n" for particles [p0, p1): each number as std::ostream inserts
// a double at setprecision(8) with the default float field, which is printf
// "%.8g" ([facet.num.put.virtuals]); std::to_chars with chars_format::general
// and precision 8 is that conversion.
void format_fluid_lines(const double* x4, double r, int64_t p0, int64_t p1, std::string& out) {
  char buf[128];
  for (int64_t p = p0; p < p1; ++p) {
    char* e = buf;
    *e++ = 'v';
    const double v[4] = {x4[4 * p], x4[4 * p + 1], x4[4 * p + 2], r};
    for (int k = 0; k < 4; ++k) {
      *e++ = ' ';
      e = std::to_chars(e, buf + sizeof(buf) - 2, v[k], std::chars_format::general, 8).ptr;
    }
    *e++ = '\n';
    out.append(buf, (size_t)(e - buf));
  }
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

bool write_all(FILE* f, const void* p, size_t n) { return n == 0 || fwrite(p, 1, n, f) == n; }

}  // namespace

extern "C" {

wc_status wc_serialize_position_only(wc_ctx* c, const char* path) {
  if (!ready(c) || !path) return fail(c, WC_ERR_ARG, "bad argument");
  const int64_t n = c->np;
  const int64_t bytes = n * 4 * (int64_t)sizeof(double);
  if (bytes > INT_MAX) return fail(c, WC_ERR_ARG, "position block exceeds the format's int byte count");
  std::vector<double> x4;
  WC_TRY(c, {
    x4.resize((size_t)(4 * n));
    wc::particles_get_soa(c, x4.data(), nullptr, nullptr, true);
  })
  File F;
  F.f = fopen(path, "wb");
  if (!F.f) return fail(c, WC_ERR_ARG, std::string("cannot open ") + path);
  const int b0 = (int)bytes, b1 = 0;  // m_d_gauss is 0 x 3 without elastic bodies
  if (!write_all(F.f, &b0, sizeof(int)) || !write_all(F.f, x4.data(), (size_t)bytes) ||
      !write_all(F.f, &b1, sizeof(int)) || fclose(F.f) != 0) {
    F.f = nullptr;
    return fail(c, WC_ERR_ARG, std::string("write failed: ") + path);
  }
  F.f = nullptr;
  return WC_OK;
}

wc_status wc_load_position_only(wc_ctx* c, const char* path) {
  if (!ready(c) || !path) return fail(c, WC_ERR_ARG, "bad argument");
  const int64_t n = c->np;
  File F;
  F.f = fopen(path, "rb");
  if (!F.f) return fail(c, WC_ERR_ARG, std::string("cannot open ") + path);
  int b0 = -1, b1 = -1;
  if (fread(&b0, sizeof(int), 1, F.f) != 1 || (int64_t)b0 != n * 4 * (int64_t)sizeof(double))
    return fail(c, WC_ERR_ARG, "position block size does not match the context's particles");
  std::vector<double> x4, v4, fB;
  try {
    x4.resize((size_t)(4 * n));
  } catch (const std::bad_alloc&) {
    return fail(c, WC_ERR_NOMEM, "host allocation failed");
  }
  if ((n > 0 && fread(x4.data(), (size_t)b0, 1, F.f) != 1) || fread(&b1, sizeof(int), 1, F.f) != 1)
    return fail(c, WC_ERR_ARG, "truncated position file");
  if (b1 != 0)
    return fail(c, WC_ERR_ARG, "the file holds Gauss-point data (elastic bodies), which a liquid context has none of");
  WC_TRY(c, {
    v4.resize((size_t)(4 * n));
    fB.resize((size_t)(9 * n));
    wc::particles_get_soa(c, nullptr, v4.data(), fB.data(), true);
    wc::particles_set_soa(c, n, x4.data(), v4.data(), fB.data(), true);
  })
  return WC_OK;
}

wc_status wc_serialize_fluid(wc_ctx* c, const char* path) {
  if (!ready(c) || !path) return fail(c, WC_ERR_ARG, "bad argument");
  const int64_t n = c->np;
  std::vector<double> x4;
  WC_TRY(c, {
    x4.resize((size_t)(4 * n));
    wc::particles_get_soa(c, x4.data(), nullptr, nullptr, true);
  })
  File F;
  F.f = fopen(path, "wb");
  if (!F.f) return fail(c, WC_ERR_ARG, std::string("cannot open ") + path);
  // formatted in parallel in chunks, written in order
  const double r = c->desc.particle_radius;
  const int64_t chunk = 1 << 18;
  const int64_t nchunk = (n + chunk - 1) / chunk;
  const int nthr = (int)std::max<int64_t>(1, std::min<int64_t>(nchunk, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
  bool ok = true;
  try {
    for (int64_t c0 = 0; c0 < nchunk && ok; c0 += nthr) {
      const int m = (int)std::min<int64_t>(nthr, nchunk - c0);
      std::vector<std::string> text((size_t)m);
      std::vector<std::thread> th;
      for (int t = 0; t < m; ++t)
        th.emplace_back([&, t] {
          const int64_t p0 = (c0 + t) * chunk, p1 = std::min(n, p0 + chunk);
          format_fluid_lines(x4.data(), r, p0, p1, text[(size_t)t]);
        });
      for (auto& t : th) t.join();
      for (int t = 0; t < m && ok; ++t) ok = write_all(F.f, text[(size_t)t].data(), text[(size_t)t].size());
    }
  } catch (const std::exception&) {
    return fail(c, WC_ERR_NOMEM, "formatting the fluid file failed");
  }
  const bool closed = fclose(F.f) == 0;
  F.f = nullptr;
  if (!ok || !closed) return fail(c, WC_ERR_ARG, std::string("write failed: ") + path);
  return WC_OK;
}


wc_status wc_refinement_passes(const wc_ctx* c, int32_t* passes) {
  if (!c || !passes) return WC_ERR_ARG;
  *passes = c->mixed ? c->pcg.passes : 0;
  return WC_OK;
}

wc_status wc_kernel_stats(const wc_ctx* c, int32_t kernel, double* ms, int64_t* launches,
                          int64_t* units) {
  if (!c || kernel < 0 || kernel >= wc::KTimer::COUNT) return WC_ERR_ARG;
  // launches still pending in the timer are harvested first (the solves do
  // not wait for them); the statistics are the context's mutable state
  wc::KTimer& t = const_cast<wc_ctx*>(c)->pcg.timer;
  if (!t.pending.empty()) {
    if (hipStreamSynchronize(c->stream) != hipSuccess) return WC_ERR_HIP;
    try {
      t.harvest();
    } catch (const wc::HipError&) {
      return WC_ERR_HIP;
    }
  }
  if (ms) *ms = t.ms[kernel];
  if (launches) *launches = t.launches[kernel];
  if (units) *units = t.units[kernel];
  return WC_OK;
}

wc_status wc_phase_times(const wc_ctx* c, double ms[7]) {
  if (!c || !ms) return WC_ERR_ARG;
  for (int i = 0; i < 7; ++i) ms[i] = c->phase_ms[i];
  return WC_OK;
}

}  // extern "C"

// ---- context-free PCGSolver drop-in ----

namespace {

struct CsrUpload {
  wc::DevBuf<uint32_t> rowptr, col;
  wc::DevBuf<double> val;
  wc::DevCsr A;
  void up(hipStream_t st, int64_t n, const uint32_t* rs, const uint32_t* c,
          const double* v) {
    const int64_t nnz = rs[n];
    rowptr.ensure(n + 1);
    col.ensure(nnz);
    val.ensure(nnz);
    WC_HIP(hipMemcpyAsync(rowptr.p, rs, sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, st));
    if (nnz) {
      WC_HIP(hipMemcpyAsync(col.p, c, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice, st));
      if (v) WC_HIP(hipMemcpyAsync(val.p, v, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
    }
    A.n = n;
    A.rowptr = rowptr.p;
    A.col = col.p;
    A.val = val.p;
    A.nnz_slots = nnz;
  }
  void release() {
    rowptr.release();
    col.release();
    val.release();
  }
};

struct FreeSession {
  hipStream_t st = nullptr;
  CsrUpload csr;
  wc::PcgWork work;
  FreeSession() { WC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)); }
  ~FreeSession() {
    if (st) (void)hipStreamSynchronize(st);
    csr.release();
    work.release();
    if (st) (void)hipStreamDestroy(st);
  }
};

wc_status check_csr(int64_t n, const uint32_t* rs, const uint32_t* col) {
  if (n < 0 || (n > 0 && (!rs || !col))) return WC_ERR_ARG;
  if (n >= 0xFFFFFFFFll) return WC_ERR_ARG;
  if (rs && rs[0] != 0) return WC_ERR_ARG;
  for (int64_t i = 0; i < n; ++i) {
    if (rs[i + 1] < rs[i]) return WC_ERR_ARG;
    for (uint32_t e = rs[i]; e < rs[i + 1]; ++e) {
      if (col[e] >= (uint64_t)n) return WC_ERR_ARG;
      if (e > rs[i] && col[e] <= col[e - 1]) return WC_ERR_ARG;
    }
  }
  return WC_OK;
}

}  // namespace

extern "C" {

wc_status wc_pcg_solve_csr(int64_t n, const uint32_t* rowstart,
                           const uint32_t* col, const double* val,
                           const double* rhs, double* result, double tol,
                           int32_t maxit, double* residual_out,
                           int32_t* iters_out) {
  if (check_csr(n, rowstart, col) != WC_OK || (n > 0 && (!val || !rhs || !result)))
    return fail(nullptr, WC_ERR_ARG, "invalid CSR matrix (rows must be sorted, columns < n)");
  if (maxit < 0) return fail(nullptr, WC_ERR_ARG, "maxit < 0");
  double res = 0;
  int it = 0;
  bool ok = true;
  try {
    FreeSession S;
    if (n > 0) {
      S.csr.up(S.st, n, rowstart, col, val);
      wc::DevBuf<double> b, x;
      b.ensure(n);
      x.ensure(n);
      WC_HIP(hipMemcpyAsync(b.p, rhs, sizeof(double) * n, hipMemcpyHostToDevice, S.st));
      wc::pcg_prepare(S.st, S.work, S.csr.A, nullptr, 0);
      ok = wc::pcg_solve(S.st, S.work, S.csr.A, b.p, x.p, tol, maxit, 0.97, 0.25,
                         &res, &it);
      WC_HIP(hipMemcpyAsync(result, x.p, sizeof(double) * n, hipMemcpyDeviceToHost, S.st));
      WC_HIP(hipStreamSynchronize(S.st));
      b.release();
      x.release();
    }
  } catch (const HipError& e) {
    return hip_fail(nullptr, e);
  }
  if (residual_out) *residual_out = res;
  if (iters_out) *iters_out = it;
  return ok ? WC_OK : WC_NOT_CONVERGED;
}

}  // extern "C"

// ---- wc_pcg_solve_csr without allocation ----

namespace {

// The device memory of wc_pcg_solve_csr_noalloc: one fixed array per GPU,
// held until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: the array's device address on each device is
// looked up once (hipGetSymbolAddress, with that device current) and only
// that address is handed to copies and kernels.
static __device__ __align__(256) unsigned char scratch[WC_NOALLOC_SCRATCH_BYTES];

constexpr int kNoAllocDevices = 64;
struct NoAllocDevice {
  std::once_flag once;
  hipError_t err = hipSuccess;
  unsigned char* ptr = nullptr;  // this device's copy of `scratch`
  hipStream_t st = nullptr;      // created with it, kept with it
};
NoAllocDevice g_noalloc[kNoAllocDevices];

// Set-up on the current device, once: the array's address and a stream.
NoAllocDevice& noalloc_device() {
  int dev = 0;
  WC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices)
    throw HipError{hipErrorInvalidDevice, "device ordinal beyond the no-allocation table",
                   __FILE__, __LINE__};
  NoAllocDevice& d = g_noalloc[dev];
  std::call_once(d.once, [&d] {
    d.err = hipGetSymbolAddress((void**)&d.ptr, HIP_SYMBOL(scratch));
    if (d.err == hipSuccess && !d.ptr) d.err = hipErrorInvalidSymbol;
    if (d.err == hipSuccess) d.err = hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking);
  });
  WC_HIP(d.err);
  return d;
}

// Every buffer of the level-scheduled PCG solve of an n-row system with nnz
// stored entries, carved from one block at 256-byte boundaries.  With
// base == nullptr only the size is computed.
struct NoAllocParts {
  uint32_t *rowptr, *col, *tpos, *lvl_count, *lvl_cursor, *lvl_list;
  double *val, *b, *x, *lv, *invdiag, *adiag, *z, *r, *s, *partial, *scal;
  int* flags;
  int32_t* level;
  size_t bytes;
};

NoAllocParts noalloc_layout(unsigned char* base, int64_t n, int64_t nnz) {
  size_t off = 0;
  auto take = [&](size_t elems, size_t elem_bytes) -> void* {
    const size_t o = (off + 255) & ~size_t(255);
    off = o + elems * elem_bytes;
    return base ? base + o : nullptr;
  };
  NoAllocParts q;
  q.rowptr = (uint32_t*)take(n + 1, 4);
  q.col = (uint32_t*)take(nnz, 4);
  q.val = (double*)take(nnz, 8);
  q.b = (double*)take(n, 8);
  q.x = (double*)take(n, 8);
  q.tpos = (uint32_t*)take(nnz, 4);
  q.lv = (double*)take(nnz, 8);
  q.invdiag = (double*)take(n, 8);
  q.adiag = (double*)take(n, 8);
  q.z = (double*)take(n, 8);
  q.r = (double*)take(n, 8);
  q.s = (double*)take(n, 8);
  q.flags = (int*)take(4, 4);
  q.level = (int32_t*)take(n, 4);
  q.lvl_count = (uint32_t*)take(n, 4);  // a row's level is below its index: n levels at most
  q.lvl_cursor = (uint32_t*)take(n, 4);
  q.lvl_list = (uint32_t*)take(n, 4);
  q.partial = (double*)take(1024, 8);
  q.scal = (double*)take(8, 8);
  q.bytes = off;
  return q;
}

}  // namespace

extern "C" {

wc_status wc_pcg_solve_csr_noalloc(int64_t n, const uint32_t* rowstart, const uint32_t* col,
                                   const double* val, const double* rhs, double* result,
                                   double tol, int32_t maxit, double* residual_out,
                                   int32_t* iters_out) {
  if (check_csr(n, rowstart, col) != WC_OK || (n > 0 && (!val || !rhs || !result)))
    return fail(nullptr, WC_ERR_ARG, "invalid CSR matrix (rows must be sorted, columns < n)");
  if (maxit < 0) return fail(nullptr, WC_ERR_ARG, "maxit < 0");
  const int64_t nnz = n > 0 ? (int64_t)rowstart[n] : 0;
  const size_t need = noalloc_layout(nullptr, n, nnz).bytes;
  if (need > (size_t)WC_NOALLOC_SCRATCH_BYTES)
    return fail(nullptr, WC_ERR_ARG,
                "wc_pcg_solve_csr_noalloc: the system (" + std::to_string(n) + " rows, " +
                    std::to_string(nnz) + " entries) needs " + std::to_string(need) +
                    " bytes of device memory, more than the fixed " +
                    std::to_string((size_t)WC_NOALLOC_SCRATCH_BYTES) +
                    " (wc_pcg_solve_csr has no such limit)");
  double res = 0;
  int it = 0;
  bool ok = true;
  NoAllocDevice* d = nullptr;
  try {
    d = &noalloc_device();
    if (n > 0) {
      const NoAllocParts q = noalloc_layout(d->ptr, n, nnz);
      const hipStream_t st = d->st;
      WC_HIP(hipMemcpyAsync(q.rowptr, rowstart, sizeof(uint32_t) * (n + 1),
                            hipMemcpyHostToDevice, st));
      if (nnz) {
        WC_HIP(hipMemcpyAsync(q.col, col, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice, st));
        WC_HIP(hipMemcpyAsync(q.val, val, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
      }
      WC_HIP(hipMemcpyAsync(q.b, rhs, sizeof(double) * n, hipMemcpyHostToDevice, st));
      wc::DevCsr A;
      A.n = n;
      A.rowptr = q.rowptr;
      A.col = q.col;
      A.val = q.val;
      A.nnz_slots = nnz;
      wc::PcgWork w;
      w.no_alloc = true;
      w.tpos.lend(q.tpos, nnz);
      w.lv.lend(q.lv, nnz);
      w.invdiag.lend(q.invdiag, n);
      w.adiag.lend(q.adiag, n);
      w.z.lend(q.z, n);
      w.r.lend(q.r, n);
      w.s.lend(q.s, n);
      w.x.lend(q.x, n);
      w.flags.lend(q.flags, 4);
      w.level.lend(q.level, n);
      w.lvl_count.lend(q.lvl_count, n);
      w.lvl_cursor.lend(q.lvl_cursor, n);
      w.lvl_list.lend(q.lvl_list, n);
      w.partial.lend(q.partial, 1024);
      w.scal.lend(q.scal, 8);
      wc::pcg_prepare(st, w, A, nullptr, 0);
      ok = wc::pcg_solve(st, w, A, q.b, q.x, tol, maxit, 0.97, 0.25, &res, &it);
      WC_HIP(hipMemcpyAsync(result, q.x, sizeof(double) * n, hipMemcpyDeviceToHost, st));
      WC_HIP(hipStreamSynchronize(st));
      w.release();
    }
  } catch (const HipError& e) {
    if (d && d->st) (void)hipStreamSynchronize(d->st);
    return hip_fail(nullptr, e);
  }
  if (residual_out) *residual_out = res;
  if (iters_out) *iters_out = it;
  return ok ? WC_OK : WC_NOT_CONVERGED;
}

}  // extern "C"

namespace {
wc_status check_ijk(int64_t n, const int32_t* ijk, int32_t ni, int32_t nj, int32_t nk) {
  if (n > 0 && !ijk) return WC_ERR_ARG;
  if (ni <= 0 || nj <= 0 || nk <= 0) return n > 0 ? WC_ERR_ARG : WC_OK;
  if ((int64_t)ni * nj * nk >= 0x7FFFFFFFll) return WC_ERR_ARG;
  for (int64_t f = 0; f < n; ++f)
    if (ijk[3 * f] < 0 || ijk[3 * f] >= ni || ijk[3 * f + 1] < 0 || ijk[3 * f + 1] >= nj ||
        ijk[3 * f + 2] < 0 || ijk[3 * f + 2] >= nk)
      return WC_ERR_ARG;
  // every DOF its own cell: a coarse DOF keeps one child per position in
  // its 2x2x2 block
  std::vector<int64_t> key((size_t)n);
  for (int64_t f = 0; f < n; ++f)
    key[f] = ((int64_t)ijk[3 * f + 2] * nj + ijk[3 * f + 1]) * ni + ijk[3 * f];
  std::sort(key.begin(), key.end());
  for (int64_t f = 1; f < n; ++f)
    if (key[f] == key[f - 1]) return WC_ERR_ARG;
  return WC_OK;
}

// the AMG hierarchy of an uploaded system: run(...) gets it built
template <class F>
wc_status with_amg(int64_t n, const uint32_t* rowstart, const uint32_t* col, const double* val,
                   const int32_t* ijk, int32_t ni, int32_t nj, int32_t nk, F run) {
  if (check_csr(n, rowstart, col) != WC_OK || (n > 0 && !val))
    return fail(nullptr, WC_ERR_ARG, "invalid CSR matrix (rows must be sorted, columns < n)");
  if (check_ijk(n, ijk, ni, nj, nk) != WC_OK)
    return fail(nullptr, WC_ERR_ARG,
                "invalid DOF cells (need 0 <= i < ni, 0 <= j < nj, 0 <= k < nk, and no two DOFs in one cell)");
  try {
    FreeSession S;
    wc::AmgHier H;
    wc::DevBuf<int32_t> dijk;
    if (n > 0) {
      S.csr.up(S.st, n, rowstart, col, val);
      dijk.ensure(3 * (size_t)n);
      WC_HIP(hipMemcpyAsync(dijk.p, ijk, sizeof(int32_t) * 3 * (size_t)n, hipMemcpyHostToDevice,
                            S.st));
      wc::amg_build(S.st, H, S.csr.A, dijk.p, ni, nj, nk);
    }
    const wc_status st = run(S, H);
    WC_HIP(hipStreamSynchronize(S.st));
    H.release();
    dijk.release();
    return st;
  } catch (const HipError& e) {
    return hip_fail(nullptr, e);
  }
}
}  // namespace

extern "C" {

wc_status wc_amgpcg_solve_csr(int64_t n, const uint32_t* rowstart, const uint32_t* col,
                              const double* val, const int32_t* dof_ijk, int32_t ni,
                              int32_t nj, int32_t nk, const double* rhs, double* result,
                              double tol, int32_t maxit, double* residual_out,
                              int32_t* iters_out) {
  if (n > 0 && (!rhs || !result)) return fail(nullptr, WC_ERR_ARG, "bad argument");
  if (maxit < 0) return fail(nullptr, WC_ERR_ARG, "maxit < 0");
  double res = 0;
  int it = 0;
  bool ok = true;
  const wc_status st = with_amg(n, rowstart, col, val, dof_ijk, ni, nj, nk,
                                [&](FreeSession& S, wc::AmgHier& H) -> wc_status {
    if (n == 0) return WC_OK;
    wc::DevBuf<double> b, x;
    b.ensure(n);
    x.ensure(n);
    WC_HIP(hipMemcpyAsync(b.p, rhs, sizeof(double) * n, hipMemcpyHostToDevice, S.st));
    wc::pcg_prepare_vectors(S.st, S.work, n);
    ok = wc::amgpcg_solve(S.st, S.work, H, S.csr.A, b.p, x.p, tol, maxit, &res, &it);
    WC_HIP(hipMemcpyAsync(result, x.p, sizeof(double) * n, hipMemcpyDeviceToHost, S.st));
    WC_HIP(hipStreamSynchronize(S.st));
    b.release();
    x.release();
    return WC_OK;
  });
  if (st != WC_OK) return st;
  if (residual_out) *residual_out = res;
  if (iters_out) *iters_out = it;
  return ok ? WC_OK : WC_NOT_CONVERGED;
}

wc_status wc_amg_apply_csr(int64_t n, const uint32_t* rowstart, const uint32_t* col,
                           const double* val, const int32_t* dof_ijk, int32_t ni, int32_t nj,
                           int32_t nk, const double* r, double* z) {
  if (n > 0 && (!r || !z)) return fail(nullptr, WC_ERR_ARG, "bad argument");
  return with_amg(n, rowstart, col, val, dof_ijk, ni, nj, nk,
                  [&](FreeSession& S, wc::AmgHier& H) -> wc_status {
    if (n == 0) return WC_OK;
    wc::DevBuf<double> dr, dz;
    dr.ensure(n);
    dz.ensure(n);
    WC_HIP(hipMemcpyAsync(dr.p, r, sizeof(double) * n, hipMemcpyHostToDevice, S.st));
    wc::amg_vcycle(S.st, H, dr.p, dz.p);
    WC_HIP(hipMemcpyAsync(z, dz.p, sizeof(double) * n, hipMemcpyDeviceToHost, S.st));
    WC_HIP(hipStreamSynchronize(S.st));
    dr.release();
    dz.release();
    return WC_OK;
  });
}

wc_status wc_mic0_factor_csr(int64_t n, const uint32_t* rowstart,
                             const uint32_t* col, const double* val, double tau,
                             double mdr, double* invdiag, double* lv) {
  if (check_csr(n, rowstart, col) != WC_OK || (n > 0 && (!val || !invdiag || !lv)))
    return fail(nullptr, WC_ERR_ARG, "invalid CSR matrix");
  if (n == 0) return WC_OK;
  try {
    FreeSession S;
    S.csr.up(S.st, n, rowstart, col, val);
    wc::pcg_prepare(S.st, S.work, S.csr.A, nullptr, 0);
    wc::pcg_factor(S.st, S.work, S.csr.A, tau, mdr);
    const int64_t nnz = rowstart[n];
    WC_HIP(hipMemcpyAsync(invdiag, S.work.invdiag.p, sizeof(double) * n, hipMemcpyDeviceToHost, S.st));
    // lv is defined only above the diagonal; the rest is returned as zero
    std::vector<double> h(nnz);
    if (nnz) WC_HIP(hipMemcpyAsync(h.data(), S.work.lv.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, S.st));
    WC_HIP(hipStreamSynchronize(S.st));
    for (int64_t i = 0; i < n; ++i)
      for (uint32_t e = rowstart[i]; e < rowstart[i + 1]; ++e)
        lv[e] = col[e] > (uint32_t)i ? h[e] : 0.0;
  } catch (const HipError& e) {
    return hip_fail(nullptr, e);
  }
  return WC_OK;
}

wc_status wc_mic0_apply_csr(int64_t n, const uint32_t* rowstart,
                            const uint32_t* col, const double* invdiag,
                            const double* lv, const double* r, double* z) {
  if (check_csr(n, rowstart, col) != WC_OK || (n > 0 && (!invdiag || !lv || !r || !z)))
    return fail(nullptr, WC_ERR_ARG, "invalid CSR matrix");
  if (n == 0) return WC_OK;
  try {
    FreeSession S;
    S.csr.up(S.st, n, rowstart, col, nullptr);
    wc::pcg_prepare(S.st, S.work, S.csr.A, nullptr, 0);
    const int64_t nnz = rowstart[n];
    WC_HIP(hipMemcpyAsync(S.work.invdiag.p, invdiag, sizeof(double) * n, hipMemcpyHostToDevice, S.st));
    if (nnz) WC_HIP(hipMemcpyAsync(S.work.lv.p, lv, sizeof(double) * nnz, hipMemcpyHostToDevice, S.st));
    wc::DevBuf<double> dr, dz;
    dr.ensure(n);
    dz.ensure(n);
    WC_HIP(hipMemcpyAsync(dr.p, r, sizeof(double) * n, hipMemcpyHostToDevice, S.st));
    wc::pcg_apply(S.st, S.work, S.csr.A, dr.p, dz.p);
    WC_HIP(hipMemcpyAsync(z, dz.p, sizeof(double) * n, hipMemcpyDeviceToHost, S.st));
    WC_HIP(hipStreamSynchronize(S.st));
    dr.release();
    dz.release();
  } catch (const HipError& e) {
    return hip_fail(nullptr, e);
  }
  return WC_OK;
}

wc_status wc_spmv_csr(int64_t n, const uint32_t* rowstart, const uint32_t* col,
                      const double* val, const double* x, double* y) {
  if (check_csr(n, rowstart, col) != WC_OK || (n > 0 && (!val || !x || !y)))
    return fail(nullptr, WC_ERR_ARG, "invalid CSR matrix");
  if (n == 0) return WC_OK;
  try {
    FreeSession S;
    S.csr.up(S.st, n, rowstart, col, val);
    wc::DevBuf<double> dx, dy;
    dx.ensure(n);
    dy.ensure(n);
    WC_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, S.st));
    wc::pcg_spmv(S.st, S.csr.A, dx.p, dy.p);
    WC_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, S.st));
    WC_HIP(hipStreamSynchronize(S.st));
    dx.release();
    dy.release();
  } catch (const HipError& e) {
    return hip_fail(nullptr, e);
  }
  return WC_OK;
}

}  // extern "C"
