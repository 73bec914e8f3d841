// Grid sweeps for gfx950: bucket activation, solid weights and liquid phi,
// the velocity prediction, pressure-DOF numbering and matrix assembly, and
// the pressure application with extrapolation and the boundary constraint.
// Every sweep is one thread per node over the bucket-major node array, so a
// wavefront covers one 4x4x4 bucket and its loads are contiguous.
#include <math.h>

#include "wc_ctx.h"

namespace wc {
namespace {

inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// expandFluidNodesMarked(1) (TwoDScene.cpp:5021-5064): a bucket is active if
// it or any of its 26 neighbours holds a stencil node of a particle.
__global__ void k_activate(Grid g, const uint8_t* touched, uint8_t* act) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= g.nbucket) return;
  const int bi = (int)(b % g.nb[0]);
  const int bj = (int)((b / g.nb[0]) % g.nb[1]);
  const int bk = (int)(b / ((int64_t)g.nb[0] * g.nb[1]));
  uint8_t a = 0;
  for (int t = -1; t <= 1; ++t)
    for (int s = -1; s <= 1; ++s)
      for (int r = -1; r <= 1; ++r) {
        const int x = bi + r, y = bj + s, z = bk + t;
        if (x < 0 || y < 0 || z < 0 || x >= g.nb[0] || y >= g.nb[1] ||
            z >= g.nb[2])
          continue;
        if (touched[bucket_index(g, x, y, z)]) a = 1;
      }
  act[b] = a;
}

// corner phi seen by updateSolidWeights (TwoDScene.cpp:1899-1911)
__device__ inline double corner_phi_w(const Grid& g, const uint8_t* act, int i,
                                      int j, int k) {
  if (!node_active(g, act, i, j, k)) return 0.5 * g.dx;
  return container_phi(g, node_pos(g, i, 0.0), node_pos(g, j, 0.0),
                       node_pos(g, k, 0.0));
}

__device__ inline double clamp01(double f) {
  return f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
}

struct NodeArgs {
  Grid g;
  const uint8_t* act;
  const unsigned long long* phikey;
  double* lphi;
  double* w[3];
  double* mass[3];  // P2G's weight sums in, masses out
  double* vol[3];   // volumes out
  double* mom[3];   // momentum in, velocity out
  double fm, fvol;  // the particle mass and volume P2G weighted with
  FrozenView F;    // mode 2: lphi is already the replayed one
  int det;         // the sums are fixed point (the deterministic P2G)
};

// a fixed-point sum of the deterministic P2G (64-bit integer bits in a
// double's storage, in units of 2^-56 for weights, 2^-48 for momentum) as
// its value
__device__ inline double from_fix(double bits, double unit) {
  return (double)__double_as_longlong(bits) * unit;
}

// updateSolidWeights (TwoDScene.cpp:1871-1962), updateLiquidPhi's solid
// override and default (:3297-3348), and of P2G (:6700-6713) M = fm sum w,
// V = fvol sum w and u = P/M.
__global__ void k_grid_nodes(NodeArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode) return;
  const double bucket_size = kBucket * g.dx;
  if (!A.act[n >> 6]) {
    if (A.F.mode != 2) A.lphi[n] = 3.0 * bucket_size;
    for (int d = 0; d < 3; ++d) {
      A.w[d][n] = 0.0;
      A.mom[d][n] = 0.0;
      A.mass[d][n] = 0.0;
      A.vol[d][n] = 0.0;
    }
    if (A.F.mode == 1) {
      A.F.liquid[n] = 0;
      A.F.flags[n] = 0;
    }
    return;
  }
  int i, j, k;
  node_coords(g, n, i, j, k);
  const uint8_t* act = A.act;
  A.w[0][n] = clamp01(1.0 - fraction_inside4(corner_phi_w(g, act, i, j, k),
                                              corner_phi_w(g, act, i, j + 1, k),
                                              corner_phi_w(g, act, i, j, k + 1),
                                              corner_phi_w(g, act, i, j + 1, k + 1)));
  A.w[1][n] = clamp01(1.0 - fraction_inside4(corner_phi_w(g, act, i, j, k),
                                              corner_phi_w(g, act, i, j, k + 1),
                                              corner_phi_w(g, act, i + 1, j, k),
                                              corner_phi_w(g, act, i + 1, j, k + 1)));
  A.w[2][n] = clamp01(1.0 - fraction_inside4(corner_phi_w(g, act, i, j, k),
                                              corner_phi_w(g, act, i, j + 1, k),
                                              corner_phi_w(g, act, i + 1, j, k),
                                              corner_phi_w(g, act, i + 1, j + 1, k)));
  if (A.F.mode != 2) {
    const unsigned long long key = A.phikey[n];
    double phi = key == 0xFFFFFFFFFFFFFFFFull ? 3.0 * bucket_size : phi_from_key(key);
    // the stencil minimum never exceeds the default (phi < 3 buckets there)
    if (phi > 3.0 * bucket_size) phi = 3.0 * bucket_size;
    if (container_phi(g, node_pos(g, i, 0.5), node_pos(g, j, 0.5),
                      node_pos(g, k, 0.5)) < 0.0)
      phi = -0.5 * g.dx;
    A.lphi[n] = phi;
    if (A.F.mode == 1) A.F.liquid[n] = phi < 0.0;
  } else if ((A.F.liquid[n] != 0) != (A.lphi[n] < 0.0)) {
    fz_depart(A.F);
  }
  uint8_t has = 0;
  for (int d = 0; d < 3; ++d) {
    const double ws = A.det ? from_fix(A.mass[d][n], 0x1p-56) : A.mass[d][n];
    const double P = A.det ? A.fm * from_fix(A.mom[d][n], 0x1p-48) : A.mom[d][n];
    const double M = A.fm * ws;
    A.mass[d][n] = M;
    A.vol[d][n] = A.fvol * ws;
    const bool hm = fz_mass(A.F, d, n, M);
    if (A.F.mode == 2 && (((A.F.flags[n] >> d) & 1) != 0) != (M > 1e-20)) fz_depart(A.F);
    has |= (uint8_t)(hm ? 1u << d : 0u);
    A.mom[d][n] = hm ? P / M : 0.0;
  }
  if (A.F.mode == 1) A.F.flags[n] = has;  // the theta clamps are added later
}

struct StepArgs {
  Grid g;
  const uint8_t* act;
  const double* mass[3];
  const double* vel[3];
  double* rhs[3];
  double* invC[3];
  double* ustar[3];
  double gravity[3];
  double dt;
  FrozenView F;
};

// stepVelocity for liquid without solids (LinearizedImplicitEuler.cpp:
// 224-337): rhs = 0 - M g (-dt) + M u; C = M + Q*0 with Q = 1; invC = 1/C or
// 1; u* = rhs/M or rhs.
__global__ void k_step_velocity(StepArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= A.g.nnode) return;
  const bool on = A.act[n >> 6] != 0;
  for (int d = 0; d < 3; ++d) {
    if (!on) {
      A.rhs[d][n] = 0.0;
      A.invC[d][n] = 0.0;
      A.ustar[d][n] = 0.0;
      continue;
    }
    const double M = A.mass[d][n];
    double rhs = 0.0;
    rhs -= M * A.gravity[d] * (-A.dt);
    rhs += M * A.vel[d][n];
    A.rhs[d][n] = rhs;
    const double Q = 1.0, Ms = 0.0;
    const double Cc = M + Q * Ms;
    const bool hm = fz_mass(A.F, d, n, M);  // M > 1e-20, and so Cc
    A.invC[d][n] = hm ? 1.0 / Cc : 1.0;
    A.ustar[d][n] = hm ? rhs / M : rhs;
  }
}

// psi_fs = ((1-psi) invC + psi invCs P) M with psi = 0 (LIE:1052-1107)
__device__ inline double psi_fs(double invC, double M) {
  const double psi = 0.0, invCs = 0.0;
  const double P = M > 1e-20 ? 0.0 / M : 1.0;
  return ((1.0 - psi) * invC + psi * invCs * P) * M;
}
// the same with a solid: invCs = 0, so the finite P = hdV / (M + hdV) adds 0
__device__ inline double psi_fs_mix(double invC, double M, double psi) {
  const double invCs = 0.0, P = 0.0;
  return ((1.0 - psi) * invC + psi * invCs * P) * M;
}

struct DofArgs {
  Grid g;
  const uint8_t* act;
  const double* lphi;
  const double* w[3];
  uint32_t* isdof;
  FrozenView F;
};

// Pressure.cpp:348-392: a cell is a DOF if phi < 0 and one of its existing
// faces has positive solid weight.
__global__ void k_dof_flags(DofArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode) return;
  uint32_t f = 0;
  if (A.act[n >> 6] && fz_liquid(A.F, A.lphi, n)) {
    int i, j, k;
    node_coords(g, n, i, j, k);
    for (int face = 0; face < 6 && !f; ++face) {
      const int axis = face >> 1;
      int q[3] = {i, j, k};
      q[axis] += face & 1;
      if (!node_active(g, A.act, q[0], q[1], q[2])) continue;
      if (A.w[axis][node_index(g, q[0], q[1], q[2])] > 0) f = 1;
    }
  }
  A.isdof[n] = f;
}

struct AsmArgs {
  Grid g;
  const uint8_t* act;
  const double* lphi;
  const double* w[3];
  const double* mass[3];
  const double* vol[3];
  const double* invC[3];
  const double* ustar[3];
  const uint32_t* isdof;
  const uint32_t* dofscan;
  uint32_t* rowlen;
  uint32_t* col;
  double* val;
  double* b;
  int32_t* level;
  int32_t level_min;
  double dt;
  FrozenView F;
  // with a solid (MIX): volume fraction, solid volume and drag ratio Q of
  // every face
  const double* psi[3];
  const double* svol[3];
  const double* Q[3];
};

__device__ inline void insert_sorted(uint32_t* cols, double* vals, int& len,
                                     uint32_t c, double v) {
  // SparseMatrix::add_to_element (SparseMatrix.h:88-103)
  int k = 0;
  for (; k < len; ++k) {
    if (cols[k] == c) {
      vals[k] += v;
      return;
    }
    if (cols[k] > c) break;
  }
  for (int t = len; t > k; --t) {
    cols[t] = cols[t - 1];
    vals[t] = vals[t - 1];
  }
  cols[k] = c;
  vals[k] = v;
  ++len;
}

// Pressure.cpp:444-791: one row per DOF, faces left, right, bottom, top,
// near, far; rhs += or -= ((u* psi_fs + u_s psi_sf) w + u_solid (1-w)) / dx;
// open faces add term = w dt/dx^2 [(1-psi) invC (Vf + Vs Q) + ...] to the
// diagonal and -term to a liquid neighbour, or term/theta for an air one.
// The resting solid's velocity u_s is zero and invCs = 0 (constructMsDVs
// runs only when the solid is solved), so psi_sf and the psi invCs terms
// add zeros.  Without a solid psi = 0, Vs = 0, Q = 1.
template <bool MIX>
__global__ void k_assemble(AsmArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode || !A.isdof[n]) return;
  const uint32_t me = A.dofscan[n];
  int i, j, k;
  node_coords(g, n, i, j, k);
  const double dx = g.dx;
  const double coeff = A.dt / (dx * dx);
  const double center_phi = A.lphi[n];
  uint32_t cols[7];
  double vals[7];
  int len = 0;
  double rhs = 0.0;
  for (int f = 0; f < 6; ++f) {
    const int axis = f >> 1, hi = f & 1;
    int q[3] = {i, j, k};
    q[axis] += hi;
    if (!node_active(g, A.act, q[0], q[1], q[2])) continue;
    const int64_t fn = node_index(g, q[0], q[1], q[2]);
    const double w = A.w[axis][fn];
    const double M = A.mass[axis][fn];
    const double invC = A.invC[axis][fn];
    const double psi = MIX ? A.psi[axis][fn] : 0.0;
    const double uf = A.ustar[axis][fn], pfs = MIX ? psi_fs_mix(invC, M, psi) : psi_fs(invC, M);
    const double us = 0.0, psf = 0.0, usolid = 0.0;
    const double flux = ((uf * pfs + us * psf) * w + usolid * (1.0 - w)) / dx;
    if (hi)
      rhs -= flux;
    else
      rhs += flux;
    if (w > 0.0) {
      const double Vs = MIX ? A.svol[axis][fn] : 0.0, Q = MIX ? A.Q[axis][fn] : 1.0;
      const double invCs = 0.0;
      // P only meets invCs = 0: any finite value gives the same term
      const double P = MIX ? 0.0 : (M > 1e-20 ? 0.0 / M : 1.0);
      const double Vf = A.vol[axis][fn];
      double term = w * coeff;
      term *= (1.0 - psi) * invC * (Vf + Vs * Q) + psi * invCs * (Vs + Vf * P);
      int nb[3] = {i, j, k};
      nb[axis] += hi ? 1 : -1;
      if (!node_active(g, A.act, nb[0], nb[1], nb[2])) continue;
      const int64_t nn = node_index(g, nb[0], nb[1], nb[2]);
      const double nphi = A.lphi[nn];
      if (fz_liquid(A.F, A.lphi, nn)) {
        if (A.isdof[nn]) {
          insert_sorted(cols, vals, len, me, term);
          insert_sorted(cols, vals, len, A.dofscan[nn], -term);
        }
      } else {
        // the face between this (liquid) cell and the air cell nn
        const double theta = hi ? fz_theta(A.F, axis, fn, true, false, center_phi, nphi, nullptr)
                                : fz_theta(A.F, axis, fn, false, true, nphi, center_phi, nullptr);
        insert_sorted(cols, vals, len, me, term / theta);
      }
    }
  }
  A.b[me] = rhs;
  A.rowlen[me] = (uint32_t)len;
  for (int t = 0; t < len; ++t) {
    A.col[7 * (int64_t)me + t] = cols[t];
    A.val[7 * (int64_t)me + t] = vals[t];
  }
  A.level[me] = i + j + k - A.level_min;
}

__global__ void k_scatter_pressure(int64_t nnode, const uint32_t* isdof,
                                   const uint32_t* dofscan, const double* p,
                                   double* pressure) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  pressure[n] = isdof[n] ? p[dofscan[n]] : 0.0;
}

struct ApplyArgs {
  Grid g;
  const uint8_t* act;
  const double* lphi;
  const double* pressure;
  const double* w[3];
  const double* vol[3];
  const double* rhs[3];
  const double* invC[3];
  double* u[3];
  uint8_t* valid[3];
  double dt;
  FrozenView F;
  const double* svol[3];  // MIX: solid volume and drag ratio Q of every face
  const double* Q[3];
};

// applyPressureGradsFluidRHS (Pressure.cpp:1137-1314) after addSolidDragRHS
// (+Q Ms us = 0: the solid rests, LIE:689-717), then performInvLocalSolve
// u = rhs invC.  Without a solid Vs = 0 and Q = 1.
template <bool MIX>
__global__ void k_apply_pressure(ApplyArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode) return;
  if (!A.act[n >> 6]) {
    for (int d = 0; d < 3; ++d) {
      A.u[d][n] = 0.0;
      A.valid[d][n] = 0;
    }
    return;
  }
  int i, j, k;
  node_coords(g, n, i, j, k);
  const double dx = g.dx;
  uint8_t clamps = 0;
  for (int d = 0; d < 3; ++d) {
    const double Q = MIX ? A.Q[d][n] : 1.0, us = 0.0, Vs = MIX ? A.svol[d][n] : 0.0;
    double rhs = A.rhs[d][n];
    rhs += Q * 0.0 * us;  // Q M_s u_s with u_s = 0
    uint8_t valid = 0;
    if (A.w[d][n] > 0.) {
      int l[3] = {i, j, k};
      l[d] -= 1;
      const bool lex = node_active(g, A.act, l[0], l[1], l[2]);
      const int64_t ln = lex ? node_index(g, l[0], l[1], l[2]) : -1;
      const double left_phi = lex ? A.lphi[ln] : 0.0;
      const double right_phi = A.lphi[n];
      const bool lliq = lex && fz_liquid(A.F, A.lphi, ln), rliq = fz_liquid(A.F, A.lphi, n);
      if (lliq || rliq) {
        double theta = 1.;
        if (!lliq || !rliq)
          theta = fz_theta(A.F, d, n, lliq, rliq, left_phi, right_phi, &clamps, true);
        double pg = 0;
        if (lex) pg -= A.pressure[ln];
        pg += A.pressure[n];
        pg /= (dx * theta);
        const double term = pg * A.dt * (A.vol[d][n] + Vs * Q);
        rhs -= term;
        valid = 1;
      }
    }
    A.u[d][n] = rhs * A.invC[d][n];
    A.valid[d][n] = valid;
  }
  if (A.F.mode == 1) A.F.flags[n] = (uint8_t)((A.F.flags[n] & 7u) | clamps);
}

struct ExtrapArgs {
  Grid g;
  const uint8_t* act;
  const double* uin[3];
  const uint8_t* vin[3];
  double* uout[3];
  uint8_t* vout[3];
  const double* w[3];
  double friction;
  int constrain;
};

// one extrapolation layer (Pressure.cpp:1826-1903) reading the previous
// layer's copies; the last layer also applies constrainLiquidVelocity
// (TwoDScene.cpp:1827-1866).
__global__ void k_extrapolate(ExtrapArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode) return;
  if (!A.act[n >> 6]) {
    for (int d = 0; d < 3; ++d) {
      A.uout[d][n] = 0.0;
      A.vout[d][n] = 0;
    }
    return;
  }
  int i, j, k;
  node_coords(g, n, i, j, k);
  for (int d = 0; d < 3; ++d) {
    double u = A.uin[d][n];
    uint8_t valid = A.vin[d][n];
    if (!valid) {
      double sum = 0;
      int count = 0;
      for (int r = 0; r < 6; ++r) {
        int q[3] = {i, j, k};
        q[r >> 1] += (r & 1) ? 1 : -1;
        if (!node_active(g, A.act, q[0], q[1], q[2])) continue;
        const int64_t m = node_index(g, q[0], q[1], q[2]);
        if (!A.vin[d][m]) continue;
        sum += A.uin[d][m];
        ++count;
      }
      if (count > 0) {
        u = sum / (double)count;
        valid = 1;
      }
    }
    if (A.constrain && A.w[d][n] == 0.0) {
      const double fr = A.friction, usolid = 0.0;
      u = (1 - fr) * u + fr * usolid;
    }
    A.uout[d][n] = u;
    A.vout[d][n] = valid;
  }
}

}  // namespace

void grid_activate(wc_ctx* ctx) {
  const Grid& g = ctx->g;
  if (ctx->frz_mode == 2) {  // the recorded activation
    WC_HIP(hipMemcpyAsync(ctx->act.p, frozen_record(ctx).act.p, g.nbucket,
                          hipMemcpyDeviceToDevice, ctx->stream));
    return;
  }
  hipLaunchKernelGGL(k_activate, dim3(blocks(g.nbucket, 256)), dim3(256), 0,
                     ctx->stream, g, ctx->touched.p, ctx->act.p);
  WC_LAUNCH_CHECK();
  if (ctx->frz_mode == 1)
    WC_HIP(hipMemcpyAsync(frozen_record(ctx).act.p, ctx->act.p, g.nbucket,
                          hipMemcpyDeviceToDevice, ctx->stream));
}

void grid_nodes(wc_ctx* ctx) {
  if (ctx->frz_mode == 2) frozen_replay_lphi(ctx);
  NodeArgs A;
  A.F = frozen_view(ctx);
  A.g = ctx->g;
  A.act = ctx->act.p;
  A.phikey = ctx->phikey.p;
  A.lphi = ctx->lphi.p;
  for (int d = 0; d < 3; ++d) {
    A.w[d] = ctx->w[d].p;
    A.mass[d] = ctx->mass[d].p;
    A.vol[d] = ctx->vol[d].p;
    A.mom[d] = ctx->mom[d].p;
  }
  // the mixed form's P2G weights with the FP32 particle mass and volume
  A.fm = ctx->mixed ? (double)(float)ctx->desc.particle_mass : ctx->desc.particle_mass;
  A.fvol = ctx->mixed ? (double)(float)ctx->desc.particle_volume : ctx->desc.particle_volume;
  A.det = ctx->deterministic ? 1 : 0;
  hipLaunchKernelGGL(k_grid_nodes, dim3(blocks(ctx->g.nnode, 256)), dim3(256),
                     0, ctx->stream, A);
  WC_LAUNCH_CHECK();
  if (ctx->frz_mode == 1) frozen_record_lphi_arg(ctx);
}

void grid_step_velocity(wc_ctx* ctx, double dt) {
  if (ctx->mix.ne > 0) {
    mix_step_velocity(ctx, dt);
    return;
  }
  StepArgs A;
  A.g = ctx->g;
  A.act = ctx->act.p;
  for (int d = 0; d < 3; ++d) {
    A.mass[d] = ctx->mass[d].p;
    A.vel[d] = ctx->mom[d].p;
    A.rhs[d] = ctx->rhs[d].p;
    A.invC[d] = ctx->invC[d].p;
    A.ustar[d] = ctx->ustar[d].p;
    A.gravity[d] = ctx->desc.gravity[d];
  }
  A.dt = dt;
  A.F = frozen_view(ctx);
  hipLaunchKernelGGL(k_step_velocity, dim3(blocks(ctx->g.nnode, 256)),
                     dim3(256), 0, ctx->stream, A);
  WC_LAUNCH_CHECK();
}

int64_t grid_assemble(wc_ctx* ctx, double dt) {
  const Grid& g = ctx->g;
  hipStream_t st = ctx->stream;
  ctx->isdof.ensure(g.nnode + 1);
  ctx->dofscan.ensure(g.nnode + 1);
  DofArgs D;
  D.g = g;
  D.act = ctx->act.p;
  D.lphi = ctx->lphi.p;
  for (int d = 0; d < 3; ++d) D.w[d] = ctx->w[d].p;
  D.isdof = ctx->isdof.p;
  D.F = frozen_view(ctx);
  hipLaunchKernelGGL(k_dof_flags, dim3(blocks(g.nnode, 256)), dim3(256), 0, st, D);
  WC_LAUNCH_CHECK();
  WC_HIP(hipMemsetAsync(ctx->isdof.p + g.nnode, 0, sizeof(uint32_t), st));
  WC_HIP(hipMemcpyAsync(ctx->dofscan.p, ctx->isdof.p,
                        sizeof(uint32_t) * (g.nnode + 1), hipMemcpyDeviceToDevice, st));
  const int64_t ndof = exclusive_scan(ctx, ctx->dofscan.p, g.nnode + 1);
  ctx->ndof = ndof;
  ctx->rowlen.ensure(ndof);
  ctx->col.ensure(7 * ndof);
  ctx->val.ensure(7 * ndof);
  ctx->b.ensure(ndof);
  ctx->p.ensure(ndof);
  ctx->dof_level.ensure(ndof);
  ctx->level_min = 0;
  ctx->level_max = g.nx[0] + g.nx[1] + g.nx[2];
  if (ndof > 0) {
    AsmArgs A;
    A.g = g;
    A.act = ctx->act.p;
    A.lphi = ctx->lphi.p;
    for (int d = 0; d < 3; ++d) {
      A.w[d] = ctx->w[d].p;
      A.mass[d] = ctx->mass[d].p;
      A.vol[d] = ctx->vol[d].p;
      A.invC[d] = ctx->invC[d].p;
      A.ustar[d] = ctx->ustar[d].p;
    }
    A.isdof = ctx->isdof.p;
    A.dofscan = ctx->dofscan.p;
    A.rowlen = ctx->rowlen.p;
    A.col = ctx->col.p;
    A.val = ctx->val.p;
    A.b = ctx->b.p;
    A.level = ctx->dof_level.p;
    A.level_min = ctx->level_min;
    A.dt = dt;
    A.F = frozen_view(ctx);
    const bool mix = ctx->mix.ne > 0;
    for (int d = 0; d < 3; ++d) {
      A.psi[d] = mix ? ctx->mix.psi[d].p : nullptr;
      A.svol[d] = mix ? ctx->mix.plane(g, d, kMixPlaneSvol) : nullptr;
      A.Q[d] = mix ? ctx->mix.Q[d].p : nullptr;
    }
    if (mix)
      hipLaunchKernelGGL(k_assemble<true>, dim3(blocks(g.nnode, 256)), dim3(256), 0, st, A);
    else
      hipLaunchKernelGGL(k_assemble<false>, dim3(blocks(g.nnode, 256)), dim3(256), 0, st, A);
    WC_LAUNCH_CHECK();
  }
  return ndof;
}

namespace {
__global__ void k_dof_cells(Grid g, const uint32_t* isdof, const uint32_t* dofscan,
                            int32_t* ijk, int o0, int o1, int o2) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= g.nnode || !isdof[n]) return;
  int i, j, k;
  node_coords(g, n, i, j, k);
  const int64_t d = dofscan[n];
  ijk[3 * d] = i - o0;
  ijk[3 * d + 1] = j - o1;
  ijk[3 * d + 2] = k - o2;
}
}  // namespace

void grid_dof_cells(wc_ctx* ctx, int32_t* ijk, const int origin[3]) {
  const Grid& g = ctx->g;
  hipLaunchKernelGGL(k_dof_cells, dim3(blocks(g.nnode, 256)), dim3(256), 0, ctx->stream, g,
                     ctx->isdof.p, ctx->dofscan.p, ijk, origin[0], origin[1], origin[2]);
  WC_LAUNCH_CHECK();
}

void grid_scatter_pressure(wc_ctx* ctx) {
  const Grid& g = ctx->g;
  if (ctx->ndof == 0) {
    WC_HIP(hipMemsetAsync(ctx->pressure.p, 0, sizeof(double) * g.nnode, ctx->stream));
    return;
  }
  hipLaunchKernelGGL(k_scatter_pressure, dim3(blocks(g.nnode, 256)), dim3(256),
                     0, ctx->stream, g.nnode, ctx->isdof.p, ctx->dofscan.p,
                     ctx->p.p, ctx->pressure.p);
  WC_LAUNCH_CHECK();
}

void grid_apply_pressure(wc_ctx* ctx, double dt) {
  const Grid& g = ctx->g;
  hipStream_t st = ctx->stream;
  ApplyArgs A;
  A.g = g;
  A.act = ctx->act.p;
  A.lphi = ctx->lphi.p;
  A.pressure = ctx->pressure.p;
  for (int d = 0; d < 3; ++d) {
    A.w[d] = ctx->w[d].p;
    A.vol[d] = ctx->vol[d].p;
    A.rhs[d] = ctx->rhs[d].p;
    A.invC[d] = ctx->invC[d].p;
    A.u[d] = ctx->unew[d].p;
    A.valid[d] = ctx->valid[d].p;
  }
  A.dt = dt;
  A.F = frozen_view(ctx);
  const bool mix = ctx->mix.ne > 0;
  for (int d = 0; d < 3; ++d) {
    A.svol[d] = mix ? ctx->mix.plane(g, d, kMixPlaneSvol) : nullptr;
    A.Q[d] = mix ? ctx->mix.Q[d].p : nullptr;
  }
  if (mix)
    hipLaunchKernelGGL(k_apply_pressure<true>, dim3(blocks(g.nnode, 256)), dim3(256), 0, st, A);
  else
    hipLaunchKernelGGL(k_apply_pressure<false>, dim3(blocks(g.nnode, 256)), dim3(256), 0, st,
                       A);
  WC_LAUNCH_CHECK();
  // layer 1: unew -> utmp; layer 2 (+ constraint): utmp -> unew
  for (int layer = 0; layer < 2; ++layer) {
    ExtrapArgs E;
    E.g = g;
    E.act = ctx->act.p;
    for (int d = 0; d < 3; ++d) {
      E.uin[d] = layer == 0 ? ctx->unew[d].p : ctx->utmp[d].p;
      E.vin[d] = layer == 0 ? ctx->valid[d].p : ctx->vtmp[d].p;
      E.uout[d] = layer == 0 ? ctx->utmp[d].p : ctx->unew[d].p;
      E.vout[d] = layer == 0 ? ctx->vtmp[d].p : ctx->valid[d].p;
      E.w[d] = ctx->w[d].p;
    }
    E.friction = ctx->desc.liquid_boundary_friction;
    E.constrain = layer == 1;
    hipLaunchKernelGGL(k_extrapolate, dim3(blocks(g.nnode, 256)), dim3(256), 0,
                       st, E);
    WC_LAUNCH_CHECK();
  }
}

}  // namespace wc
