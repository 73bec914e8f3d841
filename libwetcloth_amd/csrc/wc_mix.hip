// The mixture terms of a static porous solid (the cloth's vertices with
// LiquidInfo::solve_solid off) in the liquid path, for gfx950.
//
// The solid rests, so its P2G sums (TwoDScene.cpp:6686-6736: mass, solid
// and carried liquid volume, shape factor with its weight sum, orientation)
// and its part of the cell-centred volumes (:5952-5992) are scattered once,
// when the solid is set, with FP64 atomics into 8 planes per face axis and
// 2 cell planes.  Each substep then
//   * adds the buckets the solid's five stencils touch to the activation,
//   * scatters the cell-centred volume of the liquid particles whose cell
//     stencil can reach a cell with solid (buckets next to the solid's),
//   * forms the saturation of every face and the pore pressure of every
//     cell (one thread per node, bucket-major like the other grid sweeps),
//   * and runs the grid update with the pore pressure gradient and the drag
//     (stepVelocity, LinearizedImplicitEuler.cpp:224-337).
// The volume-fraction-weighted assembly and pressure application are the
// MIX forms of the grid kernels (wc_grid.hip).  Every expression follows
// the reference's operation order, so the FP64 form agrees with the
// sequential one up to the order of the atomic sums.
#include <math.h>

#include <algorithm>
#include <vector>

#include "wc_ctx.h"

namespace wc {
namespace {

inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

__device__ inline double clamp01(double a) { return a < 0.0 ? 0.0 : (a > 1.0 ? 1.0 : a); }

// The planes of the solid's face sums: plane(d, k) of nnode doubles.
// k: 0 solid mass sum (m + fm) w, 1 solid volume sum rest_vol rest_vf w,
// 2 carried liquid volume sum fvol w, 3 shape factor sum sf w (the shape
// factor after the static pass), 4 weight sum (the solid volume, solid +
// carried liquid, after it), 5..7 orientation sum (normalised after it).
constexpr int kPlanes = 8;
enum { P_MASS = 0, P_VSOL = 1, P_VFE = 2, P_SF = 3, P_RW = 4, P_SVOL = 4, P_ORI = 5 };

struct SolidArgs {
  Grid g;
  int64_t ne;
  const double *x, *m, *fm, *fvol, *rvol, *rvf, *ori, *sf;
  double* acc;  // kPlanes * 3 planes
  double *cvs, *cvf;
  uint8_t* touch;
};

// one thread per solid particle: the scatter of the solid's branch of
// mapParticleNodesAPIC and of mapParticleSaturationPsiNodes, and the
// buckets its five stencils touch (findNodes, TwoDScene.cpp:3683-3752)
__global__ __launch_bounds__(256) void k_solid_scatter(SolidArgs A) {
  const int64_t p = blockIdx.x * 256ll + threadIdx.x;
  if (p >= A.ne) return;
  const Grid& g = A.g;
  const int64_t nn = g.nnode, ne = A.ne;
  const double px[3] = {A.x[p], A.x[ne + p], A.x[2 * ne + p]};
  const double mass = A.m[p] + A.fm[p];
  const double vsol = A.rvol[p] * A.rvf[p];
  const double fv = A.fvol[p], sf = A.sf[p];
  const double o[3] = {A.ori[p], A.ori[ne + p], A.ori[2 * ne + p]};
  for (int d = 0; d < 3; ++d) {
    int base[3];
    double w[3][3];
    for (int a = 0; a < 3; ++a) {
      base[a] = stencil_base(g, a, px[a], face_off(d, a));
      stencil_weights(g, px[a], base[a], face_off(d, a), w[a]);
    }
    double* P = A.acc + (int64_t)d * kPlanes * nn;
    for (int c2 = 0; c2 < 3; ++c2)
      for (int c1 = 0; c1 < 3; ++c1)
        for (int c0 = 0; c0 < 3; ++c0) {
          const int64_t n = node_index(g, base[0] + c0, base[1] + c1, base[2] + c2);
          const double ww = w[0][c0] * w[1][c1] * w[2][c2];
          unsafeAtomicAdd(&P[P_MASS * nn + n], mass * ww);
          unsafeAtomicAdd(&P[P_VSOL * nn + n], vsol * ww);
          unsafeAtomicAdd(&P[P_VFE * nn + n], fv * ww);
          unsafeAtomicAdd(&P[P_SF * nn + n], sf * ww);
          unsafeAtomicAdd(&P[P_RW * nn + n], ww);
          for (int c = 0; c < 3; ++c) unsafeAtomicAdd(&P[(P_ORI + c) * nn + n], o[c] * ww);
        }
  }
  {
    int base[3];
    double w[3][3];
    for (int a = 0; a < 3; ++a) {
      base[a] = stencil_base(g, a, px[a], 0.5);
      stencil_weights(g, px[a], base[a], 0.5, w[a]);
    }
    for (int c2 = 0; c2 < 3; ++c2)
      for (int c1 = 0; c1 < 3; ++c1)
        for (int c0 = 0; c0 < 3; ++c0) {
          const int64_t n = node_index(g, base[0] + c0, base[1] + c1, base[2] + c2);
          const double ww = w[0][c0] * w[1][c1] * w[2][c2];
          unsafeAtomicAdd(&A.cvf[n], fv * ww);
          unsafeAtomicAdd(&A.cvs[n], vsol * ww);
        }
  }
  // the u, v, w faces, the corners and the cell centres
  for (int gi = 0; gi < 5; ++gi) {
    int blo[3], bhi[3];
    for (int a = 0; a < 3; ++a) {
      const double off = gi < 3 ? face_off(gi, a) : (gi == 3 ? 0.0 : 0.5);
      const int b = stencil_base(g, a, px[a], off);
      blo[a] = b >> 2;
      bhi[a] = (b + 2) >> 2;
    }
    for (int bk = blo[2]; bk <= bhi[2]; ++bk)
      for (int bj = blo[1]; bj <= bhi[1]; ++bj)
        for (int bi = blo[0]; bi <= bhi[0]; ++bi) A.touch[bucket_index(g, bi, bj, bk)] = 1;
  }
}

// the static per-face fields from the sums: shape factor (divided by the
// weight sum above 1e-20), solid volume (solid + carried liquid), volume
// fraction, orientation normalised above 1e-20 (TwoDScene.cpp:6715-6736)
__global__ void k_solid_static(Grid g, double* acc, double* psi0, double* psi1, double* psi2) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= g.nnode) return;
  const int64_t nn = g.nnode;
  const double dx = g.dx, dV = dx * dx * dx;
  double* psi[3] = {psi0, psi1, psi2};
  for (int d = 0; d < 3; ++d) {
    double* P = acc + (int64_t)d * kPlanes * nn;
    const double rw = P[P_RW * nn + n];
    if (rw > 1e-20) P[P_SF * nn + n] /= rw;
    const double vsol = P[P_VSOL * nn + n], vfe = P[P_VFE * nn + n];
    P[P_SVOL * nn + n] = vsol + vfe;
    psi[d][n] = clamp01(vsol / dV);
    const double o0 = P[P_ORI * nn + n], o1 = P[(P_ORI + 1) * nn + n],
                 o2 = P[(P_ORI + 2) * nn + n];
    const double lo = sqrt(o0 * o0 + o1 * o1 + o2 * o2);
    if (lo > 1e-20) {
      P[P_ORI * nn + n] = o0 / lo;
      P[(P_ORI + 1) * nn + n] = o1 / lo;
      P[(P_ORI + 2) * nn + n] = o2 / lo;
    }
  }
}

// the buckets next to (or at) one the solid touches
__global__ void k_near(Grid g, const uint8_t* touch, uint8_t* near) {
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
        if (x < 0 || y < 0 || z < 0 || x >= g.nb[0] || y >= g.nb[1] || z >= g.nb[2]) continue;
        if (touch[bucket_index(g, x, y, z)]) a = 1;
      }
  near[b] = a;
}

__global__ void k_or_touch(int64_t nb, const uint8_t* touch, uint8_t* touched) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b < nb && touch[b]) touched[b] = 1;
}

// The liquid particles' part of the cell-centred volume (their fluid
// volume; they carry no solid volume), for particles in buckets next to
// the solid's: only cells with solid volume have a pore pressure, and
// those cells' stencils reach no further.
template <class R>
__global__ __launch_bounds__(256) void k_liquid_cells(Grid g, const R* x, int64_t np, int64_t cap,
                                                      const uint8_t* near, double pvol,
                                                      double* cvl) {
  const int64_t p = blockIdx.x * 256ll + threadIdx.x;
  if (p >= np) return;
  const R px[3] = {x[p], x[cap + p], x[2 * cap + p]};
  int bc[3];
  for (int a = 0; a < 3; ++a) {
    int c = (int)floor(px[a] / (R)g.dx) + kGhost;
    c = c < 0 ? 0 : (c > g.nx[a] - 1 ? g.nx[a] - 1 : c);
    bc[a] = c >> 2;
  }
  if (!near[bucket_index(g, bc[0], bc[1], bc[2])]) return;
  int base[3];
  R w[3][3];
  for (int a = 0; a < 3; ++a) {
    base[a] = stencil_base_r<R>(g, a, px[a], (R)0.5);
    stencil_weights_r<R>(g, px[a], base[a], (R)0.5, w[a]);
  }
  for (int c2 = 0; c2 < 3; ++c2)
    for (int c1 = 0; c1 < 3; ++c1)
      for (int c0 = 0; c0 < 3; ++c0) {
        const R ww = w[0][c0] * w[1][c1] * w[2][c2];
        unsafeAtomicAdd(&cvl[node_index(g, base[0] + c0, base[1] + c1, base[2] + c2)],
                        pvol * (double)ww);
      }
}

// TwoDScene.cpp:563-571 (cos of the contact angle taken on the host)
__device__ inline double capillary_pressure(const wc_mix_params& p, double cos_theta,
                                            double psi) {
  if (1.0 - psi < 1e-20 || p.pore_radius < 1e-20) return 0.0;
  const double alpha = psi / (1. - psi);
  return alpha * p.surf_tension_coeff * cos_theta / p.pore_radius;
}

struct MixNodeArgs {
  Grid g;
  const uint8_t* act;
  const double* vol[3];  // the liquid's face volume
  const double* acc;
  const double *cvs, *cvf, *cvl;
  double* sat[3];
  double *psi_p, *sat_p, *pp;
  wc_mix_params prm;
  double cos_theta;
};

// the saturation of every face (TwoDScene.cpp:6727-6729) and the cell-centred
// volume fraction, saturation and pore pressure (:5933-5992)
__global__ void k_mix_nodes(MixNodeArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode) return;
  const int64_t nn = g.nnode;
  const double dx = g.dx, dV = dx * dx * dx;
  for (int d = 0; d < 3; ++d) {
    const double* P = A.acc + (int64_t)d * kPlanes * nn;
    const double vsol = P[P_VSOL * nn + n], vfe = P[P_VFE * nn + n];
    const double room = dV - vsol > 1e-20 ? dV - vsol : 1e-20;
    A.sat[d][n] = clamp01((A.vol[d][n] + vfe) / room);
  }
  const double vs = A.cvs[n], vf = A.cvf[n] + A.cvl[n];
  const double psi = clamp01(vs / dV);
  const double room = dV - vs > 1e-20 ? dV - vs : 1e-20;
  const double sat = clamp01(vf / room);
  A.psi_p[n] = psi;
  A.sat_p[n] = sat;
  A.pp[n] = capillary_pressure(A.prm, A.cos_theta, psi) * (1.0 - sat);
}

// mathutils::get_rotated_drag_x/y/z and get_rotated_pore_coeff_x/y/z
// (MathUtilities.h:792-878): row `index` of t I + (u - t) n n^T, n = a/|a|,
// summed: t + (u - t) n_index (n_x + n_y + n_z); zero for a = 0.
__device__ inline double iso_row_sum(const double a[3], double t, double u, int index) {
  const double s2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  if (s2 == 0.0) return 0.0;
  const double r = sqrt(s2);
  const double n[3] = {a[0] / r, a[1] / r, a[2] / r};
  return t + (u - t) * n[index] * (n[0] + n[1] + n[2]);
}

// TwoDScene::getDragCoeffWithOrientation (TwoDScene.cpp:4042-4097) for the
// liquid (material 0)
__device__ inline double drag_coeff(const wc_mix_params& p, double rho, double psi, double s,
                                    double dv, const double ori[3], double sf, int index) {
  if (!p.use_drag || psi == 0.0 || s == 0.0 ||
      ori[0] * ori[0] + ori[1] * ori[1] + ori[2] * ori[2] < 1e-20)
    return 0.0;
  const double ergun_coeff = p.use_nonlinear_drag ? 0.1428869017 : 0.0;
  const double di = p.yarn_diameter;
  double ka = (-log(psi) - 1.476 + 2.0 * psi - 0.5 * psi * psi) / (16.0 * psi) * di * di;
  if (ka < 1e-20) ka = 1e-20;
  double kb = (-log(psi) - 1.476 + 2.0 * psi - 1.774 * psi * psi + 4.078 * pow(psi, 3.0)) /
              (32.0 * psi) * di * di;
  if (kb < 1e-20) kb = 1e-20;
  const double mu = p.viscosity, pw = p.yazdchi_power;
  double ca = mu / ka + ergun_coeff * pow(di, pw - 1.0) * pow(mu, 1.0 - pw) /
                            (pow(1.0 - psi, 1.5) * sqrt(ka)) * pow(rho * fabs(dv), pw);
  if (ca > 1e+63) ca = 1e+63;
  double cb = mu / kb + ergun_coeff * pow(di, pw - 1.0) * pow(mu, 1.0 - pw) /
                            (pow(1.0 - psi, 1.5) * sqrt(kb)) * pow(rho * fabs(dv), pw);
  if (cb > 1e+63) cb = 1e+63;
  const double c_xy = ca * (1.0 - sf) + cb * sf;
  const double c_z = ca * sf + cb * (1.0 - sf);
  return iso_row_sum(ori, c_xy, c_z, index);
}

struct MixStepArgs {
  Grid g;
  const uint8_t* act;
  const double* mass[3];
  const double* vel[3];
  const double* vol[3];
  const double* w[3];
  const double* psi[3];
  const double* sat[3];
  const double* acc;
  const double* pp;
  double* rhs[3];
  double* invC[3];
  double* ustar[3];
  double* Q[3];
  double* hdv[3];
  double gravity[3];
  double dt, rho;
  wc_mix_params prm;
};

// stepVelocity with the solid at rest and not solved (LinearizedImplicitEuler
// .cpp:224-337): constructNodeForce (:548-659) takes gravity, the pore
// pressure gradient (Pressure.cpp:26-142) and M u; constructHDV (:846-978,
// no drag by air: the damped mass is the solid mass) the drag of the
// liquid's P2G velocity against the resting solid; constructInvMDV
// (:1109-1148) invC = 1 / (M_f + Q M_s); performLocalSolve (:128-175)
// u* = rhs / M_f.  psi_fs and psi_sf (:1052-1107, invCs = 0) are formed
// where they are used.
__global__ void k_step_velocity_mix(MixStepArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode) return;
  const int64_t nn = g.nnode;
  if (!A.act[n >> 6]) {
    for (int d = 0; d < 3; ++d) {
      A.rhs[d][n] = A.invC[d][n] = A.ustar[d][n] = 0.0;
      A.Q[d][n] = A.hdv[d][n] = 0.0;
    }
    return;
  }
  int i, j, k;
  node_coords(g, n, i, j, k);
  const double dx = g.dx, dt = A.dt;
  for (int d = 0; d < 3; ++d) {
    const double* P = A.acc + (int64_t)d * kPlanes * nn;
    const double a[3] = {P[P_ORI * nn + n], P[(P_ORI + 1) * nn + n], P[(P_ORI + 2) * nn + n]};
    const double sfac = P[P_SF * nn + n];
    const double Mf = A.mass[d][n];
    double rhs = 0.0;
    rhs -= Mf * A.gravity[d] * (-dt);
    const double vf = A.vol[d][n], wv = A.w[d][n];
    if (vf > 0.0 && wv > 0.0) {
      int l[3] = {i, j, k};
      l[d] -= 1;
      const double pp0 = node_active(g, A.act, l[0], l[1], l[2])
                             ? A.pp[node_index(g, l[0], l[1], l[2])]
                             : 0.0;
      const double pp1 = A.pp[n];
      const double grad = (pp1 - pp0) / dx;
      rhs += grad * vf * dt * A.sat[d][n] * wv * iso_row_sum(a, 2.0 - sfac, 1.0 + sfac, d);
    }
    rhs += Mf * A.vel[d][n];
    A.rhs[d][n] = rhs;
    const double dc =
        drag_coeff(A.prm, A.rho, A.psi[d][n], A.sat[d][n], A.vel[d][n] - 0.0, a, sfac, d);
    const double Vm = P[P_SVOL * nn + n] + vf;
    const double hdV = dt * dc * Vm;
    const double damped = P[P_MASS * nn + n];
    const double mshdvm = damped + hdV;
    const double Q = mshdvm > 1e-20 ? hdV / mshdvm : 1.0;
    A.hdv[d][n] = hdV;
    A.Q[d][n] = Q;
    const double Cc = Mf + Q * damped;
    A.invC[d][n] = Cc > 1e-20 ? 1.0 / Cc : 1.0;
    A.ustar[d][n] = Mf > 1e-20 ? rhs / Mf : rhs;
  }
}

}  // namespace

void MixState::release() {
  x.release();
  m.release();
  fm.release();
  fvol.release();
  rvol.release();
  rvf.release();
  ori.release();
  sf.release();
  acc.release();
  cvs.release();
  cvf.release();
  cvl.release();
  touch.release();
  near.release();
  for (int d = 0; d < 3; ++d) {
    psi[d].release();
    sat[d].release();
    Q[d].release();
    hdv[d].release();
  }
  psi_p.release();
  sat_p.release();
  pp.release();
  ne = 0;
}

double* MixState::plane(const Grid& g, int d, int k) const {
  return acc.p + ((int64_t)d * kPlanes + k) * g.nnode;
}

void mix_set(wc_ctx* ctx, int64_t ne, const double* x, const double* m, const double* fm,
             const double* fvol, const double* rvol, const double* rvf, const double* ori,
             const double* sf, const wc_mix_params& prm) {
  MixState& M = ctx->mix;
  M.release();
  if (ne <= 0) return;
  const Grid& g = ctx->g;
  const int64_t nn = g.nnode;
  hipStream_t st = ctx->stream;
  M.prm = prm;
  auto up = [&](DevBuf<double>& b, const double* src, int64_t cnt) {
    b.ensure(cnt);
    WC_HIP(hipMemcpyAsync(b.p, src, sizeof(double) * cnt, hipMemcpyHostToDevice, st));
  };
  up(M.x, x, 3 * ne);
  up(M.ori, ori, 3 * ne);
  up(M.m, m, ne);
  up(M.fm, fm, ne);
  up(M.fvol, fvol, ne);
  up(M.rvol, rvol, ne);
  up(M.rvf, rvf, ne);
  up(M.sf, sf, ne);
  M.acc.ensure(3 * kPlanes * nn);
  M.cvs.ensure(nn);
  M.cvf.ensure(nn);
  M.cvl.ensure(nn);
  M.touch.ensure(g.nbucket);
  M.near.ensure(g.nbucket);
  for (int d = 0; d < 3; ++d) {
    M.psi[d].ensure(nn);
    M.sat[d].ensure(nn);
    M.Q[d].ensure(nn);
    M.hdv[d].ensure(nn);
    WC_HIP(hipMemsetAsync(M.sat[d].p, 0, sizeof(double) * nn, st));
    WC_HIP(hipMemsetAsync(M.Q[d].p, 0, sizeof(double) * nn, st));
    WC_HIP(hipMemsetAsync(M.hdv[d].p, 0, sizeof(double) * nn, st));
  }
  M.psi_p.ensure(nn);
  M.sat_p.ensure(nn);
  M.pp.ensure(nn);
  WC_HIP(hipMemsetAsync(M.acc.p, 0, sizeof(double) * 3 * kPlanes * nn, st));
  WC_HIP(hipMemsetAsync(M.cvs.p, 0, sizeof(double) * nn, st));
  WC_HIP(hipMemsetAsync(M.cvf.p, 0, sizeof(double) * nn, st));
  WC_HIP(hipMemsetAsync(M.cvl.p, 0, sizeof(double) * nn, st));
  WC_HIP(hipMemsetAsync(M.psi_p.p, 0, sizeof(double) * nn, st));
  WC_HIP(hipMemsetAsync(M.sat_p.p, 0, sizeof(double) * nn, st));
  WC_HIP(hipMemsetAsync(M.pp.p, 0, sizeof(double) * nn, st));
  WC_HIP(hipMemsetAsync(M.touch.p, 0, g.nbucket, st));
  SolidArgs A;
  A.g = g;
  A.ne = ne;
  A.x = M.x.p;
  A.m = M.m.p;
  A.fm = M.fm.p;
  A.fvol = M.fvol.p;
  A.rvol = M.rvol.p;
  A.rvf = M.rvf.p;
  A.ori = M.ori.p;
  A.sf = M.sf.p;
  A.acc = M.acc.p;
  A.cvs = M.cvs.p;
  A.cvf = M.cvf.p;
  A.touch = M.touch.p;
  hipLaunchKernelGGL(k_solid_scatter, dim3(blocks(ne, 256)), dim3(256), 0, st, A);
  WC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_solid_static, dim3(blocks(nn, 256)), dim3(256), 0, st, g, M.acc.p,
                     M.psi[0].p, M.psi[1].p, M.psi[2].p);
  WC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_near, dim3(blocks(g.nbucket, 256)), dim3(256), 0, st, g, M.touch.p,
                     M.near.p);
  WC_LAUNCH_CHECK();
  WC_HIP(hipStreamSynchronize(st));
  M.ne = ne;
}

void mix_touch(wc_ctx* ctx) {
  const Grid& g = ctx->g;
  hipLaunchKernelGGL(k_or_touch, dim3(blocks(g.nbucket, 256)), dim3(256), 0, ctx->stream,
                     g.nbucket, ctx->mix.touch.p, ctx->touched.p);
  WC_LAUNCH_CHECK();
}

void mix_grid(wc_ctx* ctx) {
  MixState& M = ctx->mix;
  const Grid& g = ctx->g;
  hipStream_t st = ctx->stream;
  WC_HIP(hipMemsetAsync(M.cvl.p, 0, sizeof(double) * g.nnode, st));
  if (ctx->np > 0) {
    if (ctx->mixed)
      hipLaunchKernelGGL(k_liquid_cells<float>, dim3(blocks(ctx->np, 256)), dim3(256), 0, st, g,
                         (const float*)ctx->fx.p, ctx->np, ctx->pcap, (const uint8_t*)M.near.p,
                         (double)(float)ctx->desc.particle_volume, M.cvl.p);
    else
      hipLaunchKernelGGL(k_liquid_cells<double>, dim3(blocks(ctx->np, 256)), dim3(256), 0, st,
                         g, (const double*)ctx->x.p, ctx->np, ctx->pcap,
                         (const uint8_t*)M.near.p, ctx->desc.particle_volume, M.cvl.p);
    WC_LAUNCH_CHECK();
  }
  MixNodeArgs A;
  A.g = g;
  A.act = ctx->act.p;
  for (int d = 0; d < 3; ++d) {
    A.vol[d] = ctx->vol[d].p;
    A.sat[d] = M.sat[d].p;
  }
  A.acc = M.acc.p;
  A.cvs = M.cvs.p;
  A.cvf = M.cvf.p;
  A.cvl = M.cvl.p;
  A.psi_p = M.psi_p.p;
  A.sat_p = M.sat_p.p;
  A.pp = M.pp.p;
  A.prm = M.prm;
  A.cos_theta = cos(M.prm.rest_contact_angle);
  hipLaunchKernelGGL(k_mix_nodes, dim3(blocks(g.nnode, 256)), dim3(256), 0, st, A);
  WC_LAUNCH_CHECK();
}

void mix_step_velocity(wc_ctx* ctx, double dt) {
  MixState& M = ctx->mix;
  MixStepArgs A;
  A.g = ctx->g;
  A.act = ctx->act.p;
  for (int d = 0; d < 3; ++d) {
    A.mass[d] = ctx->mass[d].p;
    A.vel[d] = ctx->mom[d].p;
    A.vol[d] = ctx->vol[d].p;
    A.w[d] = ctx->w[d].p;
    A.psi[d] = M.psi[d].p;
    A.sat[d] = M.sat[d].p;
    A.rhs[d] = ctx->rhs[d].p;
    A.invC[d] = ctx->invC[d].p;
    A.ustar[d] = ctx->ustar[d].p;
    A.Q[d] = M.Q[d].p;
    A.hdv[d] = M.hdv[d].p;
    A.gravity[d] = ctx->desc.gravity[d];
  }
  A.acc = M.acc.p;
  A.pp = M.pp.p;
  A.dt = dt;
  A.rho = ctx->desc.liquid_density;
  A.prm = M.prm;
  hipLaunchKernelGGL(k_step_velocity_mix, dim3(blocks(ctx->g.nnode, 256)), dim3(256), 0,
                     ctx->stream, A);
  WC_LAUNCH_CHECK();
}

}  // namespace wc
