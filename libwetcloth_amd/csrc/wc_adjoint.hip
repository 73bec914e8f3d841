// The adjoint of the liquid substep on gfx950.  The reference has no
// differentiation; this is the reverse-mode derivative of the substep the
// other modules run (P2G, grid update, assembly, pressure solve, pressure
// application, extrapolation, boundary constraint, G2P, advection and solid
// projection), taken with the substep's discrete decisions held fixed: the
// activated buckets, each cell's liquid-phi particle, the liquid cells and
// with them the pressure DOFs, the theta clamps, the faces with mass and the
// particles the solid projection moves.  A forward substep run with its
// decisions recorded (FrozenView mode 1) or replayed (mode 2) leaves
// everything the adjoint reads on the device.
//
// Every kernel is a gather, so no result depends on the order of atomics,
// except the G2P adjoint's spread of each particle's gradient onto its
// 3 x 27 face nodes (FP64 atomics).  The adjoint pressure solve reuses the
// MIC(0)-PCG of the forward solve on the same (symmetric) matrix.
#include <math.h>

#include <algorithm>

#include "wc_ctx.h"

namespace wc {
namespace {

inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }
constexpr int kRedBlocks = 256;

// derivative of quad_kernel (MathUtilities.h:416-424)
__device__ inline double quad_kernel_d(double x) {
  const double ax = fabs(x);
  if (ax < 0.5) return -2.0 * x;
  if (ax < 1.5) return x > 0 ? -(1.5 - ax) : (1.5 - ax);
  return 0.0;
}

// the three weights of a stencil on one axis and their derivatives by x
__device__ inline void weights_d(const Grid& g, double x, int base, double off, double w[3],
                                 double dw[3]) {
  for (int a = 0; a < 3; ++a) {
    const double r = (x - node_pos(g, base + a, off)) / g.dx;
    w[a] = quad_kernel(r);
    dw[a] = quad_kernel_d(r) / g.dx;
  }
}

// ---- decisions: slot map, liquid-phi particles, replay ----

__global__ void k_slot_of(const uint32_t* pid, int64_t n, uint32_t* slot_of) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p < n) slot_of[pid[p]] = (uint32_t)p;
}

// Among the particles whose distance to an active cell equals the cell's
// minimum (the key P2G left), the smallest caller index: the particle
// updateLiquidPhi's loop (TwoDScene.cpp:3307-3332) keeps, which takes a
// distance only when it is strictly smaller.
__global__ void k_lphi_argmin(Grid g, const double* x, int64_t n, int64_t cap, double r0,
                              const uint8_t* act, const unsigned long long* phikey,
                              const uint32_t* pid, uint32_t* argmin) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double px[3] = {x[p], x[cap + p], x[2 * cap + p]};
  int base[3];
  for (int a = 0; a < 3; ++a) base[a] = stencil_base_r<double>(g, a, px[a], 0.5);
  for (int c2 = 0; c2 < 3; ++c2)
    for (int c1 = 0; c1 < 3; ++c1)
      for (int c0 = 0; c0 < 3; ++c0) {
        const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
        if (!node_active(g, act, i, j, k)) continue;
        const double d0 = px[0] - node_pos_r<double>(g, i, 0.5);
        const double d1 = px[1] - node_pos_r<double>(g, j, 0.5);
        const double d2 = px[2] - node_pos_r<double>(g, k, 0.5);
        const double phi = sqrt(d0 * d0 + d1 * d1 + d2 * d2) - r0;
        const int64_t nn = node_index(g, i, j, k);
        if (phi_key(phi) == phikey[nn]) atomicMin(&argmin[nn], pid[p]);
      }
}

// -2 for a cell the solid override sets (TwoDScene.cpp:3334-3348), -1 for
// one that keeps the default (no particle distance below three buckets, or
// an inactive cell), else the particle of k_lphi_argmin
__global__ void k_lphi_arg(Grid g, const uint8_t* act, const unsigned long long* phikey,
                           const uint32_t* argmin, int32_t* arg) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= g.nnode) return;
  int32_t a = -1;
  if (act[n >> 6]) {
    int i, j, k;
    node_coords(g, n, i, j, k);
    if (container_phi(g, node_pos(g, i, 0.5), node_pos(g, j, 0.5), node_pos(g, k, 0.5)) < 0.0) {
      a = -2;
    } else {
      const unsigned long long key = phikey[n];
      if (key != 0xFFFFFFFFFFFFFFFFull && phi_from_key(key) < 3.0 * kBucket * g.dx &&
          argmin[n] != 0xFFFFFFFFu)
        a = (int32_t)argmin[n];
    }
  }
  arg[n] = a;
}

// the liquid phi of every cell from its recorded particle
__global__ void k_lphi_replay(Grid g, const int32_t* arg, const uint32_t* slot_of,
                              const double* x, int64_t cap, double r0, double* lphi) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= g.nnode) return;
  const int32_t a = arg[n];
  double phi = 3.0 * kBucket * g.dx;
  if (a == -2) {
    phi = -0.5 * g.dx;
  } else if (a >= 0) {
    int i, j, k;
    node_coords(g, n, i, j, k);
    const uint32_t s = slot_of[a];
    const double d0 = x[s] - node_pos(g, i, 0.5);
    const double d1 = x[cap + s] - node_pos(g, j, 0.5);
    const double d2 = x[2 * cap + s] - node_pos(g, k, 0.5);
    phi = sqrt(d0 * d0 + d1 * d1 + d2 * d2) - r0;
  }
  lphi[n] = phi;
}

// ---- particle state in the caller's order ----

// dst[t * n + pid[p]] = src[t * cap + p] for planes t < nplanes
__global__ void k_to_caller(const double* src, int64_t cap, const uint32_t* pid, int64_t n,
                            int nplanes, double* dst) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t q = pid[p];
  for (int t = 0; t < nplanes; ++t) dst[t * n + q] = src[t * cap + p];
}

__global__ void k_from_caller(const double* src, int64_t n, int64_t cap, int nplanes,
                              double* dst) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int t = 0; t < nplanes; ++t) dst[t * cap + p] = src[t * n + p];
}

__global__ void k_iota(uint32_t* pid, int64_t n) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p < n) pid[p] = (uint32_t)p;
}

// ---- the adjoint kernels, in the reverse order of the substep ----

struct AdjArgs {
  Grid g;
  int64_t n, cap;
  const uint32_t* pid;
  const double* x0;  // the substep's start state (slot order)
  const double* v0;
  const double* C0;
  const double* v;   // the G2P velocity (slot order)
  const uint8_t* act;
  const double* unew[3];
  const double* w[3];
  const double* mass[3];
  const double* vel[3];  // P2G velocity
  const double* vol[3];
  const double* invC[3];
  const double* lphi;
  const double* pressure;
  const uint32_t* isdof;
  const uint32_t* dofscan;
  const double* p;       // pressure per DOF
  const uint8_t* liquid; // the decisions of the substep
  const uint8_t* flags;
  const uint8_t* proj;
  const int32_t* arg;
  double* gx;            // gradient state, caller order
  double* gv;
  double* gC;
  double* ax;            // dL/dx0 of the G2P part, slot order
  double* ub[3];
  double* ub2[3];
  double* ub0[3];
  uint8_t* code[3];
  uint8_t* cnt[3];
  double* phib;
  const double* bb;
  double* pb;
  double* partial;
  double dt, flip, invD, friction, fm;
  int solid_projection;
};

__device__ inline bool liquid_at(const AdjArgs& A, int64_t n) { return A.liquid[n] != 0; }

// 1. G2P (TwoDScene.cpp:6958-7019), advection (SceneStepper.cpp:30-31) and
// solidProjection (TwoDScene.cpp:1782-1821) of one particle, reversed: the
// gradients of its new x, v, C go to its start position (ax) and, spread by
// the transfer weights, to the face velocities (ub).
__global__ __launch_bounds__(256) void k_adj_g2p(AdjArgs A) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  const Grid& g = A.g;
  const int64_t N = A.n, cap = A.cap;
  const uint32_t id = A.pid[p];
  const double px[3] = {A.x0[p], A.x0[cap + p], A.x0[2 * cap + p]};
  const double fv[3] = {A.v[p], A.v[cap + p], A.v[2 * cap + p]};
  double nx[3];
  for (int d = 0; d < 3; ++d) nx[d] = px[d] + fv[d] * A.dt;
  const double xo[3] = {A.gx[id], A.gx[N + id], A.gx[2 * N + id]};
  double nxb[3] = {xo[0], xo[1], xo[2]};
  double fvb[3] = {A.gv[id], A.gv[N + id], A.gv[2 * N + id]};
  double pxb[3] = {0, 0, 0};
  const double dx = g.dx, invD = A.invD;
  if (A.solid_projection && A.proj[id]) {
    // x' = nx - phi_now gh, gh = gr / |gr|, phi_now = phi_ori + gh . dp
    int base[3];
    double w[3][3], dw[3][3];
    for (int a = 0; a < 3; ++a) {
      base[a] = stencil_base(g, a, px[a], 0.0);
      weights_d(g, px[a], base[a], 0.0, w[a], dw[a]);
    }
    double phi_ori = 0.0, gr[3] = {0, 0, 0};
    for (int c2 = 0; c2 < 3; ++c2)
      for (int c1 = 0; c1 < 3; ++c1)
        for (int c0 = 0; c0 < 3; ++c0) {
          const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
          const double q0 = node_pos(g, i, 0.0), q1 = node_pos(g, j, 0.0), q2 = node_pos(g, k, 0.0);
          const double phi = node_active(g, A.act, i, j, k) ? container_phi(g, q0, q1, q2) : 3.0 * dx;
          const double ww = w[0][c0] * w[1][c1] * w[2][c2];
          phi_ori += phi * ww;
          gr[0] += phi * invD * ww * (q0 - nx[0]);
          gr[1] += phi * invD * ww * (q1 - nx[1]);
          gr[2] += phi * invD * ww * (q2 - nx[2]);
        }
    const double gn = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
    double gh[3] = {gr[0], gr[1], gr[2]};
    if (gn > 1e-20)
      for (int a = 0; a < 3; ++a) gh[a] /= gn;
    const double dp[3] = {fv[0] * A.dt, fv[1] * A.dt, fv[2] * A.dt};
    const double phi_now = phi_ori + (gh[0] * dp[0] + gh[1] * dp[1] + gh[2] * dp[2]);
    const double pnb = -(gh[0] * xo[0] + gh[1] * xo[1] + gh[2] * xo[2]);
    double ghb[3], grb[3];
    for (int a = 0; a < 3; ++a) {
      ghb[a] = -phi_now * xo[a] + pnb * dp[a];
      fvb[a] += A.dt * pnb * gh[a];
    }
    if (gn > 1e-20) {
      const double dd = gh[0] * ghb[0] + gh[1] * ghb[1] + gh[2] * ghb[2];
      for (int a = 0; a < 3; ++a) grb[a] = (ghb[a] - gh[a] * dd) / gn;
    } else {
      for (int a = 0; a < 3; ++a) grb[a] = ghb[a];
    }
    for (int a = 0; a < 3; ++a) nxb[a] += -invD * phi_ori * grb[a];
    for (int c2 = 0; c2 < 3; ++c2)
      for (int c1 = 0; c1 < 3; ++c1)
        for (int c0 = 0; c0 < 3; ++c0) {
          const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
          const double q[3] = {node_pos(g, i, 0.0), node_pos(g, j, 0.0), node_pos(g, k, 0.0)};
          const double phi =
              node_active(g, A.act, i, j, k) ? container_phi(g, q[0], q[1], q[2]) : 3.0 * dx;
          const double coef = pnb + invD * ((q[0] - nx[0]) * grb[0] + (q[1] - nx[1]) * grb[1] +
                                            (q[2] - nx[2]) * grb[2]);
          pxb[0] += phi * coef * dw[0][c0] * w[1][c1] * w[2][c2];
          pxb[1] += phi * coef * w[0][c0] * dw[1][c1] * w[2][c2];
          pxb[2] += phi * coef * w[0][c0] * w[1][c1] * dw[2][c2];
        }
  }
  // nx = px + fv dt
  for (int a = 0; a < 3; ++a) {
    pxb[a] += nxb[a];
    fvb[a] += A.dt * nxb[a];
  }
  // fv_d = sum w u;  C_d = flip sum w u (x_n - x_p) invD
  for (int d = 0; d < 3; ++d) {
    int base[3];
    double w[3][3], dw[3][3];
    for (int a = 0; a < 3; ++a) {
      base[a] = stencil_base(g, a, px[a], face_off(d, a));
      weights_d(g, px[a], base[a], face_off(d, a), w[a], dw[a]);
    }
    const double Bb[3] = {A.flip * A.gC[(d * 3 + 0) * N + id], A.flip * A.gC[(d * 3 + 1) * N + id],
                          A.flip * A.gC[(d * 3 + 2) * N + id]};
    for (int c2 = 0; c2 < 3; ++c2)
      for (int c1 = 0; c1 < 3; ++c1)
        for (int c0 = 0; c0 < 3; ++c0) {
          const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
          if (!node_active(g, A.act, i, j, k)) continue;
          const int64_t n = node_index(g, i, j, k);
          const double u = A.unew[d][n];
          const double e[3] = {node_pos(g, i, face_off(d, 0)) - px[0],
                               node_pos(g, j, face_off(d, 1)) - px[1],
                               node_pos(g, k, face_off(d, 2)) - px[2]};
          const double G = fvb[d] + invD * (Bb[0] * e[0] + Bb[1] * e[1] + Bb[2] * e[2]);
          const double ww = w[0][c0] * w[1][c1] * w[2][c2];
          unsafeAtomicAdd(&A.ub[d][n], ww * G);
          pxb[0] += u * G * dw[0][c0] * w[1][c1] * w[2][c2] - ww * u * invD * Bb[0];
          pxb[1] += u * G * w[0][c0] * dw[1][c1] * w[2][c2] - ww * u * invD * Bb[1];
          pxb[2] += u * G * w[0][c0] * w[1][c1] * dw[2][c2] - ww * u * invD * Bb[2];
        }
  }
  for (int a = 0; a < 3; ++a) A.ax[a * cap + p] = pxb[a];
}

// whether the pressure application (Pressure.cpp:1184-1211) sets face (d, n)
// itself: positive solid weight and a liquid cell on either side
__device__ inline bool direct_face(const AdjArgs& A, int d, int64_t n, int i, int j, int k) {
  if (!A.act[n >> 6] || !(A.w[d][n] > 0.)) return false;
  int l[3] = {i, j, k};
  l[d] -= 1;
  const bool lex = node_active(A.g, A.act, l[0], l[1], l[2]);
  return (lex && liquid_at(A, node_index(A.g, l[0], l[1], l[2]))) || liquid_at(A, n);
}

// 2. constrainLiquidVelocity (TwoDScene.cpp:1827-1866) reversed, and the
// faces the extrapolation starts from (bit 0 of code)
__global__ void k_adj_direct(AdjArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode) return;
  int i, j, k;
  node_coords(g, n, i, j, k);
  for (int d = 0; d < 3; ++d) {
    A.code[d][n] = direct_face(A, d, n, i, j, k) ? 1 : 0;
    A.cnt[d][n] = 0;
    if (A.act[n >> 6] && A.w[d][n] == 0.0) A.ub[d][n] *= (1 - A.friction);
  }
}

// the faces layer L of the extrapolation (Pressure.cpp:1826-1903) fills
// (code bit 1 + L) and how many valid neighbours each averages (cnt, four
// bits per layer); valid before layer 0 = bit 0, before layer 1 = bits 0-1
__global__ void k_adj_fill(AdjArgs A, int L) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode || !A.act[n >> 6]) return;
  int i, j, k;
  node_coords(g, n, i, j, k);
  const uint8_t vmask = L == 0 ? 1 : 3;
  for (int d = 0; d < 3; ++d) {
    if (A.code[d][n] & vmask) continue;
    int c = 0;
    for (int r = 0; r < 6; ++r) {
      int q[3] = {i, j, k};
      q[r >> 1] += (r & 1) ? 1 : -1;
      if (!node_active(g, A.act, q[0], q[1], q[2])) continue;
      if (A.code[d][node_index(g, q[0], q[1], q[2])] & vmask) ++c;
    }
    if (c > 0) {
      A.code[d][n] |= (uint8_t)(2 << L);
      A.cnt[d][n] |= (uint8_t)(c << (4 * L));
    }
  }
}

// layer L reversed: a filled face hands its gradient to the neighbours it
// averaged, in equal shares, and keeps none (its value before the layer was
// overwritten).  Gather form: ub2 = ub with the shares added.
__global__ void k_adj_extrap(AdjArgs A, int L, int flip) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode) return;
  double* const* in = flip ? A.ub2 : A.ub;
  double* const* out = flip ? A.ub : A.ub2;
  if (!A.act[n >> 6]) {
    for (int d = 0; d < 3; ++d) out[d][n] = 0.0;
    return;
  }
  int i, j, k;
  node_coords(g, n, i, j, k);
  const uint8_t vmask = L == 0 ? 1 : 3, fbit = (uint8_t)(2 << L);
  for (int d = 0; d < 3; ++d) {
    const uint8_t c = A.code[d][n];
    double v = (c & fbit) ? 0.0 : in[d][n];
    if (c & vmask) {
      for (int r = 0; r < 6; ++r) {
        int q[3] = {i, j, k};
        q[r >> 1] += (r & 1) ? 1 : -1;
        if (!node_active(g, A.act, q[0], q[1], q[2])) continue;
        const int64_t m = node_index(g, q[0], q[1], q[2]);
        if (A.code[d][m] & fbit) v += in[d][m] / (double)((A.cnt[d][m] >> (4 * L)) & 15);
      }
    }
    out[d][n] = v;
  }
}

// theta of face (d, fn) whose left cell is ln (< 0: none) as the forward
// substep took it (the liquid flags and clamps of the substep)
__device__ inline double face_theta(const AdjArgs& A, int d, int64_t fn, int64_t ln, bool ll,
                                    bool lr) {
  if (ll && lr) return 1.0;
  if ((A.flags[fn] >> (3 + d)) & 1) return kThetaCriterion;
  const double phil = ln >= 0 ? A.lphi[ln] : 0.0, phir = A.lphi[fn];
  if (ll) return phil / (phil - phir);
  if (lr) return phir / (phir - phil);
  return 0.0;
}

// 4. the pressure application (Pressure.cpp:1137-1314, performInvLocalSolve
// LinearizedImplicitEuler.cpp:65-104), u = invC rhs - invC dt V (p_R - p_L) /
// (dx theta), reversed into the gradient of the pressures: per DOF cell,
// gathered from its six faces
__global__ void k_adj_pb(AdjArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  if (n >= g.nnode || !A.isdof[n]) return;
  int i, j, k;
  node_coords(g, n, i, j, k);
  const double dx = g.dx;
  double acc = 0.0;
  for (int d = 0; d < 3; ++d)
    for (int hi = 0; hi < 2; ++hi) {
      int q[3] = {i, j, k};
      q[d] += hi;
      if (!node_active(g, A.act, q[0], q[1], q[2])) continue;
      const int64_t fn = node_index(g, q[0], q[1], q[2]);
      if (!(A.code[d][fn] & 1)) continue;  // not set by the pressure application
      int l[3] = {q[0], q[1], q[2]};
      l[d] -= 1;
      const bool lex = node_active(g, A.act, l[0], l[1], l[2]);
      const int64_t ln = lex ? node_index(g, l[0], l[1], l[2]) : -1;
      const bool ll = lex && liquid_at(A, ln), lr = liquid_at(A, fn);
      const double theta = face_theta(A, d, fn, ln, ll, lr);
      const double c = A.invC[d][fn] * A.dt * A.vol[d][fn] / (dx * theta);
      // u -= c (p_fn - p_ln): this cell is fn's right cell (hi = 0) or left
      acc += hi ? c * A.ub[d][fn] : -(c * A.ub[d][fn]);
    }
  A.pb[A.dofscan[n]] = acc;
}

// 5-6. After the adjoint solve bb = A^-1 pb: the gradients of the grid
// update's rhs (the pressure application's invC rhs and the assembled rhs
// b: +-(u* psi_fs) w / dx with u* psi_fs = invC rhs, Pressure.cpp:444-791),
// of the P2G velocities (rhs = M (u + g dt), LinearizedImplicitEuler.cpp:
// 224-337: ub0 = M rbar where the face has mass) and the block partials of
// dL/dg_y = sum M dt rbar; and per cell the gradient of its liquid phi
// through the theta of its faces (the pressure application's and the
// diagonal's term / theta of each air face of a DOF).
__device__ inline double rbar_at(const AdjArgs& A, int d, int64_t n, int i, int j, int k) {
  const Grid& g = A.g;
  const double invC = A.invC[d][n];
  double rb = invC * A.ub[d][n];
  const double wv = A.w[d][n];
  if (A.isdof[n]) rb += A.bb[A.dofscan[n]] * invC * wv / g.dx;  // the cell's low face
  int l[3] = {i, j, k};
  l[d] -= 1;
  if (node_active(g, A.act, l[0], l[1], l[2])) {
    const int64_t ln = node_index(g, l[0], l[1], l[2]);
    if (A.isdof[ln]) rb += -A.bb[A.dofscan[ln]] * invC * wv / g.dx;  // the left cell's high face
  }
  return rb;
}

// the theta gradient of face (d, fn) into the cell on side `right` of it
__device__ inline double face_phib(const AdjArgs& A, int d, int64_t fn, int fi, int fj, int fk,
                                   bool right) {
  const Grid& g = A.g;
  int l[3] = {fi, fj, fk};
  l[d] -= 1;
  const bool lex = node_active(g, A.act, l[0], l[1], l[2]);
  const int64_t ln = lex ? node_index(g, l[0], l[1], l[2]) : -1;
  const bool ll = lex && liquid_at(A, ln), lr = liquid_at(A, fn);
  if (ll == lr) return 0.0;                   // theta = 1 (or no liquid): constant
  if ((A.flags[fn] >> (3 + d)) & 1) return 0.0;  // clamped
  const double wv = A.w[d][fn];
  if (!(wv > 0.0)) return 0.0;
  const double dx = g.dx, dt = A.dt;
  const double phil = lex ? A.lphi[ln] : 0.0, phir = A.lphi[fn];
  const double theta = ll ? phil / (phil - phir) : phir / (phir - phil);
  const double invC = A.invC[d][fn], vol = A.vol[d][fn];
  double tb = 0.0;
  // the pressure application's term (one side liquid and w > 0: a face it sets)
  {
    const double pn = A.pressure[fn], pl = lex ? A.pressure[ln] : 0.0;
    tb += A.ub[d][fn] * invC * dt * vol * (pn - pl) / (dx * theta * theta);
  }
  // the diagonal's term / theta of the liquid DOF cell next to an air cell
  const double term = wv * (dt / (dx * dx)) * (invC * vol);
  if (lr && lex && A.isdof[fn]) {
    const uint32_t me = A.dofscan[fn];
    tb += A.bb[me] * A.p[me] * term / (theta * theta);
  }
  if (ll && A.isdof[ln]) {
    const uint32_t me = A.dofscan[ln];
    tb += A.bb[me] * A.p[me] * term / (theta * theta);
  }
  // theta = a / (a - b), a the liquid cell's phi, b the air cell's
  const double a = ll ? phil : phir, b = ll ? phir : phil;
  const double den = (a - b) * (a - b);
  const double da = -b / den, db = a / den;
  const bool liquid_side = ll ? !right : right;
  return tb * (liquid_side ? da : db);
}

__global__ __launch_bounds__(256) void k_adj_post(AdjArgs A) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const Grid& g = A.g;
  double gy = 0.0;
  if (n < g.nnode) {
    if (!A.act[n >> 6]) {
      for (int d = 0; d < 3; ++d) A.ub0[d][n] = 0.0;
      A.phib[n] = 0.0;
    } else {
      int i, j, k;
      node_coords(g, n, i, j, k);
      for (int d = 0; d < 3; ++d) {
        const bool hm = ((A.flags[n] >> d) & 1) != 0;
        const double rb = rbar_at(A, d, n, i, j, k);
        const double M = A.mass[d][n];
        A.ub0[d][n] = hm ? M * rb : 0.0;
        if (d == 1 && hm) gy += M * A.dt * rb;
      }
      double pb = 0.0;
      for (int d = 0; d < 3; ++d) {
        pb += face_phib(A, d, n, i, j, k, true);  // the cell's low face
        int q[3] = {i, j, k};
        q[d] += 1;
        if (node_active(g, A.act, q[0], q[1], q[2]))
          pb += face_phib(A, d, node_index(g, q[0], q[1], q[2]), q[0], q[1], q[2], false);
      }
      A.phib[n] = pb;
    }
  }
  // block partial of dL/dg_y in a fixed order
  for (int o = 32; o > 0; o >>= 1) gy += __shfl_xor(gy, o, 64);
  __shared__ double sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = gy;
  __syncthreads();
  if (threadIdx.x == 0) A.partial[blockIdx.x] = ((sm[0] + sm[1]) + sm[2]) + sm[3];
}

__global__ void k_sum_partials(const double* partial, int64_t nblocks, double* out) {
  __shared__ double sm[256];
  double s = 0.0;
  for (int64_t b = threadIdx.x; b < nblocks; b += 256) s += partial[b];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

// 7-8. P2G (TwoDScene.cpp:6700-6713; u = P / M, P = fm sum w (v + C e),
// M = fm sum w) and the liquid phi (TwoDScene.cpp:3307-3332; phi = |x_p -
// x_c| - r0 of the cell's recorded particle) reversed, per particle: the
// gradients of its start state, written in the caller's order
__global__ __launch_bounds__(256) void k_adj_p2g(AdjArgs A) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  const Grid& g = A.g;
  const int64_t N = A.n, cap = A.cap;
  const uint32_t id = A.pid[p];
  const double px[3] = {A.x0[p], A.x0[cap + p], A.x0[2 * cap + p]};
  double pxb[3] = {A.ax[p], A.ax[cap + p], A.ax[2 * cap + p]};
  const double fm = A.fm;
  for (int d = 0; d < 3; ++d) {
    int base[3];
    double w[3][3], dw[3][3];
    for (int a = 0; a < 3; ++a) {
      base[a] = stencil_base(g, a, px[a], face_off(d, a));
      weights_d(g, px[a], base[a], face_off(d, a), w[a], dw[a]);
    }
    const double vd = A.v0[d * cap + p];
    const double b0 = A.C0[(d * 3 + 0) * cap + p];
    const double b1 = A.C0[(d * 3 + 1) * cap + p];
    const double b2 = A.C0[(d * 3 + 2) * cap + p];
    double vdb = 0, cb[3] = {0, 0, 0};
    for (int c2 = 0; c2 < 3; ++c2)
      for (int c1 = 0; c1 < 3; ++c1)
        for (int c0 = 0; c0 < 3; ++c0) {
          const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
          if (!node_active(g, A.act, i, j, k)) continue;
          const int64_t n = node_index(g, i, j, k);
          if (!((A.flags[n] >> d) & 1)) continue;
          const double M = A.mass[d][n];
          const double Pb = A.ub0[d][n] / M;
          const double Mb = -A.ub0[d][n] * A.vel[d][n] / M;
          const double e[3] = {node_pos(g, i, face_off(d, 0)) - px[0],
                               node_pos(g, j, face_off(d, 1)) - px[1],
                               node_pos(g, k, face_off(d, 2)) - px[2]};
          const double vel = vd + (b0 * e[0] + b1 * e[1] + b2 * e[2]);
          const double ww = w[0][c0] * w[1][c1] * w[2][c2];
          vdb += fm * ww * Pb;
          for (int a = 0; a < 3; ++a) cb[a] += fm * ww * Pb * e[a];
          const double G = fm * (Pb * vel + Mb);
          pxb[0] += G * dw[0][c0] * w[1][c1] * w[2][c2] - fm * ww * Pb * b0;
          pxb[1] += G * w[0][c0] * dw[1][c1] * w[2][c2] - fm * ww * Pb * b1;
          pxb[2] += G * w[0][c0] * w[1][c1] * dw[2][c2] - fm * ww * Pb * b2;
        }
    A.gv[d * N + id] = vdb;
    for (int a = 0; a < 3; ++a) A.gC[(d * 3 + a) * N + id] = cb[a];
  }
  {
    int base[3];
    for (int a = 0; a < 3; ++a) base[a] = stencil_base(g, a, px[a], 0.5);
    for (int c2 = 0; c2 < 3; ++c2)
      for (int c1 = 0; c1 < 3; ++c1)
        for (int c0 = 0; c0 < 3; ++c0) {
          const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
          if (!in_range(g, i, j, k)) continue;
          const int64_t n = node_index(g, i, j, k);
          if (A.arg[n] != (int32_t)id) continue;
          const double pb = A.phib[n];
          if (pb == 0.0) continue;
          const double d0 = px[0] - node_pos(g, i, 0.5);
          const double d1 = px[1] - node_pos(g, j, 0.5);
          const double d2 = px[2] - node_pos(g, k, 0.5);
          const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
          if (r > 0) {
            pxb[0] += pb * d0 / r;
            pxb[1] += pb * d1 / r;
            pxb[2] += pb * d2 / r;
          }
        }
  }
  for (int a = 0; a < 3; ++a) A.gx[a * N + id] = pxb[a];
}

DevCsr context_system(const wc_ctx* c) {
  DevCsr A;
  A.n = c->ndof;
  A.stride = 7;
  A.len = c->rowlen.p;
  A.col = c->col.p;
  A.val = c->val.p;
  A.nnz_slots = 7 * c->ndof;
  return A;
}

double liquid_r0(const wc_ctx* c) {
  return c->g.dx * 0.883644 > c->desc.particle_radius ? c->g.dx * 0.883644
                                                        : c->desc.particle_radius;
}

}  // namespace

void FrozenRecord::release() {
  act.release();
  liquid.release();
  flags.release();
  proj.release();
  arg.release();
  recorded = false;
  np = 0;
}

void ParticleCheckpoint::release() {
  x.release();
  v.release();
  C.release();
  n = 0;
}

void AdjointWork::release() {
  slot_of.release();
  argmin.release();
  x0.release();
  v0.release();
  C0.release();
  gx.release();
  gv.release();
  gC.release();
  ax.release();
  for (int d = 0; d < 3; ++d) {
    ub[d].release();
    ub2[d].release();
    ub0[d].release();
    code[d].release();
    cnt[d].release();
  }
  phib.release();
  pb.release();
  bb.release();
  partial.release();
  n = 0;
  have_start = false;
  ready = false;
}

void adjoint_release(wc_ctx* c) {
  for (FrozenRecord& r : c->frz) r.release();
  c->frz.clear();
  for (ParticleCheckpoint& k : c->ckpt) k.release();
  c->ckpt.clear();
  c->adj.release();
  c->frz_dep.release();
}

FrozenRecord& frozen_record(wc_ctx* c) {
  if ((int)c->frz.size() <= c->frz_slot) c->frz.resize((size_t)c->frz_slot + 1);
  FrozenRecord& r = c->frz[(size_t)c->frz_slot];
  const Grid& g = c->g;
  r.act.ensure(g.nbucket);
  r.liquid.ensure(g.nnode);
  r.flags.ensure(g.nnode);
  r.arg.ensure(g.nnode);
  r.proj.ensure(std::max<int64_t>(1, c->np));
  return r;
}

FrozenView frozen_view(wc_ctx* c) {
  FrozenView F;
  if (!c->frz_mode) return F;
  FrozenRecord& r = frozen_record(c);
  F.mode = c->frz_mode;
  F.liquid = r.liquid.p;
  F.flags = r.flags.p;
  F.proj = r.proj.p;
  F.pid = c->pid.p;
  if (c->frz_mode == 2) {
    if (!c->frz_dep.p) {
      c->frz_dep.ensure(1);
      WC_HIP(hipMemsetAsync(c->frz_dep.p, 0, sizeof(unsigned long long), c->stream));
    }
    F.dep = c->frz_dep.p;
  }
  return F;
}

void frozen_after_sort(wc_ctx* c) {
  AdjointWork& a = c->adj;
  const int64_t n = c->np, cap = c->pcap;
  hipStream_t st = c->stream;
  a.slot_of.ensure(std::max<int64_t>(1, n));
  if (n > 0) {
    hipLaunchKernelGGL(k_slot_of, dim3(blocks(n, 256)), dim3(256), 0, st, c->pid.p, n,
                       a.slot_of.p);
    WC_LAUNCH_CHECK();
  }
  a.x0.ensure(3 * (size_t)cap);
  a.v0.ensure(3 * (size_t)cap);
  a.C0.ensure(9 * (size_t)cap);
  WC_HIP(hipMemcpyAsync(a.x0.p, c->x.p, sizeof(double) * 3 * cap, hipMemcpyDeviceToDevice, st));
  WC_HIP(hipMemcpyAsync(a.v0.p, c->v.p, sizeof(double) * 3 * cap, hipMemcpyDeviceToDevice, st));
  WC_HIP(hipMemcpyAsync(a.C0.p, c->C.p, sizeof(double) * 9 * cap, hipMemcpyDeviceToDevice, st));
  a.have_start = true;
  a.ready = false;
}

void frozen_replay_lphi(wc_ctx* c) {
  const Grid& g = c->g;
  FrozenRecord& r = frozen_record(c);
  hipLaunchKernelGGL(k_lphi_replay, dim3(blocks(g.nnode, 256)), dim3(256), 0, c->stream, g,
                     r.arg.p, c->adj.slot_of.p, c->x.p, c->pcap, liquid_r0(c), c->lphi.p);
  WC_LAUNCH_CHECK();
}

void frozen_record_lphi_arg(wc_ctx* c) {
  const Grid& g = c->g;
  hipStream_t st = c->stream;
  FrozenRecord& r = frozen_record(c);
  AdjointWork& a = c->adj;
  a.argmin.ensure(g.nnode);
  WC_HIP(hipMemsetAsync(a.argmin.p, 0xFF, sizeof(uint32_t) * g.nnode, st));
  if (c->np > 0) {
    hipLaunchKernelGGL(k_lphi_argmin, dim3(blocks(c->np, 256)), dim3(256), 0, st, g, c->x.p,
                       c->np, c->pcap, liquid_r0(c), c->act.p, c->phikey.p, c->pid.p,
                       a.argmin.p);
    WC_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_lphi_arg, dim3(blocks(g.nnode, 256)), dim3(256), 0, st, g, c->act.p,
                     c->phikey.p, a.argmin.p, r.arg.p);
  WC_LAUNCH_CHECK();
}

bool adjoint_substep(wc_ctx* c, double dt, double tol, int maxit, double* gb,
                     double* residual_out, int* iters_out) {
  const Grid& g = c->g;
  hipStream_t st = c->stream;
  AdjointWork& a = c->adj;
  FrozenRecord& rec = frozen_record(c);
  const int64_t nn = g.nnode, n = c->np;
  for (int d = 0; d < 3; ++d) {
    a.ub[d].ensure(nn);
    a.ub2[d].ensure(nn);
    a.ub0[d].ensure(nn);
    a.code[d].ensure(nn);
    a.cnt[d].ensure(nn);
    WC_HIP(hipMemsetAsync(a.ub[d].p, 0, sizeof(double) * nn, st));
  }
  a.phib.ensure(nn);
  a.ax.ensure(3 * (size_t)c->pcap);
  a.pb.ensure(std::max<int64_t>(1, c->ndof));
  a.bb.ensure(std::max<int64_t>(1, c->ndof));
  const unsigned nblk = blocks(nn, 256);
  a.partial.ensure(nblk + 1);

  AdjArgs A;
  A.g = g;
  A.n = n;
  A.cap = c->pcap;
  A.pid = c->pid.p;
  A.x0 = a.x0.p;
  A.v0 = a.v0.p;
  A.C0 = a.C0.p;
  A.v = c->v.p;
  A.act = c->act.p;
  for (int d = 0; d < 3; ++d) {
    A.unew[d] = c->unew[d].p;
    A.w[d] = c->w[d].p;
    A.mass[d] = c->mass[d].p;
    A.vel[d] = c->mom[d].p;
    A.vol[d] = c->vol[d].p;
    A.invC[d] = c->invC[d].p;
    A.ub[d] = a.ub[d].p;
    A.ub2[d] = a.ub2[d].p;
    A.ub0[d] = a.ub0[d].p;
    A.code[d] = a.code[d].p;
    A.cnt[d] = a.cnt[d].p;
  }
  A.lphi = c->lphi.p;
  A.pressure = c->pressure.p;
  A.isdof = c->isdof.p;
  A.dofscan = c->dofscan.p;
  A.p = c->p.p;
  A.liquid = rec.liquid.p;
  A.flags = rec.flags.p;
  A.proj = rec.proj.p;
  A.arg = rec.arg.p;
  A.gx = a.gx.p;
  A.gv = a.gv.p;
  A.gC = a.gC.p;
  A.ax = a.ax.p;
  A.phib = a.phib.p;
  A.bb = a.bb.p;
  A.pb = a.pb.p;
  A.partial = a.partial.p;
  A.dt = dt;
  A.flip = c->desc.flip_coeff;
  A.invD = 4.0 / (g.dx * g.dx);
  A.friction = c->desc.liquid_boundary_friction;
  A.fm = c->desc.particle_mass;
  A.solid_projection = c->desc.solid_projection;

  const unsigned pb_ = blocks(std::max<int64_t>(1, n), 256);
  if (n > 0) {
    hipLaunchKernelGGL(k_adj_g2p, dim3(pb_), dim3(256), 0, st, A);
    WC_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_adj_direct, dim3(nblk), dim3(256), 0, st, A);
  hipLaunchKernelGGL(k_adj_fill, dim3(nblk), dim3(256), 0, st, A, 0);
  hipLaunchKernelGGL(k_adj_fill, dim3(nblk), dim3(256), 0, st, A, 1);
  // layer 1, then layer 0: ub -> ub2 -> ub
  hipLaunchKernelGGL(k_adj_extrap, dim3(nblk), dim3(256), 0, st, A, 1, 0);
  hipLaunchKernelGGL(k_adj_extrap, dim3(nblk), dim3(256), 0, st, A, 0, 1);
  WC_LAUNCH_CHECK();
  bool ok = true;
  double res = 0;
  int it = 0;
  if (c->ndof > 0) {
    hipLaunchKernelGGL(k_adj_pb, dim3(nblk), dim3(256), 0, st, A);
    WC_LAUNCH_CHECK();
    // bbar = A^-1 pbar: the matrix is symmetric, so the forward's MIC(0)-PCG
    const DevCsr S = context_system(c);
    pcg_prepare(st, c->pcg, S, c->dof_level.p, c->level_max - c->level_min + 1);
    ok = pcg_solve(st, c->pcg, S, a.pb.p, a.bb.p, tol, maxit, 0.97, 0.25, &res, &it);
  }
  hipLaunchKernelGGL(k_adj_post, dim3(nblk), dim3(256), 0, st, A);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, a.partial.p, (int64_t)nblk,
                     a.partial.p + nblk);
  WC_LAUNCH_CHECK();
  if (n > 0) {
    hipLaunchKernelGGL(k_adj_p2g, dim3(pb_), dim3(256), 0, st, A);
    WC_LAUNCH_CHECK();
  }
  double gy = 0.0;
  WC_HIP(hipMemcpyAsync(&gy, a.partial.p + nblk, sizeof(double), hipMemcpyDeviceToHost, st));
  WC_HIP(hipStreamSynchronize(st));
  *gb += -gy;  // gravity (0, -|g|, 0)
  if (residual_out) *residual_out = res;
  if (iters_out) *iters_out = it;
  return ok;
}

void checkpoint_save(wc_ctx* c, int slot) {
  if ((int)c->ckpt.size() <= slot) c->ckpt.resize((size_t)slot + 1);
  ParticleCheckpoint& k = c->ckpt[(size_t)slot];
  const int64_t n = c->np;
  k.x.ensure(std::max<int64_t>(1, 3 * n));
  k.v.ensure(std::max<int64_t>(1, 3 * n));
  k.C.ensure(std::max<int64_t>(1, 9 * n));
  k.n = n;
  if (n == 0) return;
  hipStream_t st = c->stream;
  const unsigned b = blocks(n, 256);
  hipLaunchKernelGGL(k_to_caller, dim3(b), dim3(256), 0, st, c->x.p, c->pcap, c->pid.p, n, 3, k.x.p);
  hipLaunchKernelGGL(k_to_caller, dim3(b), dim3(256), 0, st, c->v.p, c->pcap, c->pid.p, n, 3, k.v.p);
  hipLaunchKernelGGL(k_to_caller, dim3(b), dim3(256), 0, st, c->C.p, c->pcap, c->pid.p, n, 9, k.C.p);
  WC_LAUNCH_CHECK();
}

void checkpoint_restore(wc_ctx* c, int slot) {
  const ParticleCheckpoint& k = c->ckpt[(size_t)slot];
  const int64_t n = k.n;
  if (n > c->pcap) throw HipError{hipErrorInvalidValue, "checkpoint larger than the particle capacity", __FILE__, __LINE__};
  c->np = n;
  c->sorted = false;
  c->xw_valid = false;
  c->adj.have_start = false;
  c->adj.ready = false;
  if (n == 0) return;
  hipStream_t st = c->stream;
  const unsigned b = blocks(n, 256);
  hipLaunchKernelGGL(k_from_caller, dim3(b), dim3(256), 0, st, k.x.p, n, c->pcap, 3, c->x.p);
  hipLaunchKernelGGL(k_from_caller, dim3(b), dim3(256), 0, st, k.v.p, n, c->pcap, 3, c->v.p);
  hipLaunchKernelGGL(k_from_caller, dim3(b), dim3(256), 0, st, k.C.p, n, c->pcap, 9, c->C.p);
  hipLaunchKernelGGL(k_iota, dim3(b), dim3(256), 0, st, c->pid.p, n);
  WC_LAUNCH_CHECK();
}

}  // namespace wc
