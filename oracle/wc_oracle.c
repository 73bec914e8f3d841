/*
 * FP64 CPU oracle for the libWetCloth liquid substep -- TEST INFRASTRUCTURE.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this; the product library never does.  See wc_oracle.h for the parity
 * status of each function.  Reference paths are relative to libWetCloth/Core.
 *
 * Build: gcc -O2 -ffp-contract=off (no fused multiply-add: the reference's
 * x86-64 SSE2 build has none, and the HIP kernels are built the same way).
 */
#include "wc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define THETA_CRITERION 0.01 /* Pressure.cpp:20 */

/* Planted mistakes, for the tests only: tests/test_oracle_pins.py compiles
 * this file again with -DWCO_PLANT=k to show that each pin of the oracle
 * fails on mistake k.  The default build (WCO_PLANT 0) contains none of them.
 *   1 activation without the one-bucket ring     2 activation with two rings
 *   3 stencil buckets from the base node only     4 no solid override of the liquid phi
 *   5 DOF test w >= 0 instead of w > 0            6 DOFs numbered i-fastest over the grid
 *   7 extrapolation in place (no copies)          8 one extrapolation layer
 *   9 flip_coeff applied to v instead of C        10 gravity with the wrong sign
 *   11 u* without the dt factor on gravity        12 liquid phi radius without the
 *                                                    0.883644 dx floor
 *   adjoint (wco_substep_adjoint) without:  13 the theta terms  14 the solid
 *   projection  15 the liquid phi's particle term  16 P2G's weight derivative
 *   mixture:  20 assembly without the (1 - psi) factor  21 pressure application
 *   without the solid volume V_s Q  22 pore pressure gradient with its sign
 *   flipped  23 Q formed from the liquid mass  24 the solid's P2G on the
 *   cell-centre stencil for every axis */
#ifndef WCO_PLANT
#define WCO_PLANT 0
#endif

/* the static porous solid (the mixture section at the end of this file) */
struct wco_mix;
static void mix_touch(const wco_scene* s, uint8_t* touched);
static void mix_p2g(wco_scene* s);
static void mix_step_velocity(wco_scene* s, double dt);
static void mix_free(struct wco_mix* m);
/* a face's solid terms: volume fraction psi, solid volume Vs, drag ratios Q
 * (mshdvm_hdvm) and P (mfhdvm_hdvm), psi_sf and solid mass */
static void mix_face(const wco_scene* s, int d, int64_t n, double* psi, double* Vs, double* Q,
                     double* P, double* psf, double* Ms);

/* ------------------------------------------------------------------ grid */

void wco_grid_init(wco_grid* g, int ni, int nj, int nk, double dx) {
  g->n[0] = ni;
  g->n[1] = nj;
  g->n[2] = nk;
  g->dx = dx;
  for (int d = 0; d < 3; ++d) {
    /* two buckets of ghost cells below, the interior, then two more buckets
     * above: the stencils of particles in the container reach one cell past
     * it and the activation dilates by one more bucket
     * (expandFluidNodesMarked(1), TwoDScene.cpp:5021-5064); the reference's
     * grid extends three buckets past the particles' bounding box
     * (TwoDScene.cpp:1657-1690), so it never clips that dilation, and
     * neither does this one */
    g->nb[d] = (g->n[d] + WCO_BUCKET - 1) / WCO_BUCKET + 4;
    g->nx[d] = g->nb[d] * WCO_BUCKET;
  }
  g->nbucket = (int64_t)g->nb[0] * g->nb[1] * g->nb[2];
  g->nnode = g->nbucket * 64;
}

int64_t wco_bucket_index(const wco_grid* g, int bi, int bj, int bk) {
  return ((int64_t)bk * g->nb[1] + bj) * g->nb[0] + bi;
}

int64_t wco_node_index(const wco_grid* g, int i, int j, int k) {
  return (wco_bucket_index(g, i >> 2, j >> 2, k >> 2) << 6) |
         ((int64_t)(k & 3) << 4) | ((j & 3) << 2) | (i & 3);
}

static int in_range(const wco_grid* g, int i, int j, int k) {
  return i >= 0 && j >= 0 && k >= 0 && i < g->nx[0] && j < g->nx[1] &&
         k < g->nx[2];
}

static int node_active(const wco_scene* s, int i, int j, int k) {
  if (!in_range(&s->g, i, j, k)) return 0;
  return s->act[wco_bucket_index(&s->g, i >> 2, j >> 2, k >> 2)] != 0;
}

/* world coordinate of dense node index i on an axis, shifted by off cells */
static double node_pos(const wco_grid* g, int i, double off) {
  return ((double)(i - WCO_GHOST) + off) * g->dx;
}

/* The decisions a replay holds fixed (wco_frozen_set): whether a cell is
 * liquid and whether a face has mass. */
static int is_liquid(const wco_scene* s, int64_t n) {
  return s->frz_mode == 2 ? s->frz_liquid[n] : s->lphi[n] < 0.0;
}
static int has_mass(const wco_scene* s, int d, int64_t n) {
  return s->frz_mode == 2 ? s->frz_mass[d][n] : s->mass[d][n] > 1e-20;
}

/* theta of the face (axis d, node fn) between cells nl and nr (nl < 0: no
 * cell) with liquid phis phil, phir: fraction_inside(phil, phir)
 * (MathUtilities.h:286-293) by the cells' liquid flags, clamped at the
 * theta criterion (Pressure.cpp:20); the clamp is a decision a replay keeps */
static double face_theta(wco_scene* s, int d, int64_t fn, int64_t nl, int64_t nr, double phil,
                         double phir) {
  const int ll = nl >= 0 && is_liquid(s, nl), lr = is_liquid(s, nr);
  double th;
  if (ll && lr)
    th = 1;
  else if (ll)
    th = phil / (phil - phir);
  else if (lr)
    th = phir / (phir - phil);
  else
    th = 0;
  int clamp;
  if (s->frz_mode == 2) {
    clamp = s->frz_clamp[d][fn];
  } else {
    clamp = th < THETA_CRITERION;
    if (s->frz_mode == 1) s->frz_clamp[d][fn] = (uint8_t)clamp;
  }
  return clamp ? THETA_CRITERION : th;
}

/* ----------------------------------------------------------- math kernels */

/* MathUtilities.h:416-424 */
double wco_quad_kernel(double x) {
  const double ax = fabs(x);
  if (ax < 0.5) return 0.75 - ax * ax;
  if (ax < 1.5) return 0.5 * (1.5 - ax) * (1.5 - ax);
  return 0.0;
}

/* MathUtilities.h:286-293 */
double wco_fraction_inside2(double phi_left, double phi_right) {
  if (phi_left < 0 && phi_right < 0) return 1;
  if (phi_left < 0 && phi_right >= 0) return phi_left / (phi_left - phi_right);
  if (phi_left >= 0 && phi_right < 0) return phi_right / (phi_right - phi_left);
  return 0;
}

static void rotate4(double* a) {
  const double t = a[0];
  a[0] = a[1];
  a[1] = a[2];
  a[2] = a[3];
  a[3] = t;
}

/* MathUtilities.h:301-375: area fraction of a square with corner values
 * (bl, br, tl, tr) that lies where phi < 0; corners are walked in the cyclic
 * order bl, br, tr, tl. */
double wco_fraction_inside4(double bl, double br, double tl, double tr) {
  const int inside = (bl < 0) + (tl < 0) + (br < 0) + (tr < 0);
  double c[4] = {bl, br, tr, tl};
  if (inside == 4) return 1.0;
  if (inside == 3) {
    while (c[0] < 0) rotate4(c);
    const double s0 = 1.0 - wco_fraction_inside2(c[0], c[3]);
    const double s1 = 1.0 - wco_fraction_inside2(c[0], c[1]);
    return 1.0 - 0.5 * s0 * s1;
  }
  if (inside == 2) {
    while (c[0] >= 0 || !(c[1] < 0 || c[2] < 0)) rotate4(c);
    if (c[1] < 0) {
      const double sl = wco_fraction_inside2(c[0], c[3]);
      const double sr = wco_fraction_inside2(c[1], c[2]);
      return 0.5 * (sl + sr);
    }
    const double mid = 0.25 * (c[0] + c[1] + c[2] + c[3]);
    if (mid < 0) {
      double area = 0.0;
      const double s1 = 1.0 - wco_fraction_inside2(c[0], c[3]);
      const double s3 = 1.0 - wco_fraction_inside2(c[2], c[3]);
      area += 0.5 * s1 * s3;
      const double s2 = 1 - wco_fraction_inside2(c[2], c[1]);
      const double s0 = 1 - wco_fraction_inside2(c[0], c[1]);
      area += 0.5 * s0 * s2;
      return 1.0 - area;
    } else {
      double area = 0.0;
      const double s0 = wco_fraction_inside2(c[0], c[1]);
      const double s1 = wco_fraction_inside2(c[0], c[3]);
      area += 0.5 * s0 * s1;
      const double s2 = wco_fraction_inside2(c[2], c[1]);
      const double s3 = wco_fraction_inside2(c[2], c[3]);
      area += 0.5 * s2 * s3;
      return area;
    }
  }
  if (inside == 1) {
    while (c[0] >= 0) rotate4(c);
    const double s0 = wco_fraction_inside2(c[0], c[3]);
    const double s1 = wco_fraction_inside2(c[0], c[1]);
    return 0.5 * s0 * s1;
  }
  return 0;
}

/* DistanceFields.cpp:68-78 box_phi with zero corner radius, negated for a
 * container (inside=1 gives sign -1, DistanceFields.cpp:486; identity
 * rotation, DistanceFields.cpp:601-609): positive inside the box. */
double wco_box_container_phi(const double pos[3], const double lo[3],
                             const double hi[3]) {
  double dd[3], da[3];
  for (int d = 0; d < 3; ++d) {
    const double c = 0.5 * (lo[d] + hi[d]);
    const double e = 0.5 * (hi[d] - lo[d]);
    dd[d] = fabs(pos[d] - c) - e;
    da[d] = dd[d] > 0.0 ? dd[d] : 0.0;
  }
  double m = dd[0] > dd[1] ? dd[0] : dd[1];
  m = m > dd[2] ? m : dd[2];
  if (m > 0.0) m = 0.0;
  const double phi = m + sqrt(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]);
  return -phi;
}

/* ---------------------------------------------------------------- seeding */

static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t wco_hash3(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(mix64(mix64(seed) ^ a) ^ b);
}

double wco_uniform(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

static void block_cells(const wco_grid* g, const double lo[3],
                        const double hi[3], int c0[3], int cn[3]) {
  for (int d = 0; d < 3; ++d) {
    c0[d] = (int)floor(lo[d] / g->dx);
    const int c1 = (int)ceil(hi[d] / g->dx);
    cn[d] = c1 > c0[d] ? c1 - c0[d] : 0;
  }
}

int64_t wco_seed_block(const wco_grid* g, const double lo[3], const double hi[3],
                       uint64_t seed, double* x, int64_t cap) {
  int c0[3], cn[3];
  block_cells(g, lo, hi, c0, cn);
  int64_t count = 0;
  const double dx = g->dx;
  for (int kk = 0; kk < cn[2]; ++kk)
    for (int jj = 0; jj < cn[1]; ++jj)
      for (int ii = 0; ii < cn[0]; ++ii)
        for (int sub = 0; sub < 8; ++sub) {
          const uint64_t id =
              ((((uint64_t)kk * cn[1] + jj) * cn[0] + ii) << 3) | sub;
          const int cc[3] = {c0[0] + ii, c0[1] + jj, c0[2] + kk};
          const int sc[3] = {sub & 1, (sub >> 1) & 1, (sub >> 2) & 1};
          double p[3];
          int keep = 1;
          for (int d = 0; d < 3; ++d) {
            const double u = wco_uniform(wco_hash3(seed, id, (uint64_t)d));
            p[d] = ((double)cc[d] + ((double)sc[d] + u) * 0.5) * dx;
            if (!(p[d] >= lo[d] && p[d] < hi[d])) keep = 0;
          }
          if (!keep) continue;
          if (x && count < cap) {
            for (int d = 0; d < 3; ++d) x[(int64_t)d * cap + count] = p[d];
          }
          ++count;
        }
  return count;
}

/* Irwin-Hall: the sum of 12 uniforms minus 6 has mean 0 and variance 1. */
static double noise_cell(uint64_t seed, int i, int j, int k, int comp) {
  const uint64_t key = ((uint64_t)(k + (1 << 20)) << 42) |
                       ((uint64_t)(j + (1 << 20)) << 21) |
                       (uint64_t)(i + (1 << 20));
  double s = 0.0;
  for (int t = 0; t < 12; ++t)
    s += wco_uniform(wco_hash3(seed ^ 0xA076D1F0C3B5E927ull, key,
                               (uint64_t)(comp * 16 + t)));
  return s - 6.0;
}

/* Stencil of three nodes along one axis for a grid offset by off cells
 * (TwoDScene.cpp:3700-3713): base node and the clamp that keeps it inside
 * the dense grid. */
static int stencil_base(const wco_grid* g, int axis, double x, double off) {
  const double t = x / g->dx + ((double)WCO_GHOST - off);
  const double fl = floor(t);
  const double fr = t - fl;
  int b = (int)fl + (fr > 0.5 ? 0 : -1);
  if (b < 0) b = 0;
  if (b > g->nx[axis] - 3) b = g->nx[axis] - 3;
  return b;
}

static void stencil_weights(const wco_grid* g, double x, int base, double off,
                            double w[3]) {
  for (int a = 0; a < 3; ++a)
    w[a] = wco_quad_kernel((x - node_pos(g, base + a, off)) / g->dx);
}

void wco_seed_velocity(const wco_grid* g, uint64_t seed, double sigma,
                       int64_t np, const double* x, double* v) {
  /* filtered field on world cells [-1, n] stored with offset 1 */
  const int m[3] = {g->n[0] + 2, g->n[1] + 2, g->n[2] + 2};
  const int q[3] = {g->n[0] + 4, g->n[1] + 4, g->n[2] + 4};
  const int64_t mm = (int64_t)m[0] * m[1] * m[2];
  const int64_t qq = (int64_t)q[0] * q[1] * q[2];
  double* raw = (double*)malloc(sizeof(double) * 3 * qq);
  double* fil = (double*)malloc(sizeof(double) * 3 * mm);
  const double scale = sigma / sqrt(27.0);
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < q[2]; ++k)
      for (int j = 0; j < q[1]; ++j)
        for (int i = 0; i < q[0]; ++i)
          raw[c * qq + ((int64_t)k * q[1] + j) * q[0] + i] =
              noise_cell(seed, i - 2, j - 2, k - 2, c);
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < m[2]; ++k)
      for (int j = 0; j < m[1]; ++j)
        for (int i = 0; i < m[0]; ++i) {
          double s = 0.0;
          for (int dk = 0; dk < 3; ++dk)
            for (int dj = 0; dj < 3; ++dj)
              for (int di = 0; di < 3; ++di)
                s += raw[c * qq + ((int64_t)(k + dk) * q[1] + (j + dj)) * q[0] +
                         (i + di)];
          fil[c * mm + ((int64_t)k * m[1] + j) * m[0] + i] = s * scale;
        }
  for (int64_t p = 0; p < np; ++p) {
    const double px[3] = {x[p], x[np + p], x[2 * np + p]};
    int base[3];
    double w[3][3];
    for (int d = 0; d < 3; ++d) {
      base[d] = stencil_base(g, d, px[d], 0.5);
      stencil_weights(g, px[d], base[d], 0.5, w[d]);
    }
    for (int c = 0; c < 3; ++c) {
      double acc = 0.0;
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            /* dense cell -> world cell -> filtered-array index */
            int wi = base[0] + c0 - WCO_GHOST + 1;
            int wj = base[1] + c1 - WCO_GHOST + 1;
            int wk = base[2] + c2 - WCO_GHOST + 1;
            double f = 0.0;
            if (wi >= 0 && wj >= 0 && wk >= 0 && wi < m[0] && wj < m[1] &&
                wk < m[2])
              f = fil[c * mm + ((int64_t)wk * m[1] + wj) * m[0] + wi];
            const double ww = w[0][c0] * w[1][c1] * w[2][c2];
            acc += ww * f;
          }
      v[(int64_t)c * np + p] = acc;
    }
  }
  free(raw);
  free(fil);
}

/* ------------------------------------------------------------------ scene */

static void* xcalloc(size_t n, size_t sz) {
  void* p = calloc(n ? n : 1, sz);
  if (!p) abort();
  return p;
}

wco_scene* wco_scene_create(int ni, int nj, int nk, double dx, double rho,
                            double flip, double gx, double gy, double gz,
                            int64_t np, const double* x, const double* v,
                            const double* C, double pmass, double pvol,
                            double radius) {
  wco_scene* s = (wco_scene*)xcalloc(1, sizeof(wco_scene));
  wco_grid_init(&s->g, ni, nj, nk, dx);
  s->rho = rho;
  s->flip = flip;
  s->friction = 1.0;
  s->radius = radius;
  s->pmass = pmass;
  s->pvol = pvol;
  s->gravity[0] = gx;
  s->gravity[1] = gy;
  s->gravity[2] = gz;
  s->solid_projection = 1;
  for (int d = 0; d < 3; ++d) {
    s->box_lo[d] = 0.0;
    s->box_hi[d] = (double)s->g.n[d] * dx;
  }
  s->np = np;
  s->x = (double*)xcalloc(3 * np, sizeof(double));
  s->v = (double*)xcalloc(3 * np, sizeof(double));
  s->C = (double*)xcalloc(9 * np, sizeof(double));
  if (x) memcpy(s->x, x, sizeof(double) * 3 * np);
  if (v) memcpy(s->v, v, sizeof(double) * 3 * np);
  if (C) memcpy(s->C, C, sizeof(double) * 9 * np);
  const int64_t nn = s->g.nnode;
  s->act = (uint8_t*)xcalloc(s->g.nbucket, 1);
  s->lphi = (double*)xcalloc(nn, sizeof(double));
  s->sphi = (double*)xcalloc(nn, sizeof(double));
  for (int d = 0; d < 3; ++d) {
    s->w[d] = (double*)xcalloc(nn, sizeof(double));
    s->mass[d] = (double*)xcalloc(nn, sizeof(double));
    s->vel[d] = (double*)xcalloc(nn, sizeof(double));
    s->vol[d] = (double*)xcalloc(nn, sizeof(double));
    s->rhs[d] = (double*)xcalloc(nn, sizeof(double));
    s->invC[d] = (double*)xcalloc(nn, sizeof(double));
    s->ustar[d] = (double*)xcalloc(nn, sizeof(double));
    s->psi_fs[d] = (double*)xcalloc(nn, sizeof(double));
    s->unew[d] = (double*)xcalloc(nn, sizeof(double));
    s->valid[d] = (uint8_t*)xcalloc(nn, 1);
  }
  s->dof = (int32_t*)xcalloc(nn, sizeof(int32_t));
  s->pressure = (double*)xcalloc(nn, sizeof(double));
  s->lphi_arg = (int32_t*)xcalloc(nn, sizeof(int32_t));
  return s;
}

static void free_system(wco_scene* s) {
  free(s->rowstart);
  free(s->colidx);
  free(s->val);
  free(s->b);
  free(s->p);
  s->rowstart = NULL;
  s->colidx = NULL;
  s->val = NULL;
  s->b = NULL;
  s->p = NULL;
}

void wco_scene_destroy(wco_scene* s) {
  if (!s) return;
  free(s->x);
  free(s->v);
  free(s->C);
  free(s->act);
  free(s->lphi);
  free(s->sphi);
  for (int d = 0; d < 3; ++d) {
    free(s->w[d]);
    free(s->mass[d]);
    free(s->vel[d]);
    free(s->vol[d]);
    free(s->rhs[d]);
    free(s->invC[d]);
    free(s->ustar[d]);
    free(s->psi_fs[d]);
    free(s->unew[d]);
    free(s->valid[d]);
  }
  free(s->dof);
  free(s->pressure);
  free(s->lphi_arg);
  free(s->frz_act);
  free(s->frz_arg);
  free(s->frz_liquid);
  free(s->frz_dof);
  for (int d = 0; d < 3; ++d) {
    free(s->frz_clamp[d]);
    free(s->frz_mass[d]);
  }
  free(s->frz_proj);
  mix_free(s->mix);
  free_system(s);
  free(s);
}

/* Bucket activation: buckets holding any stencil node of a liquid particle on
 * the u, v, w, corner or centre grids (findNodes, TwoDScene.cpp:3683-3752,
 * called from resampleNodes :5076-5085), then one ring of neighbours
 * (expandFluidNodesMarked(1), TwoDScene.cpp:5021-5064). */
void wco_activate(wco_scene* s) {
  const wco_grid* g = &s->g;
  if (s->frz_mode == 2) {
    memcpy(s->act, s->frz_act, g->nbucket);
    return;
  }
  uint8_t* touched = (uint8_t*)xcalloc(g->nbucket, 1);
  static const double offs[5][3] = {{0.0, 0.5, 0.5},
                                    {0.5, 0.0, 0.5},
                                    {0.5, 0.5, 0.0},
                                    {0.0, 0.0, 0.0},
                                    {0.5, 0.5, 0.5}};
  if (s->mix) mix_touch(s, touched);
  for (int64_t p = 0; p < s->np; ++p) {
    const double px[3] = {s->x[p], s->x[s->np + p], s->x[2 * s->np + p]};
    for (int gi = 0; gi < 5; ++gi) {
      int blo[3], bhi[3];
      for (int d = 0; d < 3; ++d) {
        const int b = stencil_base(g, d, px[d], offs[gi][d]);
        blo[d] = b >> 2;
        bhi[d] = (WCO_PLANT == 3 ? b : b + 2) >> 2;
      }
      for (int bk = blo[2]; bk <= bhi[2]; ++bk)
        for (int bj = blo[1]; bj <= bhi[1]; ++bj)
          for (int bi = blo[0]; bi <= bhi[0]; ++bi)
            touched[wco_bucket_index(g, bi, bj, bk)] = 1;
    }
  }
  for (int bk = 0; bk < g->nb[2]; ++bk)
    for (int bj = 0; bj < g->nb[1]; ++bj)
      for (int bi = 0; bi < g->nb[0]; ++bi) {
        int a = 0;
        const int ring = WCO_PLANT == 1 ? 0 : (WCO_PLANT == 2 ? 2 : 1);
        for (int t = -ring; t <= ring && !a; ++t)
          for (int u = -ring; u <= ring && !a; ++u)
            for (int r = -ring; r <= ring && !a; ++r) {
              const int x = bi + r, y = bj + u, z = bk + t;
              if (x < 0 || y < 0 || z < 0 || x >= g->nb[0] || y >= g->nb[1] ||
                  z >= g->nb[2])
                continue;
              if (touched[wco_bucket_index(g, x, y, z)]) a = 1;
            }
        s->act[wco_bucket_index(g, bi, bj, bk)] = (uint8_t)a;
      }
  free(touched);
  if (s->frz_mode == 1) memcpy(s->frz_act, s->act, g->nbucket);
}

static double cell_solid_phi(const wco_scene* s, int i, int j, int k) {
  const double p[3] = {node_pos(&s->g, i, 0.5), node_pos(&s->g, j, 0.5),
                       node_pos(&s->g, k, 0.5)};
  return wco_box_container_phi(p, s->box_lo, s->box_hi);
}

/* TwoDScene.cpp:3286-3373 updateLiquidPhi (no surface tension/viscosity) */
void wco_update_liquid_phi(wco_scene* s) {
  const wco_grid* g = &s->g;
  const double dx = g->dx;
  const double bucket_size = WCO_BUCKET * dx;
  for (int64_t c = 0; c < g->nnode; ++c) {
    s->lphi[c] = 3.0 * bucket_size;
    s->lphi_arg[c] = -1;
  }
  const double r0 = WCO_PLANT == 12 ? s->radius
                                     : (dx * 0.883644 > s->radius ? dx * 0.883644 : s->radius);
  if (s->frz_mode == 2) {
    /* the recorded particle of each cell gives its distance */
    for (int k = 0; k < g->nx[2]; ++k)
      for (int j = 0; j < g->nx[1]; ++j)
        for (int i = 0; i < g->nx[0]; ++i) {
          const int64_t n = wco_node_index(g, i, j, k);
          const int32_t a = s->frz_arg[n];
          s->lphi_arg[n] = a;
          if (a == -2) {
            s->lphi[n] = -0.5 * dx;
          } else if (a >= 0) {
            const double d0 = s->x[a] - node_pos(g, i, 0.5);
            const double d1 = s->x[s->np + a] - node_pos(g, j, 0.5);
            const double d2 = s->x[2 * s->np + a] - node_pos(g, k, 0.5);
            s->lphi[n] = sqrt(d0 * d0 + d1 * d1 + d2 * d2) - r0;
          }
        }
    return;
  }
  for (int64_t p = 0; p < s->np; ++p) {
    const double px[3] = {s->x[p], s->x[s->np + p], s->x[2 * s->np + p]};
    int base[3];
    for (int d = 0; d < 3; ++d) base[d] = stencil_base(g, d, px[d], 0.5);
    for (int c2 = 0; c2 < 3; ++c2)
      for (int c1 = 0; c1 < 3; ++c1)
        for (int c0 = 0; c0 < 3; ++c0) {
          const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
          if (!node_active(s, i, j, k)) continue;
          const double d0 = px[0] - node_pos(g, i, 0.5);
          const double d1 = px[1] - node_pos(g, j, 0.5);
          const double d2 = px[2] - node_pos(g, k, 0.5);
          const double phi = sqrt(d0 * d0 + d1 * d1 + d2 * d2) - r0;
          const int64_t n = wco_node_index(g, i, j, k);
          if (phi < s->lphi[n]) {
            s->lphi[n] = phi;
            s->lphi_arg[n] = (int32_t)p;
          }
        }
  }
  for (int k = 0; k < g->nx[2]; ++k)
    for (int j = 0; j < g->nx[1]; ++j)
      for (int i = 0; i < g->nx[0]; ++i) {
        if (!node_active(s, i, j, k) || WCO_PLANT == 4) continue;
        if (cell_solid_phi(s, i, j, k) < 0.0) {
          s->lphi[wco_node_index(g, i, j, k)] = -0.5 * dx;
          s->lphi_arg[wco_node_index(g, i, j, k)] = -2;
        }
      }
  if (s->frz_mode == 1) {
    memcpy(s->frz_arg, s->lphi_arg, sizeof(int32_t) * g->nnode);
    for (int64_t n = 0; n < g->nnode; ++n) s->frz_liquid[n] = s->lphi[n] < 0.0;
  }
}

/* TwoDScene.cpp:7940-7970 updateSolidPhi: corner phi of the container in
 * activated buckets (the container is static: solid face velocity 0). */
void wco_update_solid_phi(wco_scene* s) {
  const wco_grid* g = &s->g;
  for (int k = 0; k < g->nx[2]; ++k)
    for (int j = 0; j < g->nx[1]; ++j)
      for (int i = 0; i < g->nx[0]; ++i) {
        const int64_t n = wco_node_index(g, i, j, k);
        if (!node_active(s, i, j, k)) {
          s->sphi[n] = 0.0;
          continue;
        }
        const double p[3] = {node_pos(g, i, 0.0), node_pos(g, j, 0.0),
                             node_pos(g, k, 0.0)};
        s->sphi[n] = wco_box_container_phi(p, s->box_lo, s->box_hi);
      }
}

/* corner phi as seen by updateSolidWeights: 0.5 dx where the corner's bucket
 * is not activated (TwoDScene.cpp:1899-1911) */
static double corner_phi_w(const wco_scene* s, int i, int j, int k) {
  if (!node_active(s, i, j, k)) return 0.5 * s->g.dx;
  return s->sphi[wco_node_index(&s->g, i, j, k)];
}

/* TwoDScene.cpp:1871-1962 updateSolidWeights with the corner order of
 * connectSolidPhiNodes (TwoDScene.cpp:4447-4590). */
void wco_update_solid_weights(wco_scene* s) {
  const wco_grid* g = &s->g;
  for (int k = 0; k < g->nx[2]; ++k)
    for (int j = 0; j < g->nx[1]; ++j)
      for (int i = 0; i < g->nx[0]; ++i) {
        const int64_t n = wco_node_index(g, i, j, k);
        if (!node_active(s, i, j, k)) {
          s->w[0][n] = s->w[1][n] = s->w[2][n] = 0.0;
          continue;
        }
        double f;
        f = wco_fraction_inside4(
            corner_phi_w(s, i, j, k), corner_phi_w(s, i, j + 1, k),
            corner_phi_w(s, i, j, k + 1), corner_phi_w(s, i, j + 1, k + 1));
        f = 1.0 - f;
        s->w[0][n] = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
        f = wco_fraction_inside4(
            corner_phi_w(s, i, j, k), corner_phi_w(s, i, j, k + 1),
            corner_phi_w(s, i + 1, j, k), corner_phi_w(s, i + 1, j, k + 1));
        f = 1.0 - f;
        s->w[1][n] = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
        f = wco_fraction_inside4(
            corner_phi_w(s, i, j, k), corner_phi_w(s, i, j + 1, k),
            corner_phi_w(s, i + 1, j, k), corner_phi_w(s, i + 1, j + 1, k));
        f = 1.0 - f;
        s->w[2][n] = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
      }
}

/* face-grid offsets (in cells) of the u, v, w grids: TwoDScene.cpp:5076-5081 */
static const double face_off[3][3] = {
    {0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}};

/* TwoDScene.cpp:6609-6919 mapParticleNodesAPIC, fluid branch (:6700-6713,
 * :6706-6712): P = sum fm w (v_d + fB.row(d).(x_n - x_p)), M = sum fm w,
 * V = sum fvol w; u = P/M where M > 1e-20. */
void wco_p2g(wco_scene* s) {
  const wco_grid* g = &s->g;
  const int64_t np = s->np;
  double* mom[3];
  for (int d = 0; d < 3; ++d) {
    memset(s->mass[d], 0, sizeof(double) * g->nnode);
    memset(s->vel[d], 0, sizeof(double) * g->nnode);
    memset(s->vol[d], 0, sizeof(double) * g->nnode);
    mom[d] = (double*)xcalloc(g->nnode, sizeof(double));
  }
  const double fm = s->pmass, fvol = s->pvol;
  for (int64_t p = 0; p < np; ++p) {
    const double px[3] = {s->x[p], s->x[np + p], s->x[2 * np + p]};
    for (int d = 0; d < 3; ++d) {
      int base[3];
      double w[3][3];
      for (int a = 0; a < 3; ++a) {
        base[a] = stencil_base(g, a, px[a], face_off[d][a]);
        stencil_weights(g, px[a], base[a], face_off[d][a], w[a]);
      }
      const double vd = s->v[(int64_t)d * np + p];
      const double b0 = s->C[(int64_t)(d * 3 + 0) * np + p];
      const double b1 = s->C[(int64_t)(d * 3 + 1) * np + p];
      const double b2 = s->C[(int64_t)(d * 3 + 2) * np + p];
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
            if (!node_active(s, i, j, k)) continue;
            const double ww = w[0][c0] * w[1][c1] * w[2][c2];
            const double e0 = node_pos(g, i, face_off[d][0]) - px[0];
            const double e1 = node_pos(g, j, face_off[d][1]) - px[1];
            const double e2 = node_pos(g, k, face_off[d][2]) - px[2];
            const double vel = vd + (b0 * e0 + b1 * e1 + b2 * e2);
            const int64_t n = wco_node_index(g, i, j, k);
            mom[d][n] += vel * fm * ww;
            s->mass[d][n] += fm * ww;
            s->vol[d][n] += fvol * ww;
          }
    }
  }
  for (int d = 0; d < 3; ++d) {
    for (int64_t n = 0; n < g->nnode; ++n) {
      if (s->frz_mode == 1) s->frz_mass[d][n] = s->mass[d][n] > 1e-20;
      if (has_mass(s, d, n)) s->vel[d][n] = mom[d][n] / s->mass[d][n];
    }
    free(mom[d]);
  }
  if (s->mix) mix_p2g(s);
}

/* LinearizedImplicitEuler.cpp:224-337 stepVelocity for liquid without
 * solids (psi = 0, no elasto mass, drag coefficient 0 by TwoDScene.cpp:4047):
 *   rhs = 0 - M g (-dt) + M u          (LIE:548-659, SimpleGravityForce:26-54;
 *                                        the pore-pressure term adds 0)
 *   C = M + Q*0 with Q = 1 (fallback, LIE:950-955); invC = 1/C or 1 (LIE:1122)
 *   u* = rhs/M if M > 1e-20 else rhs   (LIE:128-175)
 *   psi_fs = ((1-psi) invC + psi invCs P) M = invC M (LIE:1052-1107) */
void wco_step_velocity(wco_scene* s, double dt) {
  const wco_grid* g = &s->g;
  if (s->mix) {
    mix_step_velocity(s, dt);
    return;
  }
  for (int d = 0; d < 3; ++d) {
    const double gd = s->gravity[d];
    for (int64_t n = 0; n < g->nnode; ++n) {
      if (!s->act[n >> 6]) {
        s->rhs[d][n] = s->invC[d][n] = s->ustar[d][n] = s->psi_fs[d][n] = 0.0;
        continue;
      }
      const double M = s->mass[d][n];
      double rhs = 0.0;
      if (WCO_PLANT == 10)
        rhs += M * gd * (-dt);
      else if (WCO_PLANT == 11)
        rhs -= M * gd * (-1.0);
      else
        rhs -= M * gd * (-dt);
      rhs += M * s->vel[d][n];
      s->rhs[d][n] = rhs;
      const double Q = 1.0, Ms = 0.0;
      const double Cc = M + Q * Ms;
      const int hm = has_mass(s, d, n); /* M > 1e-20 (and so Cc) */
      const double invC = hm ? 1.0 / Cc : 1.0;
      s->invC[d][n] = invC;
      s->ustar[d][n] = hm ? rhs / M : rhs;
      const double psi = 0.0, invCs = 0.0;
      const double P = hm ? 0.0 / M : 1.0;
      s->psi_fs[d][n] = ((1.0 - psi) * invC + psi * invCs * P) * M;
    }
  }
}

/* ------------------------------------------------------- pressure system */

typedef struct {
  uint32_t col;
  double val;
} entry;

static void insert_sorted(entry* row, int* len, uint32_t col, double v) {
  /* SparseMatrix::add_to_element (SparseMatrix.h:88-103) */
  int k = 0;
  for (; k < *len; ++k) {
    if (row[k].col == col) {
      row[k].val += v;
      return;
    }
    if (row[k].col > col) break;
  }
  for (int t = *len; t > k; --t) row[t] = row[t - 1];
  row[k].col = col;
  row[k].val = v;
  ++*len;
}

static int cell_exists(const wco_scene* s, int i, int j, int k) {
  return node_active(s, i, j, k);
}

/* Pressure.cpp:285-791 solveNodePressure: DOF classification (:348-392),
 * bucket-major numbering (:394-439, partial sum over buckets in bucket
 * order), then the 7-point rows in face order left, right, bottom, top,
 * near, far (:444-791). */
int64_t wco_assemble(wco_scene* s, double dt) {
  const wco_grid* g = &s->g;
  const double dx = g->dx;
  const double coeff = dt / (dx * dx);
  free_system(s);
  int64_t ndof = 0;
  for (int64_t n = 0; n < g->nnode; ++n) {
    s->dof[n] = -1;
    s->pressure[n] = 0.0;
  }
  /* storage order is bucket-major, so a running count is the partial sum */
  for (int64_t bucket = 0; bucket < g->nbucket && s->frz_mode != 2; ++bucket) {
    if (!s->act[bucket]) continue;
    const int bk = (int)(bucket / ((int64_t)g->nb[0] * g->nb[1]));
    const int bj = (int)((bucket / g->nb[0]) % g->nb[1]);
    const int bi = (int)(bucket % g->nb[0]);
    for (int l = 0; l < 64; ++l) {
      const int i = bi * 4 + (l & 3), j = bj * 4 + ((l >> 2) & 3),
                k = bk * 4 + (l >> 4);
      const int64_t n = (bucket << 6) | l;
      if (!(s->lphi[n] < 0)) continue;
      /* faces: x at (i),(i+1); y at (j),(j+1); z at (k),(k+1) */
      const int fi[6][3] = {{i, j, k},     {i + 1, j, k}, {i, j, k},
                            {i, j + 1, k}, {i, j, k},     {i, j, k + 1}};
      int open = 0;
      for (int f = 0; f < 6 && !open; ++f) {
        const int* q = fi[f];
        if (!node_active(s, q[0], q[1], q[2])) continue;
        const double wf = s->w[f >> 1][wco_node_index(g, q[0], q[1], q[2])];
        if (WCO_PLANT == 5 ? wf >= 0 : wf > 0) open = 1;
      }
      if (open) s->dof[n] = (int32_t)ndof++;
    }
  }
  if (s->frz_mode == 2) {
    memcpy(s->dof, s->frz_dof, sizeof(int32_t) * g->nnode);
    ndof = s->frz_ndof;
  } else if (s->frz_mode == 1) {
    memcpy(s->frz_dof, s->dof, sizeof(int32_t) * g->nnode);
    s->frz_ndof = ndof;
  }
  if (WCO_PLANT == 6) { /* planted: numbered i-fastest over the whole grid */
    int64_t c = 0;
    for (int k = 0; k < g->nx[2]; ++k)
      for (int j = 0; j < g->nx[1]; ++j)
        for (int i = 0; i < g->nx[0]; ++i) {
          const int64_t n = wco_node_index(g, i, j, k);
          if (s->dof[n] >= 0) s->dof[n] = (int32_t)c++;
        }
  }
  s->ndof = ndof;
  s->rowstart = (uint32_t*)xcalloc(ndof + 1, sizeof(uint32_t));
  s->colidx = (uint32_t*)xcalloc(7 * ndof, sizeof(uint32_t));
  s->val = (double*)xcalloc(7 * ndof, sizeof(double));
  s->b = (double*)xcalloc(ndof, sizeof(double));
  s->p = (double*)xcalloc(ndof, sizeof(double));
  uint32_t nnz = 0;
  for (int64_t bucket = 0; bucket < g->nbucket; ++bucket) {
    if (!s->act[bucket]) continue;
    const int bk = (int)(bucket / ((int64_t)g->nb[0] * g->nb[1]));
    const int bj = (int)((bucket / g->nb[0]) % g->nb[1]);
    const int bi = (int)(bucket % g->nb[0]);
    for (int l = 0; l < 64; ++l) {
      const int64_t n = (bucket << 6) | l;
      const int32_t me = s->dof[n];
      if (me < 0) continue;
      const int i = bi * 4 + (l & 3), j = bj * 4 + ((l >> 2) & 3),
                k = bk * 4 + (l >> 4);
      const double center_phi = s->lphi[n];
      entry row[7];
      int len = 0;
      double rhs = 0.0;
      /* face f: axis f>>1; low side (f even) adds, high side subtracts */
      for (int f = 0; f < 6; ++f) {
        const int axis = f >> 1, hi = f & 1;
        int q[3] = {i, j, k};
        q[axis] += hi;
        if (!node_active(s, q[0], q[1], q[2])) continue;
        const int64_t fn = wco_node_index(g, q[0], q[1], q[2]);
        const double w = s->w[axis][fn];
        const double uf = s->ustar[axis][fn], pfs = s->psi_fs[axis][fn];
        double psi = 0.0, Vs = 0.0, Q = 1.0, Pm = 0.0, psf = 0.0, Ms = 0.0;
        if (s->mix) mix_face(s, axis, fn, &psi, &Vs, &Q, &Pm, &psf, &Ms);
        /* the solid's grid velocity m_node_v_plus: zero (not solved) */
        const double us = 0.0, usolid = 0.0;
        const double flux = ((uf * pfs + us * psf) * w + usolid * (1.0 - w)) / dx;
        if (hi)
          rhs -= flux;
        else
          rhs += flux;
        if (w > 0.0) {
          const double invCs = 0.0; /* constructMsDVs runs only when the solid is solved */
          const double M = s->mass[axis][fn];
          const double P = s->mix ? Pm : (M > 1e-20 ? 0.0 / M : 1.0);
          const double Vf = s->vol[axis][fn];
          double term = w * coeff;
          term *= (WCO_PLANT == 20 ? 1.0 : 1.0 - psi) * s->invC[axis][fn] * (Vf + Vs * Q) +
                  psi * invCs * (Vs + Vf * P);
          int nb[3] = {i, j, k};
          nb[axis] += hi ? 1 : -1;
          if (!cell_exists(s, nb[0], nb[1], nb[2])) continue;
          const int64_t nn = wco_node_index(g, nb[0], nb[1], nb[2]);
          const double nphi = s->lphi[nn];
          if (is_liquid(s, nn)) {
            const int32_t other = s->dof[nn];
            if (other >= 0) {
              insert_sorted(row, &len, (uint32_t)me, term);
              insert_sorted(row, &len, (uint32_t)other, -term);
            }
          } else {
            /* the face between this cell and the air cell nn */
            const double theta = hi ? face_theta(s, axis, fn, n, nn, center_phi, nphi)
                                    : face_theta(s, axis, fn, nn, n, nphi, center_phi);
            insert_sorted(row, &len, (uint32_t)me, term / theta);
          }
        }
      }
      s->b[me] = rhs;
      s->rowstart[me] = nnz;
      for (int t = 0; t < len; ++t) {
        s->colidx[nnz] = row[t].col;
        s->val[nnz] = row[t].val;
        ++nnz;
      }
    }
  }
  s->rowstart[ndof] = nnz;
  s->nnz_cap = 7 * ndof;
  return ndof;
}

/* ------------------------------------------------------------ PCGSolver */

/* SparseMatrix.h:298-322 multiply */
void wco_spmv(int64_t n, const uint32_t* rowstart, const uint32_t* col,
              const double* val, const double* x, double* y) {
  for (int64_t i = 0; i < n; ++i) {
    double r = 0;
    for (uint32_t e = rowstart[i]; e < rowstart[i + 1]; ++e) r += val[e] * x[col[e]];
    y[i] = r;
  }
}

/* first entry of row i strictly above the diagonal */
static uint32_t upper_start(const uint32_t* rowstart, const uint32_t* col,
                            int64_t i) {
  uint32_t e = rowstart[i];
  while (e < rowstart[i + 1] && col[e] <= (uint32_t)i) ++e;
  return e;
}

/* PCGSolver.h:97-205 factor_modified_incomplete_cholesky0.  The factor's
 * column k is row k's strictly-upper part, so lv is indexed like val. */
void wco_mic0_factor(int64_t n, const uint32_t* rowstart, const uint32_t* col,
                     const double* val, double tau, double mdr,
                     double* invdiag, double* lv) {
  double* adiag = (double*)xcalloc(n, sizeof(double));
  uint32_t* cs = (uint32_t*)xcalloc(n + 1, sizeof(uint32_t));
  for (int64_t i = 0; i < n; ++i) {
    invdiag[i] = 0;
    cs[i] = upper_start(rowstart, col, i);
    for (uint32_t e = rowstart[i]; e < rowstart[i + 1]; ++e) {
      if (col[e] > (uint32_t)i)
        lv[e] = val[e];
      else if (col[e] == (uint32_t)i)
        invdiag[i] = adiag[i] = val[e];
    }
  }
  for (int64_t k = 0; k < n; ++k) {
    if (adiag[k] == 0) continue;
    if (invdiag[k] < mdr * adiag[k])
      invdiag[k] = 1 / sqrt(adiag[k]);
    else
      invdiag[k] = 1 / sqrt(invdiag[k]);
    const uint32_t c0 = cs[k], c1 = rowstart[k + 1];
    for (uint32_t p = c0; p < c1; ++p) lv[p] *= invdiag[k];
    for (uint32_t p = c0; p < c1; ++p) {
      const uint32_t j = col[p];
      const double multiplier = lv[p];
      double missing = 0;
      uint32_t a = c0;
      uint32_t b = rowstart[j];
      while (a < c1 && col[a] < j) {
        while (b < rowstart[j + 1]) {
          if (col[b] < col[a])
            ++b;
          else if (col[b] == col[a])
            break;
          else {
            missing += lv[a];
            break;
          }
        }
        ++a;
      }
      if (a < c1 && col[a] == j) invdiag[j] -= multiplier * lv[a];
      ++a;
      b = cs[j];
      const uint32_t bend = rowstart[j + 1];
      while (a < c1 && b < bend) {
        if (col[b] < col[a])
          ++b;
        else if (col[b] == col[a]) {
          lv[b] -= multiplier * lv[a];
          ++a;
          ++b;
        } else {
          missing += lv[a];
          ++a;
        }
      }
      while (a < c1) {
        missing += lv[a];
        ++a;
      }
      invdiag[j] -= tau * multiplier * missing;
    }
  }
  free(adiag);
  free(cs);
}

/* PCGSolver.h:211-223 solve_lower (column-oriented forward substitution) */
void wco_solve_lower(int64_t n, const uint32_t* rowstart, const uint32_t* col,
                     const double* invdiag, const double* lv,
                     const double* rhs, double* out) {
  memcpy(out, rhs, sizeof(double) * n);
  for (int64_t i = 0; i < n; ++i) {
    out[i] *= invdiag[i];
    for (uint32_t e = upper_start(rowstart, col, i); e < rowstart[i + 1]; ++e)
      out[col[e]] -= lv[e] * out[i];
  }
}

/* PCGSolver.h:226-239 solve_lower_transpose_in_place */
void wco_solve_lower_transpose(int64_t n, const uint32_t* rowstart,
                               const uint32_t* col, const double* invdiag,
                               const double* lv, double* x) {
  for (int64_t i = n - 1; i >= 0; --i) {
    for (uint32_t e = upper_start(rowstart, col, i); e < rowstart[i + 1]; ++e)
      x[i] -= lv[e] * x[col[e]];
    x[i] *= invdiag[i];
  }
}

/* BLASWrapper.h:24-29 (sequential sum; Eigen's dot is SIMD-blocked) */
double wco_dot(int64_t n, const double* a, const double* b) {
  double s = 0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* BLASWrapper.h:33-50 */
double wco_abs_max(int64_t n, const double* a) {
  double m = 0;
  for (int64_t i = 0; i < n; ++i)
    if (fabs(a[i]) > m) m = fabs(a[i]);
  return m;
}

/* PCGSolver.h:260-308 PCGSolver::solve */
int wco_pcg_solve(int64_t n, const uint32_t* rowstart, const uint32_t* col,
                  const double* val, const double* rhs, double* x, double tol,
                  int maxit, double tau, double mdr, double* residual_out,
                  int* iters_out) {
  if (tol < 1e-30) tol = 1e-30; /* PCGSolver.h:253 */
  for (int64_t i = 0; i < n; ++i) x[i] = 0;
  if (n == 0) {
    *residual_out = 0;
    *iters_out = 0;
    return 1;
  }
  double* r = (double*)xcalloc(n, sizeof(double));
  double* z = (double*)xcalloc(n, sizeof(double));
  double* sv = (double*)xcalloc(n, sizeof(double));
  double* invdiag = (double*)xcalloc(n, sizeof(double));
  double* lv = (double*)xcalloc(rowstart[n] ? rowstart[n] : 1, sizeof(double));
  memcpy(r, rhs, sizeof(double) * n);
  int ok = 0, it = 0;
  double res = wco_abs_max(n, r);
  *residual_out = res;
  if (res < 1e-30) {
    *iters_out = 0;
    ok = 1;
    goto done;
  }
  {
    const double tolr = tol * res;
    wco_mic0_factor(n, rowstart, col, val, tau, mdr, invdiag, lv);
    wco_solve_lower(n, rowstart, col, invdiag, lv, r, z);
    wco_solve_lower_transpose(n, rowstart, col, invdiag, lv, z);
    double rho = wco_dot(n, z, r);
    if (rho == 0 || rho != rho) {
      *iters_out = 0;
      ok = 0;
      goto done;
    }
    memcpy(sv, z, sizeof(double) * n);
    for (it = 0; it < maxit; ++it) {
      wco_spmv(n, rowstart, col, val, sv, z);
      const double alpha = rho / wco_dot(n, sv, z);
      for (int64_t i = 0; i < n; ++i) x[i] += sv[i] * alpha;
      for (int64_t i = 0; i < n; ++i) r[i] += z[i] * -alpha;
      res = wco_abs_max(n, r);
      *residual_out = res;
      if (res <= tolr) {
        *iters_out = it + 1;
        ok = 1;
        goto done;
      }
      wco_solve_lower(n, rowstart, col, invdiag, lv, r, z);
      wco_solve_lower_transpose(n, rowstart, col, invdiag, lv, z);
      const double rho_new = wco_dot(n, z, r);
      const double beta = rho_new / rho;
      for (int64_t i = 0; i < n; ++i) z[i] += sv[i] * beta;
      double* t = sv;
      sv = z;
      z = t;
      rho = rho_new;
    }
    *iters_out = it;
    ok = 0;
  }
done:
  free(r);
  free(z);
  free(sv);
  free(invdiag);
  free(lv);
  return ok;
}

int wco_solve_pressure(wco_scene* s, double tol, int maxit) {
  double res = 0;
  int it = 0;
  const int ok = wco_pcg_solve(s->ndof, s->rowstart, s->colidx, s->val, s->b,
                               s->p, tol, maxit, 0.97, 0.25, &res, &it);
  s->last_residual = res;
  s->last_iters = it;
  s->last_ok = ok;
  for (int64_t n = 0; n < s->g.nnode; ++n)
    if (s->dof[n] >= 0) s->pressure[n] = s->p[s->dof[n]];
  return ok;
}

/* LIE:2435-2504 applyPressureDragFluid (drag_by_future_solid = false):
 * addSolidDragRHS adds Q*Ms*us = 0 (LIE:689-717); applyPressureGradsFluidRHS
 * (Pressure.cpp:1137-1314); performInvLocalSolve u = rhs*invC (LIE:65-104);
 * extrapolate x3 (Pressure.cpp:1826-1903); then acceptVelocity (LIE:2176) and
 * constrainLiquidVelocity (TwoDScene.cpp:1827-1866). */
void wco_apply_pressure(wco_scene* s, double dt) {
  const wco_grid* g = &s->g;
  const double dx = g->dx;
  for (int d = 0; d < 3; ++d) {
    for (int k = 0; k < g->nx[2]; ++k)
      for (int j = 0; j < g->nx[1]; ++j)
        for (int i = 0; i < g->nx[0]; ++i) {
          const int64_t n = wco_node_index(g, i, j, k);
          if (!s->act[n >> 6]) {
            s->unew[d][n] = 0.0;
            s->valid[d][n] = 0;
            continue;
          }
          double psi = 0.0, Vs = 0.0, Q = 1.0, Pm = 0.0, psf = 0.0, Ms = 0.0;
          if (s->mix) mix_face(s, d, n, &psi, &Vs, &Q, &Pm, &psf, &Ms);
          const double us = 0.0; /* the solid's P2G velocity: at rest */
          double rhs = s->rhs[d][n];
          rhs += Q * Ms * us;
          uint8_t valid = 0;
          if (s->w[d][n] > 0.) {
            int l[3] = {i, j, k};
            l[d] -= 1;
            const int lex = cell_exists(s, l[0], l[1], l[2]);
            const int rex = 1; /* the right cell shares the face's bucket */
            const int64_t ln = lex ? wco_node_index(g, l[0], l[1], l[2]) : -1;
            const double left_phi = lex ? s->lphi[ln] : 0.0;
            const double right_phi = rex ? s->lphi[n] : 0.0;
            const int lliq = lex && is_liquid(s, ln), rliq = is_liquid(s, n);
            if (lliq || rliq) {
              double theta = 1.;
              if (!rliq || !lliq) theta = face_theta(s, d, n, ln, n, left_phi, right_phi);
              double pg = 0;
              if (lex) pg -= s->pressure[ln];
              if (rex) pg += s->pressure[n];
              pg /= (dx * theta);
              const double term = pg * dt * (s->vol[d][n] + (WCO_PLANT == 21 ? 0.0 : Vs * Q));
              rhs -= term;
              valid = 1;
            }
          }
          s->unew[d][n] = rhs * s->invC[d][n];
          s->valid[d][n] = valid;
        }
  }
  wco_extrapolate(s);
  wco_constrain_liquid_velocity(s);
}

/* Pressure.cpp:1826-1903 extrapolate, called twice per axis from
 * applyPressureDragFluid: each layer gives every invalid face of an
 * activated bucket the mean of its valid 6-neighbours, read from copies of
 * the velocity and the valid flags made before the layer. */
void wco_extrapolate(wco_scene* s) {
  const wco_grid* g = &s->g;
  double* ucopy = (double*)xcalloc(g->nnode, sizeof(double));
  uint8_t* vcopy = (uint8_t*)xcalloc(g->nnode, 1);
  static const int off[6][3] = {{-1, 0, 0}, {1, 0, 0},  {0, -1, 0},
                                {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};
  const int layers = WCO_PLANT == 8 ? 1 : 2;
  for (int d = 0; d < 3; ++d) {
    for (int layer = 0; layer < layers; ++layer) {
      const double* uread = ucopy;
      const uint8_t* vread = vcopy;
      if (WCO_PLANT == 7) { /* planted: reads the faces updated in this layer */
        uread = s->unew[d];
        vread = s->valid[d];
      }
      memcpy(ucopy, s->unew[d], sizeof(double) * g->nnode);
      memcpy(vcopy, s->valid[d], g->nnode);
      for (int k = 0; k < g->nx[2]; ++k)
        for (int j = 0; j < g->nx[1]; ++j)
          for (int i = 0; i < g->nx[0]; ++i) {
            const int64_t n = wco_node_index(g, i, j, k);
            if (!s->act[n >> 6] || vread[n]) continue;
            double sum = 0;
            int count = 0;
            for (int r = 0; r < 6; ++r) {
              const int a = i + off[r][0], b = j + off[r][1], c = k + off[r][2];
              if (!node_active(s, a, b, c)) continue;
              const int64_t m = wco_node_index(g, a, b, c);
              if (!vread[m]) continue;
              sum += uread[m];
              ++count;
            }
            if (count > 0) {
              s->unew[d][n] = sum / (double)count;
              s->valid[d][n] = 1;
            }
          }
    }
  }
  free(ucopy);
  free(vcopy);
}

/* TwoDScene.cpp:1827-1866 constrainLiquidVelocity: faces with zero solid
 * weight are lerped toward the (static) solid velocity by the boundary
 * friction. */
void wco_constrain_liquid_velocity(wco_scene* s) {
  const wco_grid* g = &s->g;
  for (int d = 0; d < 3; ++d)
    for (int64_t n = 0; n < g->nnode; ++n) {
      if (!s->act[n >> 6]) continue;
      if (s->w[d][n] == 0.0) {
        const double fr = s->friction, usolid = 0.0;
        s->unew[d][n] = (1 - fr) * s->unew[d][n] + fr * usolid;
      }
    }
}

/* TwoDScene.cpp:6936-7090 mapNodeParticlesAPIC (fluid branch), then
 * SceneStepper.cpp:17-41 advectScene and TwoDScene.cpp:1776-1822
 * solidProjection (weights of the corner grid from the pre-advection
 * position, gradient from the advected one). */
void wco_g2p_advect(wco_scene* s, double dt) {
  const wco_grid* g = &s->g;
  const int64_t np = s->np;
  const double dx = g->dx;
  const double invD = 4.0 / (dx * dx); /* MathUtilities.h:728-737 */
  for (int64_t p = 0; p < np; ++p) {
    const double px[3] = {s->x[p], s->x[np + p], s->x[2 * np + p]};
    double fv[3] = {0, 0, 0};
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int d = 0; d < 3; ++d) {
      int base[3];
      double w[3][3];
      for (int a = 0; a < 3; ++a) {
        base[a] = stencil_base(g, a, px[a], face_off[d][a]);
        stencil_weights(g, px[a], base[a], face_off[d][a], w[a]);
      }
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
            if (!node_active(s, i, j, k)) continue;
            const double ww = w[0][c0] * w[1][c1] * w[2][c2];
            const double fnv = s->unew[d][wco_node_index(g, i, j, k)];
            fv[d] += fnv * ww;
            const double e0 = node_pos(g, i, face_off[d][0]) - px[0];
            const double e1 = node_pos(g, j, face_off[d][1]) - px[1];
            const double e2 = node_pos(g, k, face_off[d][2]) - px[2];
            B[d * 3 + 0] += fnv * ww * e0 * invD;
            B[d * 3 + 1] += fnv * ww * e1 * invD;
            B[d * 3 + 2] += fnv * ww * e2 * invD;
          }
    }
    for (int t = 0; t < 9; ++t) s->C[(int64_t)t * np + p] = WCO_PLANT == 9 ? B[t] : B[t] * s->flip;
    double nx[3];
    for (int d = 0; d < 3; ++d) {
      if (WCO_PLANT == 9) fv[d] *= s->flip;
      s->v[(int64_t)d * np + p] = fv[d];
      nx[d] = px[d] + fv[d] * dt;
    }
    if (s->solid_projection) {
      int base[3];
      double w[3][3];
      for (int a = 0; a < 3; ++a) {
        base[a] = stencil_base(g, a, px[a], 0.0);
        stencil_weights(g, px[a], base[a], 0.0, w[a]);
      }
      double phi_ori = 0.0, gr[3] = {0, 0, 0};
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
            const double phi = node_active(s, i, j, k)
                                   ? s->sphi[wco_node_index(g, i, j, k)]
                                   : 3.0 * dx;
            const double ww = w[0][c0] * w[1][c1] * w[2][c2];
            phi_ori += phi * ww;
            gr[0] += phi * invD * ww * (node_pos(g, i, 0.0) - nx[0]);
            gr[1] += phi * invD * ww * (node_pos(g, j, 0.0) - nx[1]);
            gr[2] += phi * invD * ww * (node_pos(g, k, 0.0) - nx[2]);
          }
      const double gn = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
      if (gn > 1e-20) {
        gr[0] /= gn;
        gr[1] /= gn;
        gr[2] /= gn;
      }
      const double dp[3] = {fv[0] * dt, fv[1] * dt, fv[2] * dt};
      const double phi_now = phi_ori + (gr[0] * dp[0] + gr[1] * dp[1] + gr[2] * dp[2]);
      int proj;
      if (s->frz_mode == 2) {
        proj = s->frz_proj[p];
      } else {
        proj = phi_now < 0.0;
        if (s->frz_mode == 1) s->frz_proj[p] = (uint8_t)proj;
      }
      if (proj) {
        nx[0] -= phi_now * gr[0];
        nx[1] -= phi_now * gr[1];
        nx[2] -= phi_now * gr[2];
      }
    }
    for (int d = 0; d < 3; ++d) s->x[(int64_t)d * np + p] = nx[d];
  }
}

void wco_copy_vel_to_unew(wco_scene* s) {
  for (int d = 0; d < 3; ++d)
    memcpy(s->unew[d], s->vel[d], sizeof(double) * s->g.nnode);
}

/* Core/WetClothCore.cpp:77-300 substep order for liquid-only scenes with
 * correctionStrength = 0 and no splitting/merging (parity settings). */
int wco_substep(wco_scene* s, double dt, double tol, int maxit) {
  wco_activate(s);               /* rebucketize + resampleNodes :115-117 */
  wco_update_liquid_phi(s);      /* :134 */
  wco_update_solid_phi(s);       /* :146 */
  wco_update_solid_weights(s);   /* :150 */
  wco_p2g(s);                    /* :162 */
  wco_step_velocity(s, dt);      /* :179 */
  wco_assemble(s, dt);           /* :191 projectFine -> solveNodePressure */
  const int ok = wco_solve_pressure(s, tol, maxit);
  wco_apply_pressure(s, dt);     /* :224, :230, :239 */
  wco_g2p_advect(s, dt);         /* :248, :258, :262 */
  return ok;
}

/* ------------------------------------------------------------- adjoint */

void wco_frozen_set(wco_scene* s, int mode) {
  const int64_t nn = s->g.nnode;
  if (!s->frz_act) {
    s->frz_act = (uint8_t*)xcalloc(s->g.nbucket, 1);
    s->frz_arg = (int32_t*)xcalloc(nn, sizeof(int32_t));
    s->frz_liquid = (uint8_t*)xcalloc(nn, 1);
    s->frz_dof = (int32_t*)xcalloc(nn, sizeof(int32_t));
    for (int d = 0; d < 3; ++d) {
      s->frz_clamp[d] = (uint8_t*)xcalloc(nn, 1);
      s->frz_mass[d] = (uint8_t*)xcalloc(nn, 1);
    }
    s->frz_proj = (uint8_t*)xcalloc(s->np, 1);
  }
  s->frz_mode = mode;
}

/* derivative of MathUtilities.h:416-424 quad_kernel */
static double quad_kernel_d(double x) {
  const double ax = fabs(x);
  if (ax < 0.5) return -2.0 * x;
  if (ax < 1.5) return x > 0 ? -(1.5 - ax) : (1.5 - ax);
  return 0.0;
}

/* the three per-axis weights of a stencil and their derivatives by x */
static void stencil_weights_d(const wco_grid* g, double x, int base, double off, double w[3],
                              double dw[3]) {
  for (int a = 0; a < 3; ++a) {
    const double r = (x - node_pos(g, base + a, off)) / g->dx;
    w[a] = wco_quad_kernel(r);
    dw[a] = quad_kernel_d(r) / g->dx;
  }
}

/* whether the pressure application (Pressure.cpp:1184-1211) sets face (d, n)
 * directly: a positive solid weight and a liquid cell on either side */
static int direct_face(const wco_scene* s, int d, int i, int j, int k) {
  const int64_t n = wco_node_index(&s->g, i, j, k);
  if (!s->act[n >> 6] || !(s->w[d][n] > 0.)) return 0;
  int l[3] = {i, j, k};
  l[d] -= 1;
  const int lex = node_active(s, l[0], l[1], l[2]);
  const int64_t ln = lex ? wco_node_index(&s->g, l[0], l[1], l[2]) : -1;
  return (lex && is_liquid(s, ln)) || is_liquid(s, n);
}

/* d theta / d phi of a face between cells nl (may be absent: < 0) and nr:
 * adds tb * dtheta/dphi to phib of each cell */
static void theta_adjoint(const wco_scene* s, int d, int64_t fn, int64_t nl, int64_t nr,
                          double tb, double* phib) {
  if (WCO_PLANT == 13) return;
  const int ll = nl >= 0 && is_liquid(s, nl), lr = is_liquid(s, nr);
  if (ll == lr) return;  /* theta = 1 (or no liquid): constant */
  if (s->frz_clamp[d][fn]) return; /* recorded (mode 1) or replayed (mode 2) */
  const double pl = nl >= 0 ? s->lphi[nl] : 0.0, pr = s->lphi[nr];
  const double a = ll ? pl : pr, b = ll ? pr : pl; /* liquid and air phi */
  const double den = (a - b) * (a - b);
  const double da = -b / den, db = a / den; /* theta = a / (a - b) */
  if (ll) {
    if (nl >= 0) phib[nl] += tb * da;
    phib[nr] += tb * db;
  } else {
    phib[nr] += tb * da;
    if (nl >= 0) phib[nl] += tb * db;
  }
}

/* Adjoint of wco_substep in the replayed (or recorded) decisions. */
int wco_substep_adjoint(wco_scene* s, double dt, double tol, int maxit, const double* x0,
                        const double* v0, const double* C0, double* xb, double* vb, double* Cb,
                        double* gb) {
  if (s->mix) return 0; /* the adjoint covers the liquid-only substep */
  const wco_grid* g = &s->g;
  const int64_t np = s->np, nn = g->nnode;
  const double dx = g->dx, invD = 4.0 / (dx * dx), fm = s->pmass;
  double* ub[3]; /* dL/d(final face velocity), then dL/d(P2G face velocity) */
  double* ub0[3];
  for (int d = 0; d < 3; ++d) {
    ub[d] = (double*)xcalloc(nn, sizeof(double));
    ub0[d] = (double*)xcalloc(nn, sizeof(double));
  }
  double* xin = (double*)xcalloc(3 * np, sizeof(double)); /* dL/dx0 */
  double* vin = (double*)xcalloc(3 * np, sizeof(double));
  double* Cin = (double*)xcalloc(9 * np, sizeof(double));

  /* 1. G2P (TDS:6936-7090), advection and solid projection, per particle */
  for (int64_t p = 0; p < np; ++p) {
    const double px[3] = {x0[p], x0[np + p], x0[2 * np + p]};
    const double fv[3] = {s->v[p], s->v[np + p], s->v[2 * np + p]};
    double nx[3];
    for (int d = 0; d < 3; ++d) nx[d] = px[d] + fv[d] * dt;
    double xbar_out[3] = {xb[p], xb[np + p], xb[2 * np + p]};
    double nxb[3] = {xbar_out[0], xbar_out[1], xbar_out[2]};
    double fvb[3] = {vb[p], vb[np + p], vb[2 * np + p]};
    double pxb[3] = {0, 0, 0};
    if (s->solid_projection && s->frz_proj[p] && WCO_PLANT != 14) {
      int base[3];
      double w[3][3], dw[3][3];
      for (int a = 0; a < 3; ++a) {
        base[a] = stencil_base(g, a, px[a], 0.0);
        stencil_weights_d(g, px[a], base[a], 0.0, w[a], dw[a]);
      }
      double phi_ori = 0.0, gr[3] = {0, 0, 0};
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
            const double phi = node_active(s, i, j, k) ? s->sphi[wco_node_index(g, i, j, k)]
                                                       : 3.0 * dx;
            const double ww = w[0][c0] * w[1][c1] * w[2][c2];
            phi_ori += phi * ww;
            gr[0] += phi * invD * ww * (node_pos(g, i, 0.0) - nx[0]);
            gr[1] += phi * invD * ww * (node_pos(g, j, 0.0) - nx[1]);
            gr[2] += phi * invD * ww * (node_pos(g, k, 0.0) - nx[2]);
          }
      const double gn = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
      double gh[3] = {gr[0], gr[1], gr[2]};
      if (gn > 1e-20)
        for (int a = 0; a < 3; ++a) gh[a] /= gn;
      const double dp[3] = {fv[0] * dt, fv[1] * dt, fv[2] * dt};
      const double phi_now = phi_ori + (gh[0] * dp[0] + gh[1] * dp[1] + gh[2] * dp[2]);
      /* x' = nx - phi_now gh */
      const double pnb = -(gh[0] * xbar_out[0] + gh[1] * xbar_out[1] + gh[2] * xbar_out[2]);
      double ghb[3], grb[3];
      for (int a = 0; a < 3; ++a) {
        ghb[a] = -phi_now * xbar_out[a] + pnb * dp[a];
        fvb[a] += dt * pnb * gh[a]; /* dp = fv dt */
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
            const double phi = node_active(s, i, j, k) ? s->sphi[wco_node_index(g, i, j, k)]
                                                       : 3.0 * dx;
            const double q[3] = {node_pos(g, i, 0.0), node_pos(g, j, 0.0), node_pos(g, k, 0.0)};
            const double coef =
                pnb + invD * ((q[0] - nx[0]) * grb[0] + (q[1] - nx[1]) * grb[1] +
                              (q[2] - nx[2]) * grb[2]);
            pxb[0] += phi * coef * dw[0][c0] * w[1][c1] * w[2][c2];
            pxb[1] += phi * coef * w[0][c0] * dw[1][c1] * w[2][c2];
            pxb[2] += phi * coef * w[0][c0] * w[1][c1] * dw[2][c2];
          }
    }
    /* nx = px + fv dt */
    for (int a = 0; a < 3; ++a) {
      pxb[a] += nxb[a];
      fvb[a] += dt * nxb[a];
    }
    /* fv_d = sum w u;  C_d = flip sum w u (x_n - x_p) invD */
    for (int d = 0; d < 3; ++d) {
      int base[3];
      double w[3][3], dw[3][3];
      for (int a = 0; a < 3; ++a) {
        base[a] = stencil_base(g, a, px[a], face_off[d][a]);
        stencil_weights_d(g, px[a], base[a], face_off[d][a], w[a], dw[a]);
      }
      const double Bb[3] = {s->flip * Cb[(int64_t)(d * 3 + 0) * np + p],
                            s->flip * Cb[(int64_t)(d * 3 + 1) * np + p],
                            s->flip * Cb[(int64_t)(d * 3 + 2) * np + p]};
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
            if (!node_active(s, i, j, k)) continue;
            const int64_t n = wco_node_index(g, i, j, k);
            const double u = s->unew[d][n];
            const double e[3] = {node_pos(g, i, face_off[d][0]) - px[0],
                                 node_pos(g, j, face_off[d][1]) - px[1],
                                 node_pos(g, k, face_off[d][2]) - px[2]};
            const double G = fvb[d] + invD * (Bb[0] * e[0] + Bb[1] * e[1] + Bb[2] * e[2]);
            const double ww = w[0][c0] * w[1][c1] * w[2][c2];
            ub[d][n] += ww * G;
            pxb[0] += u * G * dw[0][c0] * w[1][c1] * w[2][c2] - ww * u * invD * Bb[0];
            pxb[1] += u * G * w[0][c0] * dw[1][c1] * w[2][c2] - ww * u * invD * Bb[1];
            pxb[2] += u * G * w[0][c0] * w[1][c1] * dw[2][c2] - ww * u * invD * Bb[2];
          }
    }
    for (int a = 0; a < 3; ++a) xin[(int64_t)a * np + p] = pxb[a];
  }

  /* 2. constrainLiquidVelocity (TDS:1827-1866): faces with zero weight */
  for (int d = 0; d < 3; ++d)
    for (int64_t n = 0; n < nn; ++n)
      if (s->act[n >> 6] && s->w[d][n] == 0.0) ub[d][n] *= (1 - s->friction);

  /* 3. the two extrapolation layers (Pressure.cpp:1826-1903) in reverse:
   * layer L fills each invalid face with the mean of its neighbours valid
   * before the layer (vset[L]); its adjoint hands each filled face's
   * gradient to those neighbours in equal shares, and the face's value
   * before the layer, overwritten, gets none */
  {
    static const int off[6][3] = {{-1, 0, 0}, {1, 0, 0},  {0, -1, 0},
                                  {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};
    uint8_t* vset[3];
    uint8_t* fill[2];
    int32_t* cnt[2];
    for (int L = 0; L < 3; ++L) vset[L] = (uint8_t*)xcalloc(nn, 1);
    for (int L = 0; L < 2; ++L) {
      fill[L] = (uint8_t*)xcalloc(nn, 1);
      cnt[L] = (int32_t*)xcalloc(nn, sizeof(int32_t));
    }
    for (int d = 0; d < 3; ++d) {
      for (int k = 0; k < g->nx[2]; ++k)
        for (int j = 0; j < g->nx[1]; ++j)
          for (int i = 0; i < g->nx[0]; ++i)
            vset[0][wco_node_index(g, i, j, k)] = (uint8_t)direct_face(s, d, i, j, k);
      for (int L = 0; L < 2; ++L) {
        memcpy(vset[L + 1], vset[L], nn);
        for (int k = 0; k < g->nx[2]; ++k)
          for (int j = 0; j < g->nx[1]; ++j)
            for (int i = 0; i < g->nx[0]; ++i) {
              const int64_t n = wco_node_index(g, i, j, k);
              fill[L][n] = 0;
              if (!s->act[n >> 6] || vset[L][n]) continue;
              int c = 0;
              for (int r = 0; r < 6; ++r) {
                const int a = i + off[r][0], b = j + off[r][1], q = k + off[r][2];
                if (node_active(s, a, b, q) && vset[L][wco_node_index(g, a, b, q)]) ++c;
              }
              if (c > 0) {
                fill[L][n] = 1;
                cnt[L][n] = c;
                vset[L + 1][n] = 1;
              }
            }
      }
      for (int L = 1; L >= 0; --L)
        for (int k = 0; k < g->nx[2]; ++k)
          for (int j = 0; j < g->nx[1]; ++j)
            for (int i = 0; i < g->nx[0]; ++i) {
              const int64_t n = wco_node_index(g, i, j, k);
              if (!fill[L][n]) continue;
              const double share = ub[d][n] / (double)cnt[L][n];
              for (int r = 0; r < 6; ++r) {
                const int a = i + off[r][0], b = j + off[r][1], q = k + off[r][2];
                if (!node_active(s, a, b, q)) continue;
                const int64_t m = wco_node_index(g, a, b, q);
                if (vset[L][m]) ub[d][m] += share;
              }
              ub[d][n] = 0.0;
            }
    }
    for (int L = 0; L < 3; ++L) free(vset[L]);
    for (int L = 0; L < 2; ++L) {
      free(fill[L]);
      free(cnt[L]);
    }
  }

  /* 4. the pressure application (Pressure.cpp:1137-1314, LIE:65-104):
   * u0 = invC rhs - invC dt V (p_R - p_L) / (dx theta), rhs = M (u + g dt) */
  const int64_t ndof = s->ndof;
  double* pb = (double*)xcalloc(ndof ? ndof : 1, sizeof(double));
  double* phib = (double*)xcalloc(nn, sizeof(double));
  double* rb[3]; /* dL/d rhs of the grid update per face */
  for (int d = 0; d < 3; ++d) rb[d] = (double*)xcalloc(nn, sizeof(double));
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < g->nx[2]; ++k)
      for (int j = 0; j < g->nx[1]; ++j)
        for (int i = 0; i < g->nx[0]; ++i) {
          const int64_t n = wco_node_index(g, i, j, k);
          if (!s->act[n >> 6]) continue;
          const double ubn = ub[d][n];
          const double invC = s->invC[d][n];
          rb[d][n] += invC * ubn;
          if (!direct_face(s, d, i, j, k)) continue;
          int l[3] = {i, j, k};
          l[d] -= 1;
          const int lex = node_active(s, l[0], l[1], l[2]);
          const int64_t ln = lex ? wco_node_index(g, l[0], l[1], l[2]) : -1;
          const int lliq = lex && is_liquid(s, ln), rliq = is_liquid(s, n);
          double theta = 1.0;
          if (!lliq || !rliq) theta = face_theta(s, d, n, ln, n, lex ? s->lphi[ln] : 0.0, s->lphi[n]);
          const double c = invC * dt * s->vol[d][n] / (dx * theta);
          /* u0 -= c (p_n - p_ln) */
          if (s->dof[n] >= 0) pb[s->dof[n]] -= c * ubn;
          if (lex && s->dof[ln] >= 0) pb[s->dof[ln]] += c * ubn;
          if (!lliq || !rliq) {
            const double pn = s->pressure[n], pl = lex ? s->pressure[ln] : 0.0;
            const double tb = ubn * invC * dt * s->vol[d][n] * (pn - pl) / (dx * theta * theta);
            theta_adjoint(s, d, n, ln, n, tb, phib);
          }
        }

  /* 5. the pressure solve: bbar = A^-1 pbar (A symmetric); A_ii holds
   * term / theta for each air face of a DOF */
  double* bb = (double*)xcalloc(ndof ? ndof : 1, sizeof(double));
  double res = 0;
  int it = 0;
  int ok = 1;
  if (ndof > 0)
    ok = wco_pcg_solve(ndof, s->rowstart, s->colidx, s->val, pb, bb, tol, maxit, 0.97, 0.25,
                       &res, &it);
  const double coeff = dt / (dx * dx);
  for (int64_t bucket = 0; bucket < g->nbucket; ++bucket) {
    if (!s->act[bucket]) continue;
    const int bk = (int)(bucket / ((int64_t)g->nb[0] * g->nb[1]));
    const int bj = (int)((bucket / g->nb[0]) % g->nb[1]);
    const int bi = (int)(bucket % g->nb[0]);
    for (int l = 0; l < 64; ++l) {
      const int64_t n = (bucket << 6) | l;
      const int32_t me = s->dof[n];
      if (me < 0) continue;
      const int i = bi * 4 + (l & 3), j = bj * 4 + ((l >> 2) & 3), k = bk * 4 + (l >> 4);
      for (int f = 0; f < 6; ++f) {
        const int axis = f >> 1, hi = f & 1;
        int q[3] = {i, j, k};
        q[axis] += hi;
        if (!node_active(s, q[0], q[1], q[2])) continue;
        const int64_t fn = wco_node_index(g, q[0], q[1], q[2]);
        const double w = s->w[axis][fn];
        /* b: rhs += (lo face) or -= (hi face) u* psi w / dx, u* psi = invC rhs */
        const double sgn = hi ? -1.0 : 1.0;
        rb[axis][fn] += sgn * bb[me] * s->invC[axis][fn] * w / dx;
        if (!(w > 0.0)) continue;
        int nb[3] = {i, j, k};
        nb[axis] += hi ? 1 : -1;
        if (!node_active(s, nb[0], nb[1], nb[2])) continue;
        const int64_t nnb = wco_node_index(g, nb[0], nb[1], nb[2]);
        if (is_liquid(s, nnb)) continue;
        const double term = w * coeff * (s->invC[axis][fn] * s->vol[axis][fn]);
        const int64_t nl = hi ? n : nnb, nr = hi ? nnb : n;
        const double theta = face_theta(s, axis, fn, nl, nr, s->lphi[nl], s->lphi[nr]);
        /* A_ii += term / theta: dL = -bbar_i dA_ii p_i */
        const double tb = bb[me] * s->p[me] * term / (theta * theta);
        theta_adjoint(s, axis, fn, nl, nr, tb, phib);
      }
    }
  }

  /* 6. the grid update (LIE:224-337): rhs = M (u + g dt) */
  double gbar_y = 0.0;
  for (int d = 0; d < 3; ++d)
    for (int64_t n = 0; n < nn; ++n) {
      if (!s->act[n >> 6] || !has_mass(s, d, n)) {
        ub0[d][n] = 0.0;
        continue;
      }
      const double M = s->mass[d][n];
      ub0[d][n] = M * rb[d][n];
      if (d == 1) gbar_y += M * dt * rb[d][n];
    }
  *gb += -gbar_y; /* gravity (0, -|g|, 0) */

  /* 7. P2G (TDS:6700-6713): u = P / M, P = fm sum w (v + C e), M = fm sum w */
  for (int64_t p = 0; p < np; ++p) {
    const double px[3] = {x0[p], x0[np + p], x0[2 * np + p]};
    double pxb[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
      int base[3];
      double w[3][3], dw[3][3];
      for (int a = 0; a < 3; ++a) {
        base[a] = stencil_base(g, a, px[a], face_off[d][a]);
        stencil_weights_d(g, px[a], base[a], face_off[d][a], w[a], dw[a]);
      }
      const double vd = v0[(int64_t)d * np + p];
      const double b0 = C0[(int64_t)(d * 3 + 0) * np + p];
      const double b1 = C0[(int64_t)(d * 3 + 1) * np + p];
      const double b2 = C0[(int64_t)(d * 3 + 2) * np + p];
      double vdb = 0, cb[3] = {0, 0, 0};
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
            if (!node_active(s, i, j, k)) continue;
            const int64_t n = wco_node_index(g, i, j, k);
            if (!has_mass(s, d, n)) continue;
            const double M = s->mass[d][n];
            const double Pb = ub0[d][n] / M;
            const double Mb = -ub0[d][n] * s->vel[d][n] / M;
            const double e[3] = {node_pos(g, i, face_off[d][0]) - px[0],
                                 node_pos(g, j, face_off[d][1]) - px[1],
                                 node_pos(g, k, face_off[d][2]) - px[2]};
            const double vel = vd + (b0 * e[0] + b1 * e[1] + b2 * e[2]);
            const double ww = w[0][c0] * w[1][c1] * w[2][c2];
            vdb += fm * ww * Pb;
            for (int a = 0; a < 3; ++a) cb[a] += fm * ww * Pb * e[a];
            const double G = WCO_PLANT == 16 ? 0.0 : fm * (Pb * vel + Mb);
            pxb[0] += G * dw[0][c0] * w[1][c1] * w[2][c2] - fm * ww * Pb * b0;
            pxb[1] += G * w[0][c0] * dw[1][c1] * w[2][c2] - fm * ww * Pb * b1;
            pxb[2] += G * w[0][c0] * w[1][c1] * dw[2][c2] - fm * ww * Pb * b2;
          }
      vin[(int64_t)d * np + p] += vdb;
      for (int a = 0; a < 3; ++a) Cin[(int64_t)(d * 3 + a) * np + p] += cb[a];
    }
    for (int a = 0; a < 3; ++a) xin[(int64_t)a * np + p] += pxb[a];
  }

  /* 8. the liquid phi (TDS:3307-3348): phi_c = |x_p - x_c| - r0 for the
   * cell's recorded particle */
  for (int k = 0; k < g->nx[2]; ++k)
    for (int j = 0; j < g->nx[1]; ++j)
      for (int i = 0; i < g->nx[0]; ++i) {
        const int64_t n = wco_node_index(g, i, j, k);
        const int32_t a = s->lphi_arg[n];
        if (a < 0 || phib[n] == 0.0 || WCO_PLANT == 15) continue;
        const double d0 = x0[a] - node_pos(g, i, 0.5);
        const double d1 = x0[np + a] - node_pos(g, j, 0.5);
        const double d2 = x0[2 * np + a] - node_pos(g, k, 0.5);
        const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        if (r > 0) {
          xin[a] += phib[n] * d0 / r;
          xin[np + a] += phib[n] * d1 / r;
          xin[2 * np + a] += phib[n] * d2 / r;
        }
      }

  memcpy(xb, xin, sizeof(double) * 3 * np);
  memcpy(vb, vin, sizeof(double) * 3 * np);
  memcpy(Cb, Cin, sizeof(double) * 9 * np);
  for (int d = 0; d < 3; ++d) {
    free(ub[d]);
    free(ub0[d]);
    free(rb[d]);
  }
  free(xin);
  free(vin);
  free(Cin);
  free(pb);
  free(bb);
  free(phib);
  return ok;
}

/* ---------------------------------------- mixture: static porous solid */

struct wco_mix {
  wco_mix_params prm;
  int64_t ne;
  double *x, *m, *fm, *fvol, *rvol, *rvf, *ori, *sf; /* x, ori: 3 * ne SoA */
  double *smass[3], *svol[3], *psi[3], *sat[3], *sfac[3], *ori_n[3]; /* ori_n: 3 * nnode */
  double *Q[3], *P[3], *hdv[3], *psi_sf[3];
  double *psi_p, *sat_p, *pp; /* cell centred */
};

static void mix_free(struct wco_mix* m) {
  if (!m) return;
  free(m->x);
  free(m->m);
  free(m->fm);
  free(m->fvol);
  free(m->rvol);
  free(m->rvf);
  free(m->ori);
  free(m->sf);
  for (int d = 0; d < 3; ++d) {
    free(m->smass[d]);
    free(m->svol[d]);
    free(m->psi[d]);
    free(m->sat[d]);
    free(m->sfac[d]);
    free(m->ori_n[d]);
    free(m->Q[d]);
    free(m->P[d]);
    free(m->hdv[d]);
    free(m->psi_sf[d]);
  }
  free(m->psi_p);
  free(m->sat_p);
  free(m->pp);
  free(m);
}

void wco_set_elasto(wco_scene* s, int64_t ne, const double* x, const double* m,
                    const double* fm, const double* fvol, const double* rest_vol,
                    const double* rest_vf, const double* ori, const double* sf,
                    const wco_mix_params* prm) {
  mix_free(s->mix);
  s->mix = NULL;
  if (ne <= 0) return;
  struct wco_mix* M = (struct wco_mix*)xcalloc(1, sizeof(struct wco_mix));
  M->prm = *prm;
  M->ne = ne;
  M->x = (double*)xcalloc(3 * ne, sizeof(double));
  M->ori = (double*)xcalloc(3 * ne, sizeof(double));
  memcpy(M->x, x, sizeof(double) * 3 * ne);
  memcpy(M->ori, ori, sizeof(double) * 3 * ne);
  double** per[6] = {&M->m, &M->fm, &M->fvol, &M->rvol, &M->rvf, &M->sf};
  const double* src[6] = {m, fm, fvol, rest_vol, rest_vf, sf};
  for (int t = 0; t < 6; ++t) {
    *per[t] = (double*)xcalloc(ne, sizeof(double));
    memcpy(*per[t], src[t], sizeof(double) * ne);
  }
  const int64_t nn = s->g.nnode;
  for (int d = 0; d < 3; ++d) {
    M->smass[d] = (double*)xcalloc(nn, sizeof(double));
    M->svol[d] = (double*)xcalloc(nn, sizeof(double));
    M->psi[d] = (double*)xcalloc(nn, sizeof(double));
    M->sat[d] = (double*)xcalloc(nn, sizeof(double));
    M->sfac[d] = (double*)xcalloc(nn, sizeof(double));
    M->ori_n[d] = (double*)xcalloc(3 * nn, sizeof(double));
    M->Q[d] = (double*)xcalloc(nn, sizeof(double));
    M->P[d] = (double*)xcalloc(nn, sizeof(double));
    M->hdv[d] = (double*)xcalloc(nn, sizeof(double));
    M->psi_sf[d] = (double*)xcalloc(nn, sizeof(double));
  }
  M->psi_p = (double*)xcalloc(nn, sizeof(double));
  M->sat_p = (double*)xcalloc(nn, sizeof(double));
  M->pp = (double*)xcalloc(nn, sizeof(double));
  s->mix = M;
}

double* wco_mix_field(wco_scene* s, const char* name, int axis) {
  struct wco_mix* M = s->mix;
  if (!M || axis < 0) return NULL;
  if (!strcmp(name, "ori")) return axis < 9 ? M->ori_n[axis / 3] + (int64_t)(axis % 3) * s->g.nnode : NULL;
  if (!strcmp(name, "psi_p")) return M->psi_p;
  if (!strcmp(name, "sat_p")) return M->sat_p;
  if (!strcmp(name, "pp")) return M->pp;
  if (axis > 2) return NULL;
  if (!strcmp(name, "smass")) return M->smass[axis];
  if (!strcmp(name, "svol")) return M->svol[axis];
  if (!strcmp(name, "psi")) return M->psi[axis];
  if (!strcmp(name, "sat")) return M->sat[axis];
  if (!strcmp(name, "sf")) return M->sfac[axis];
  if (!strcmp(name, "Q")) return M->Q[axis];
  if (!strcmp(name, "P")) return M->P[axis];
  if (!strcmp(name, "hdv")) return M->hdv[axis];
  if (!strcmp(name, "psi_sf")) return M->psi_sf[axis];
  return NULL;
}

static double clamp01(double a) { return a < 0.0 ? 0.0 : (a > 1.0 ? 1.0 : a); }

/* The orientation terms: get_rotated_drag_x/y/z and get_rotated_pore_coeff_x/y/z
 * (MathUtilities.h:792-878) are the component sums of one row of a tensor
 * that is transversely isotropic about n = a / |a|, T = t I + (u - t) n n^T:
 * row k sums to t + (u - t) n_k (n_x + n_y + n_z).  The drag tensor has
 * t = d(0) = d(1) (c_xy: every caller passes c_xy twice) and u = d(2)
 * (c_z); the pore tensor has t = 2 - sf and u = 1 + sf.  Both are zero for
 * a = 0 (the reference divides a zero numerator by max(1e-63, |a|^2)). */
static double iso_row_sum(const double a[3], double t, double u, int k) {
  const double s2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  if (s2 == 0.0) return 0.0;
  const double r = sqrt(s2);
  const double n[3] = {a[0] / r, a[1] / r, a[2] / r};
  return t + (u - t) * n[k] * (n[0] + n[1] + n[2]);
}

double wco_rotated_drag(const double a[3], const double d[3], int index) {
  return iso_row_sum(a, d[0], d[2], index);
}

double wco_rotated_pore_coeff(const double a[3], double sf, int index) {
  return iso_row_sum(a, 2.0 - sf, 1.0 + sf, index);
}

/* TwoDScene.cpp:563-571 */
double wco_capillary_pressure(const wco_mix_params* prm, double psi) {
  if (1.0 - psi < 1e-20 || prm->pore_radius < 1e-20) return 0.0;
  const double alpha = psi / (1. - psi);
  const double contact_angle = cos(prm->rest_contact_angle);
  return alpha * prm->surf_tension_coeff * contact_angle / prm->pore_radius;
}

/* TwoDScene.cpp:4042-4097, material 0 (the liquid: its viscosity and
 * density): the permeabilities along and across the yarns ka, kb of volume
 * fraction psi and yarn diameter, the coefficients mu / k plus the
 * nonlinear (Ergun) term, mixed by the shape factor and rotated by the
 * orientation. */
double wco_drag_coeff(const wco_mix_params* prm, double rho, double psi, double sat, double dv,
                      const double ori[3], double sf, int index) {
  const double o2 = ori[0] * ori[0] + ori[1] * ori[1] + ori[2] * ori[2];
  if (!prm->use_drag || psi == 0.0 || sat == 0.0 || o2 < 1e-20) return 0.0;
  const double ergun_coeff = prm->use_nonlinear_drag ? 0.1428869017 : 0.0;
  const double di = prm->yarn_diameter;
  double ka = (-log(psi) - 1.476 + 2.0 * psi - 0.5 * psi * psi) / (16.0 * psi) * di * di;
  if (ka < 1e-20) ka = 1e-20;
  double kb = (-log(psi) - 1.476 + 2.0 * psi - 1.774 * psi * psi + 4.078 * pow(psi, 3.0)) /
              (32.0 * psi) * di * di;
  if (kb < 1e-20) kb = 1e-20;
  const double mu = prm->viscosity;
  const double pw = prm->yazdchi_power;
  double ca = mu / ka + ergun_coeff * pow(di, pw - 1.0) * pow(mu, 1.0 - pw) /
                            (pow(1.0 - psi, 1.5) * sqrt(ka)) * pow(rho * fabs(dv), pw);
  if (ca > 1e+63) ca = 1e+63;
  double cb = mu / kb + ergun_coeff * pow(di, pw - 1.0) * pow(mu, 1.0 - pw) /
                            (pow(1.0 - psi, 1.5) * sqrt(kb)) * pow(rho * fabs(dv), pw);
  if (cb > 1e+63) cb = 1e+63;
  const double c_xy = ca * (1.0 - sf) + cb * sf;
  const double c_z = ca * sf + cb * (1.0 - sf);
  const double cv[3] = {c_xy, c_xy, c_z};
  return wco_rotated_drag(ori, cv, index);
}

/* the five stencils of every solid particle touch their buckets (findNodes
 * runs over all particles, TwoDScene.cpp:3683-3752) */
static void mix_touch(const wco_scene* s, uint8_t* touched) {
  const struct wco_mix* M = s->mix;
  const wco_grid* g = &s->g;
  static const double offs[5][3] = {
      {0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}, {0.0, 0.0, 0.0}, {0.5, 0.5, 0.5}};
  for (int64_t p = 0; p < M->ne; ++p) {
    const double px[3] = {M->x[p], M->x[M->ne + p], M->x[2 * M->ne + p]};
    for (int gi = 0; gi < 5; ++gi) {
      int blo[3], bhi[3];
      for (int d = 0; d < 3; ++d) {
        const int b = stencil_base(g, d, px[d], offs[gi][d]);
        blo[d] = b >> 2;
        bhi[d] = (b + 2) >> 2;
      }
      for (int bk = blo[2]; bk <= bhi[2]; ++bk)
        for (int bj = blo[1]; bj <= bhi[1]; ++bj)
          for (int bi = blo[0]; bi <= bhi[0]; ++bi) touched[wco_bucket_index(g, bi, bj, bk)] = 1;
    }
  }
}

/* The solid's branch of mapParticleNodesAPIC (TwoDScene.cpp:6686-6699,
 * 6715-6736: the orientation sum normalised where its norm exceeds 1e-20) and mapParticleSaturationPsiNodes (:5952-5992) with
 * updatePorePressureNodes (:5933-5947). */
static void mix_p2g(wco_scene* s) {
  struct wco_mix* M = s->mix;
  const wco_grid* g = &s->g;
  const int64_t nn = g->nnode, ne = M->ne;
  const double dx = g->dx, dV = dx * dx * dx;
  double* vsol = (double*)xcalloc(nn, sizeof(double));
  double* vfe = (double*)xcalloc(nn, sizeof(double));
  double* rw = (double*)xcalloc(nn, sizeof(double));
  for (int d = 0; d < 3; ++d) {
    memset(M->smass[d], 0, sizeof(double) * nn);
    memset(M->sfac[d], 0, sizeof(double) * nn);
    memset(M->ori_n[d], 0, sizeof(double) * 3 * nn);
    memset(vsol, 0, sizeof(double) * nn);
    memset(vfe, 0, sizeof(double) * nn);
    memset(rw, 0, sizeof(double) * nn);
    for (int64_t p = 0; p < ne; ++p) {
      const double px[3] = {M->x[p], M->x[ne + p], M->x[2 * ne + p]};
      int base[3];
      double w[3][3];
      for (int a = 0; a < 3; ++a) {
        const double off = WCO_PLANT == 24 ? 0.5 : face_off[d][a];
        base[a] = stencil_base(g, a, px[a], off);
        stencil_weights(g, px[a], base[a], off, w[a]);
      }
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
            if (!node_active(s, i, j, k)) continue;
            const int64_t n = wco_node_index(g, i, j, k);
            const double ww = w[0][c0] * w[1][c1] * w[2][c2];
            M->smass[d][n] += (M->m[p] + M->fm[p]) * ww;
            vsol[n] += M->rvol[p] * M->rvf[p] * ww;
            vfe[n] += M->fvol[p] * ww;
            M->sfac[d][n] += M->sf[p] * ww;
            rw[n] += ww;
            for (int c = 0; c < 3; ++c) M->ori_n[d][c * nn + n] += M->ori[c * ne + p] * ww;
          }
    }
    for (int64_t n = 0; n < nn; ++n) {
      if (rw[n] > 1e-20) M->sfac[d][n] /= rw[n];
      double* o = M->ori_n[d];
      const double lo = sqrt(o[n] * o[n] + o[nn + n] * o[nn + n] + o[2 * nn + n] * o[2 * nn + n]);
      if (lo > 1e-20)
        for (int c = 0; c < 3; ++c) o[c * nn + n] /= lo;
      M->svol[d][n] = vsol[n] + vfe[n];
      M->psi[d][n] = clamp01(vsol[n] / dV);
      const double room = dV - vsol[n] > 1e-20 ? dV - vsol[n] : 1e-20;
      M->sat[d][n] = clamp01((s->vol[d][n] + vfe[n]) / room);
    }
  }
  /* cell centres: the liquid and solid volume of every particle's cell
   * stencil (liquid particles carry no solid volume), then the pore pressure */
  memset(vsol, 0, sizeof(double) * nn);
  memset(vfe, 0, sizeof(double) * nn);
  for (int pass = 0; pass < 2; ++pass) {
    const int64_t cnt = pass == 0 ? ne : s->np;
    for (int64_t p = 0; p < cnt; ++p) {
      const double* X = pass == 0 ? M->x : s->x;
      const double px[3] = {X[p], X[cnt + p], X[2 * cnt + p]};
      const double fv = pass == 0 ? M->fvol[p] : s->pvol;
      const double sv = pass == 0 ? M->rvol[p] * M->rvf[p] : 0.0;
      int base[3];
      double w[3][3];
      for (int a = 0; a < 3; ++a) {
        base[a] = stencil_base(g, a, px[a], 0.5);
        stencil_weights(g, px[a], base[a], 0.5, w[a]);
      }
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
            if (!node_active(s, i, j, k)) continue;
            const int64_t n = wco_node_index(g, i, j, k);
            const double ww = w[0][c0] * w[1][c1] * w[2][c2];
            vfe[n] += fv * ww;
            vsol[n] += sv * ww;
          }
    }
  }
  for (int64_t n = 0; n < nn; ++n) {
    M->psi_p[n] = clamp01(vsol[n] / dV);
    const double room = dV - vsol[n] > 1e-20 ? dV - vsol[n] : 1e-20;
    M->sat_p[n] = clamp01(vfe[n] / room);
    M->pp[n] = wco_capillary_pressure(&M->prm, M->psi_p[n]) * (1.0 - M->sat_p[n]);
  }
  free(vsol);
  free(vfe);
  free(rw);
}

static void mix_face(const wco_scene* s, int d, int64_t n, double* psi, double* Vs, double* Q,
                     double* P, double* psf, double* Ms) {
  const struct wco_mix* M = s->mix;
  *psi = M->psi[d][n];
  *Vs = M->svol[d][n];
  *Q = M->Q[d][n];
  *P = M->P[d][n];
  *psf = M->psi_sf[d][n];
  *Ms = M->smass[d][n];
}

/* stepVelocity with the solid not solved (LinearizedImplicitEuler.cpp:
 * 224-337): constructNodeForce (:548-659: gravity, then the pore pressure
 * gradient of Pressure.cpp:26-142, then M u), constructHDV (:846-978, no
 * drag by air: the damped mass is the solid mass), constructInvMDV
 * (:1109-1148), performLocalSolve (:128-175) and constructPsiSF
 * (:1052-1107; invCs stays zero: constructMsDVs runs only when the solid is
 * solved). */
static void mix_step_velocity(wco_scene* s, double dt) {
  struct wco_mix* M = s->mix;
  const wco_grid* g = &s->g;
  const int64_t nn = g->nnode;
  const double dx = g->dx;
  for (int d = 0; d < 3; ++d) {
    const double gd = s->gravity[d];
    for (int k = 0; k < g->nx[2]; ++k)
      for (int j = 0; j < g->nx[1]; ++j)
        for (int i = 0; i < g->nx[0]; ++i) {
          const int64_t n = wco_node_index(g, i, j, k);
          if (!s->act[n >> 6]) {
            s->rhs[d][n] = s->invC[d][n] = s->ustar[d][n] = s->psi_fs[d][n] = 0.0;
            M->Q[d][n] = M->P[d][n] = M->hdv[d][n] = M->psi_sf[d][n] = 0.0;
            continue;
          }
          const double Mf = s->mass[d][n];
          double rhs = 0.0;
          rhs -= Mf * gd * (-dt);
          /* computePorePressureGrads: (p_right - p_left) / dx V_f dt sat w
           * times the rotated pore coefficient, on faces with liquid volume
           * and solid weight */
          const double vf = s->vol[d][n], wv = s->w[d][n];
          if (vf > 0.0 && wv > 0.0) {
            int l[3] = {i, j, k};
            l[d] -= 1;
            const double pp0 = node_active(s, l[0], l[1], l[2])
                                   ? M->pp[wco_node_index(g, l[0], l[1], l[2])]
                                   : 0.0;
            const double pp1 = M->pp[n];
            const double grad = (WCO_PLANT == 22 ? pp0 - pp1 : pp1 - pp0) / dx;
            const double a[3] = {M->ori_n[d][n], M->ori_n[d][nn + n], M->ori_n[d][2 * nn + n]};
            rhs += grad * vf * dt * M->sat[d][n] * wv *
                   wco_rotated_pore_coeff(a, M->sfac[d][n], d);
          }
          rhs += Mf * s->vel[d][n];
          s->rhs[d][n] = rhs;
          /* drag: the solid is at rest, so dv is the liquid's P2G velocity */
          const double a[3] = {M->ori_n[d][n], M->ori_n[d][nn + n], M->ori_n[d][2 * nn + n]};
          const double dc = wco_drag_coeff(&M->prm, s->rho, M->psi[d][n], M->sat[d][n],
                                           s->vel[d][n] - 0.0, a, M->sfac[d][n], d);
          const double Vm = M->svol[d][n] + vf;
          const double hdV = dt * dc * Vm;
          const double damped = M->smass[d][n];
          const double mfhdvm = Mf + hdV, mshdvm = damped + hdV;
          const double P = mfhdvm > 1e-20 ? hdV / mfhdvm : 1.0;
          const double Q = WCO_PLANT == 23 ? P : (mshdvm > 1e-20 ? hdV / mshdvm : 1.0);
          M->hdv[d][n] = hdV;
          M->P[d][n] = P;
          M->Q[d][n] = Q;
          const double Cc = Mf + Q * damped;
          const double invC = Cc > 1e-20 ? 1.0 / Cc : 1.0;
          s->invC[d][n] = invC;
          s->ustar[d][n] = Mf > 1e-20 ? rhs / Mf : rhs;
          const double psi = M->psi[d][n], invCs = 0.0;
          s->psi_fs[d][n] = ((1.0 - psi) * invC + psi * invCs * P) * Mf;
          M->psi_sf[d][n] = (psi * invCs + (1.0 - psi) * invC * Q) * damped;
        }
  }
}
