// Particle-side kernels for gfx950: seeding, the bucket sort, P2G and G2P.
//
// Particles live in HBM as structure-of-arrays planes and are kept sorted by
// the 4x4x4 bucket of their cell (the reference re-buckets every substep,
// TwoDScene.cpp:1657-1708).  P2G runs one workgroup per bucket: the bucket's
// particles scatter their 3x3x3 quadratic B-spline stencils into an LDS tile
// of 8x8x8 nodes (the bucket plus a two-node halo) for the u, v and w grids,
// and the tile is then added to HBM with one atomic per touched node.  The
// same pass takes the stencil-limited minimum of the liquid distance
// (updateLiquidPhi) and marks the buckets the stencils touch (resampleNodes'
// activation).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <type_traits>

#include <vector>

#include "wc_ctx.h"

namespace wc {
namespace {

constexpr int kP2GThreads = 256;  // the particle permutation in k_p2g assumes 256
constexpr int kTile = 8;  // LDS nodes per axis: bucket (4) + halo 2 below, 2 above
constexpr int kTileNodes = kTile * kTile * kTile;
// P2G's LDS node layout: k planes 68 doubles apart.  A double spans two of
// the 64 banks; with i, j strides of 1 and 8 doubles and this k stride, the
// nodes of a bucket's 4x4x4 cells shift by 2 i + 16 j + 8 k banks, so a
// wave's 64 lanes meet at most 2-way conflicts (a stride of 64 aliases every
// k, 72 aliases k with j)
constexpr int kPlane = 68;
constexpr int kLdsNodes = kTile * kPlane;
__device__ inline int p2g_lds(int li, int lj, int lk) { return lk * kPlane + lj * kTile + li; }
constexpr unsigned long long kPhiUntouched = 0xFFFFFFFFFFFFFFFFull;

__device__ inline int clampi(int a, int lo, int hi) {
  return a < lo ? lo : (a > hi ? hi : a);
}

template <class R>
__device__ inline int particle_bucket_axis(const Grid& g, int axis, R x) {
  const int c = (int)floor(x / (R)g.dx) + kGhost;
  return clampi(c, 0, g.nx[axis] - 1) >> 2;
}

// Runs of equal keys in consecutive lanes of a wave (the particles stay
// nearly sorted from one substep to the next, so a wave's particles mostly
// share a few buckets).  Each run's first lane claims space for the whole
// run with one atomic; head is the run's first lane, len its length (valid in
// the head lane), rank the lane's position in its run.
struct WaveRun {
  int head, len, rank;
};
__device__ inline WaveRun wave_run(uint32_t key) {
  const int lane = (int)__lane_id();
  const uint32_t prev = __shfl_up(key, 1u);
  const bool is_head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(is_head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  WaveRun r;
  r.head = 63 - __clzll(heads & upto);
  const unsigned long long after = heads & ~upto;
  const int next = after ? __ffsll(after) - 1 : 64;
  r.len = next - lane;
  r.rank = lane - r.head;
  return r;
}

constexpr uint32_t kNoKey = 0xFFFFFFFFu;  // lanes past the last particle

template <class R>
__global__ __launch_bounds__(256) void k_sort_keys(Grid g, const R* x, int64_t n, int64_t cap,
                                                   uint32_t* key, uint32_t* count) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  uint32_t b = kNoKey;
  if (p < n) {
    const int bi = particle_bucket_axis<R>(g, 0, x[p]);
    const int bj = particle_bucket_axis<R>(g, 1, x[cap + p]);
    const int bk = particle_bucket_axis<R>(g, 2, x[2 * cap + p]);
    b = (uint32_t)bucket_index(g, bi, bj, bk);
    key[p] = b;
  }
  const WaveRun r = wave_run(b);
  if (b != kNoKey && r.rank == 0) atomicAdd(&count[b], (uint32_t)r.len);
}

template <class R>
__global__ __launch_bounds__(256) void k_sort_scatter(int64_t n, int64_t cap, const uint32_t* key,
                                                      const uint32_t* start, uint32_t* cursor,
                                                      const R* x, const R* v, const R* C,
                                                      const uint32_t* pid, R* x2, R* v2, R* C2,
                                                      uint32_t* pid2) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const uint32_t b = p < n ? key[p] : kNoKey;
  // the particle's state is read before the run takes its space, so the
  // loads are in flight during the atomic's round trip
  R xs[3], vs[3], cs[9];
  uint32_t id = 0, sb = 0;
  if (b != kNoKey) {
    sb = start[b];
    for (int d = 0; d < 3; ++d) {
      xs[d] = x[d * cap + p];
      vs[d] = v[d * cap + p];
    }
    for (int t = 0; t < 9; ++t) cs[t] = C[t * cap + p];
    id = pid[p];
  }
  const WaveRun r = wave_run(b);
  uint32_t base = 0;
  if (b != kNoKey && r.rank == 0) base = atomicAdd(&cursor[b], (uint32_t)r.len);
  base = __shfl(base, r.head);
  if (b == kNoKey) return;
  const int64_t q = (int64_t)sb + base + r.rank;
  for (int d = 0; d < 3; ++d) {
    x2[d * cap + q] = xs[d];
    v2[d * cap + q] = vs[d];
  }
  for (int t = 0; t < 9; ++t) C2[t * cap + q] = cs[t];
  pid2[q] = id;
}

template <class R>
struct P2GArgs {
  Grid g;
  const R* x;
  const R* v;
  const R* C;
  int64_t cap;
  const uint32_t* bstart;
  double* wsum[3];  // sum of the weights (the node mass and volume are fm and fvol times it)
  double* mom[3];
  unsigned long long* phikey;
  uint8_t* touched;
  R fm, fvol, r0;
};

// grid offsets of the five stencils resampleNodes builds (TwoDScene.cpp:
// 5076-5085): u, v, w faces, corners (solid phi) and cell centres (pressure)
__device__ inline double grid5_off(int gi, int a) {
  if (gi < 3) return face_off(gi, a);
  return gi == 3 ? 0.0 : 0.5;
}

// The deterministic form's fixed-point sums: each contribution rounded to
// a multiple of 2^-56 (weights) or 2^-48 (weight x velocity) and added as a
// 64-bit integer, so a node's sum does not depend on the order the
// particles reach it in (integer addition is associative).  A node's
// weight sum is about the particles per cell (the B-spline weights of a
// particle sum to one over its stencil), far below the 2^7 the format
// holds; its momentum sum stays below the 2^15 it holds for speeds below
// 1000 m/s.
constexpr double kFixW = 0x1p56, kFixP = 0x1p48;
__device__ inline unsigned long long to_fix(double v, double scale) {
  return (unsigned long long)__double2ll_rn(v * scale);
}

// R is the particle arithmetic (double, or float in the mixed-precision
// form); each particle's contributions are summed in FP64 (LDS, then the
// FP64 grid), or with DET in fixed point (LDS, then the grid's arrays
// holding 64-bit integers until k_grid_nodes converts them).
template <class R, bool DET>
__global__ __launch_bounds__(kP2GThreads) void k_p2g(P2GArgs<R> A) {
  const int64_t b = blockIdx.x;
  const uint32_t pbeg = A.bstart[b], pend = A.bstart[b + 1];
  if (pbeg == pend) return;
  const Grid& g = A.g;
  const int bi = (int)(b % g.nb[0]);
  const int bj = (int)((b / g.nb[0]) % g.nb[1]);
  const int bk = (int)(b / ((int64_t)g.nb[0] * g.nb[1]));
  const int o0 = bi * 4 - 2, o1 = bj * 4 - 2, o2 = bk * 4 - 2;

  // [axis][sum of weights, sum of weight * velocity], in FP64 for either R.
  // Particle mass and volume are uniform, so a node's mass and volume are
  // fm and fvol times its weight sum (TwoDScene.cpp:6700-6713 sums fm w and
  // fvol w): two sums per node instead of three, in LDS and in the flush
  // (the grid kernel that follows forms the mass and the volume).
  using Acc = typename std::conditional<DET, unsigned long long, double>::type;
  __shared__ Acc sm[3][2][kLdsNodes];
  __shared__ unsigned long long sphi[kLdsNodes];
  __shared__ int stouch[27];
  for (int i = threadIdx.x; i < kLdsNodes; i += kP2GThreads) {
    for (int d = 0; d < 3; ++d) {
      sm[d][0][i] = (Acc)0;
      sm[d][1][i] = (Acc)0;
    }
    sphi[i] = kPhiUntouched;
  }
  if (threadIdx.x < 27) stouch[threadIdx.x] = 0;
  __syncthreads();

  // The bucket's particles in chunks of 256; within a chunk thread t takes
  // particle (t % 32) * 8 + t / 32 (a bijection of 0..255), so the lanes of a
  // wave hold particles 8 apart: the bucket keeps its particles in their
  // previous order, eight per cell when seeded, and lanes on different cells
  // collide less on the same LDS nodes.  (Chunks of 512 with all 64 lanes on
  // different cells were slower: the waves then fetch each particle line
  // more often.)
  const uint32_t tperm = (threadIdx.x & 31) * 8 + (threadIdx.x >> 5);
  for (uint32_t c0 = pbeg; c0 < pend; c0 += kP2GThreads) {
    const uint32_t p = c0 + tperm;
    if (p >= pend) continue;
    const R px[3] = {A.x[p], A.x[A.cap + p], A.x[2 * A.cap + p]};
    // u, v, w faces: TwoDScene.cpp:6700-6713
    for (int d = 0; d < 3; ++d) {
      int base[3];
      R w[3][3];
      for (int a = 0; a < 3; ++a) {
        base[a] = stencil_base_r<R>(g, a, px[a], face_off_r<R>(d, a));
        stencil_weights_r<R>(g, px[a], base[a], face_off_r<R>(d, a), w[a]);
      }
      const R vd = A.v[d * A.cap + p];
      const R b0 = A.C[(d * 3 + 0) * A.cap + p];
      const R b1 = A.C[(d * 3 + 1) * A.cap + p];
      const R b2 = A.C[(d * 3 + 2) * A.cap + p];
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
            const R ww = w[0][c0] * w[1][c1] * w[2][c2];
            const R e0 = node_pos_r<R>(g, i, face_off_r<R>(d, 0)) - px[0];
            const R e1 = node_pos_r<R>(g, j, face_off_r<R>(d, 1)) - px[1];
            const R e2 = node_pos_r<R>(g, k, face_off_r<R>(d, 2)) - px[2];
            const R vel = vd + (b0 * e0 + b1 * e1 + b2 * e2);
            const int li = i - o0, lj = j - o1, lk = k - o2;
            if (li >= 0 && lj >= 0 && lk >= 0 && li < kTile && lj < kTile &&
                lk < kTile) {
              const int l = p2g_lds(li, lj, lk);
              if constexpr (DET) {
                atomicAdd(&sm[d][0][l], to_fix((double)ww, kFixW));
                atomicAdd(&sm[d][1][l], to_fix((double)(vel * ww), kFixP));
              } else {
                atomicAdd(&sm[d][0][l], (double)ww);
                atomicAdd(&sm[d][1][l], (double)(vel * ww));
              }
            } else {
              const int64_t n = node_index(g, i, j, k);
              if constexpr (DET) {  // the particle mass is applied by k_grid_nodes
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.wsum[d][n]),
                          to_fix((double)ww, kFixW));
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.mom[d][n]),
                          to_fix((double)(vel * ww), kFixP));
              } else {
                unsafeAtomicAdd(&A.wsum[d][n], (double)ww);
                unsafeAtomicAdd(&A.mom[d][n], (double)A.fm * (double)(vel * ww));
              }
            }
          }
    }
    // liquid phi at cell centres: TwoDScene.cpp:3307-3332
    {
      int base[3];
      for (int a = 0; a < 3; ++a) base[a] = stencil_base_r<R>(g, a, px[a], (R)0.5);
      for (int c2 = 0; c2 < 3; ++c2)
        for (int c1 = 0; c1 < 3; ++c1)
          for (int c0 = 0; c0 < 3; ++c0) {
            const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
            const R d0 = px[0] - node_pos_r<R>(g, i, (R)0.5);
            const R d1 = px[1] - node_pos_r<R>(g, j, (R)0.5);
            const R d2 = px[2] - node_pos_r<R>(g, k, (R)0.5);
            const R phi = sqrt(d0 * d0 + d1 * d1 + d2 * d2) - A.r0;
            const unsigned long long key = phi_key((double)phi);
            const int li = i - o0, lj = j - o1, lk = k - o2;
            if (li >= 0 && lj >= 0 && lk >= 0 && li < kTile && lj < kTile &&
                lk < kTile)
              atomicMin(&sphi[p2g_lds(li, lj, lk)], key);
            else
              atomicMin(&A.phikey[node_index(g, i, j, k)], key);
          }
    }
    // buckets touched by the five stencils (findNodes, TwoDScene.cpp:3683)
    for (int gi = 0; gi < 5; ++gi) {
      int blo[3], bhi[3];
      for (int a = 0; a < 3; ++a) {
        const int base = stencil_base_r<R>(g, a, px[a], (R)grid5_off(gi, a));
        blo[a] = base >> 2;
        bhi[a] = (base + 2) >> 2;
      }
      for (int zk = blo[2]; zk <= bhi[2]; ++zk)
        for (int zj = blo[1]; zj <= bhi[1]; ++zj)
          for (int zi = blo[0]; zi <= bhi[0]; ++zi) {
            const int di = zi - bi, dj = zj - bj, dk = zk - bk;
            if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 &&
                dk <= 1)
              stouch[(dk + 1) * 9 + (dj + 1) * 3 + (di + 1)] = 1;
            else
              A.touched[bucket_index(g, zi, zj, zk)] = 1;
          }
    }
  }
  __syncthreads();

  for (int t = threadIdx.x; t < kTileNodes; t += kP2GThreads) {
    const int li = t % kTile, lj = (t / kTile) % kTile, lk = t / (kTile * kTile);
    const int i = o0 + li, j = o1 + lj, k = o2 + lk;
    if (!in_range(g, i, j, k)) continue;
    const int64_t n = node_index(g, i, j, k);
    const int l = p2g_lds(li, lj, lk);
    for (int d = 0; d < 3; ++d) {
      const Acc sw = sm[d][0][l];
      if (sw != (Acc)0) {
        if constexpr (DET) {
          atomicAdd(reinterpret_cast<unsigned long long*>(&A.wsum[d][n]), sw);
          atomicAdd(reinterpret_cast<unsigned long long*>(&A.mom[d][n]), sm[d][1][l]);
        } else {
          unsafeAtomicAdd(&A.wsum[d][n], sw);
          unsafeAtomicAdd(&A.mom[d][n], (double)A.fm * sm[d][1][l]);
        }
      }
    }
    const unsigned long long key = sphi[l];
    if (key != kPhiUntouched) atomicMin(&A.phikey[n], key);
  }
  if (threadIdx.x < 27 && stouch[threadIdx.x]) {
    const int di = threadIdx.x % 3 - 1, dj = (threadIdx.x / 3) % 3 - 1,
              dk = threadIdx.x / 9 - 1;
    const int zi = bi + di, zj = bj + dj, zk = bk + dk;
    if (zi >= 0 && zj >= 0 && zk >= 0 && zi < g.nb[0] && zj < g.nb[1] &&
        zk < g.nb[2])
      A.touched[bucket_index(g, zi, zj, zk)] = 1;
  }
}

template <class R>
struct G2PArgs {
  Grid g;
  R* x;
  R* v;
  R* C;
  int64_t n, cap;
  const double* u[3];
  const uint8_t* act;
  const uint32_t* bstart;  // bucket ranges of the sorted particles (k_g2p_bucket)
  R dt, flip, invD;
  int solid_projection;
  FrozenView F;  // mode 1 records, mode 2 replays which particles the projection moves
};

// mapNodeParticlesAPIC fluid branch (TwoDScene.cpp:6958-7019), advectScene
// (SceneStepper.cpp:30-31) and solidProjection (TwoDScene.cpp:1782-1821).
//
// fetch(d, i, j, k, u) gives the face-d velocity of node (i, j, k) and
// returns true, or returns false for a node outside the active buckets
// (whose term the reference skips).  A fetch may instead give 0 for such a
// node: the weight is >= 0, so the term adds +0 or -0, which leaves every
// sum unchanged (a sum that starts at +0 never becomes -0).
// region(d, base) may return an LDS array holding face d's values at
// [(c2 * kTile + c1) * kTile + c0] for the 27 stencil nodes base + (c0, c1,
// c2) (then no node is fetched one by one), or null.
// g2p_particle_at takes the particle's position as already loaded.
// solidProjection (TwoDScene.cpp:1782-1821) of one particle: the corner-grid
// weights of the position px the transfer weights were computed at, the
// gradient toward the current position nx (advected), the displacement
// fv dt; nx is moved out of the solid when phi_now < 0.
template <class R>
__device__ __attribute__((always_inline)) inline void solid_project(const G2PArgs<R>& A,
                                                                    int64_t p, const R px[3],
                                                                    R nx[3], const R fv[3]) {
  const Grid& g = A.g;
  uint8_t* proj = A.F.mode ? &A.F.proj[A.F.pid[p]] : nullptr;
  // solidProjection moves the particle only if phi_now < 0.  The box distance
  // is 1-Lipschitz, so every stencil node's phi (or the inactive value 3 dx)
  // is at least min(3 dx, phi(px) - 2.6 dx), phi_ori, their convex
  // combination, too, and phi_now >= phi_ori - |dp| with |gr| <= 1: a
  // particle with phi(px) > 4 dx + |v|_1 dt and |v|_1 dt < 2 dx is never
  // moved, and skips the evaluation.
  {
    const R step = (fabs(fv[0]) + fabs(fv[1]) + fabs(fv[2])) * A.dt;
    const R cdx = (R)g.dx;
    if (A.F.mode != 2 && container_phi_r<R>(g, px[0], px[1], px[2]) > (R)4 * cdx + step &&
        step < (R)2 * cdx) {
      if (proj) *proj = 0;
      return;
    }
  }
  int base[3];
  R w[3][3];
  for (int a = 0; a < 3; ++a) {
    base[a] = stencil_base_r<R>(g, a, px[a], (R)0);
    stencil_weights_r<R>(g, px[a], base[a], (R)0, w[a]);
  }
  R phi_ori = 0, gr[3] = {0, 0, 0};
  for (int c2 = 0; c2 < 3; ++c2)
    for (int c1 = 0; c1 < 3; ++c1)
      for (int c0 = 0; c0 < 3; ++c0) {
        const int i = base[0] + c0, j = base[1] + c1, k = base[2] + c2;
        const R q0 = node_pos_r<R>(g, i, (R)0), q1 = node_pos_r<R>(g, j, (R)0),
                q2 = node_pos_r<R>(g, k, (R)0);
        const R phi = node_active(g, A.act, i, j, k) ? container_phi_r<R>(g, q0, q1, q2)
                                                       : (R)3 * (R)g.dx;
        const R ww = w[0][c0] * w[1][c1] * w[2][c2];
        phi_ori += phi * ww;
        gr[0] += phi * A.invD * ww * (q0 - nx[0]);
        gr[1] += phi * A.invD * ww * (q1 - nx[1]);
        gr[2] += phi * A.invD * ww * (q2 - nx[2]);
      }
  const R gn = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
  if (gn > (R)1e-20) {
    gr[0] /= gn;
    gr[1] /= gn;
    gr[2] /= gn;
  }
  const R dp0 = fv[0] * A.dt, dp1 = fv[1] * A.dt, dp2 = fv[2] * A.dt;
  const R phi_now = phi_ori + (gr[0] * dp0 + gr[1] * dp1 + gr[2] * dp2);
  const bool move = A.F.mode == 2 ? *proj != 0 : phi_now < (R)0;
  if (A.F.mode == 1) *proj = move;
  if (A.F.mode == 2 && move != (phi_now < (R)0)) fz_depart(A.F);
  if (move) {
    nx[0] -= phi_now * gr[0];
    nx[1] -= phi_now * gr[1];
    nx[2] -= phi_now * gr[2];
  }
}

// With ADVECT false only v and C are written (mapNodeParticlesAPIC alone).
template <class R, bool ADVECT = true, class Fetch, class Region>
__device__ __attribute__((always_inline)) inline void g2p_particle_at(const G2PArgs<R>& A,
                                                                      int64_t p, R x0, R x1,
                                                                      R x2, Fetch fetch,
                                                                      Region region) {
  const Grid& g = A.g;
  const R px[3] = {x0, x1, x2};
  R fv[3] = {0, 0, 0};
  R B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int base[3];
    R w[3][3];
    for (int a = 0; a < 3; ++a) {
      base[a] = stencil_base_r<R>(g, a, px[a], face_off_r<R>(d, a));
      stencil_weights_r<R>(g, px[a], base[a], face_off_r<R>(d, a), w[a]);
    }
    // node offsets from the particle along each axis (the same operations
    // for every node that shares the coordinate)
    R e[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c) e[a][c] = node_pos_r<R>(g, base[a] + c, face_off_r<R>(d, a)) - px[a];
    auto add = [&](int c0, int c1, int c2, R fnv) {
      const R ww = w[0][c0] * w[1][c1] * w[2][c2];
      fv[d] += fnv * ww;
      B[d * 3 + 0] += fnv * ww * e[0][c0] * A.invD;
      B[d * 3 + 1] += fnv * ww * e[1][c1] * A.invD;
      B[d * 3 + 2] += fnv * ww * e[2][c2] * A.invD;
    };
    if (const double* reg = region(d, base)) {
#pragma unroll
      for (int c2 = 0; c2 < 3; ++c2)
#pragma unroll
        for (int c1 = 0; c1 < 3; ++c1)
#pragma unroll
          for (int c0 = 0; c0 < 3; ++c0) add(c0, c1, c2, (R)reg[(c2 * kTile + c1) * kTile + c0]);
    } else {
#pragma unroll
      for (int c2 = 0; c2 < 3; ++c2)
#pragma unroll
        for (int c1 = 0; c1 < 3; ++c1)
#pragma unroll
          for (int c0 = 0; c0 < 3; ++c0) {
            R fnv;
            if (fetch(d, base[0] + c0, base[1] + c1, base[2] + c2, fnv)) add(c0, c1, c2, fnv);
          }
    }
  }
  for (int t = 0; t < 9; ++t) A.C[t * A.cap + p] = B[t] * A.flip;
  for (int d = 0; d < 3; ++d) A.v[d * A.cap + p] = fv[d];
  if constexpr (ADVECT) {
    R nx[3];
    for (int d = 0; d < 3; ++d) nx[d] = px[d] + fv[d] * A.dt;
    if (A.solid_projection) solid_project<R>(A, p, px, nx, fv);
    for (int d = 0; d < 3; ++d) A.x[d * A.cap + p] = nx[d];
  }
}

// One thread per particle, gathers from global memory.
template <class R, bool ADVECT, class Fetch, class Region>
__device__ __attribute__((always_inline)) inline void g2p_particle(const G2PArgs<R>& A, int64_t p,
                                                                   Fetch fetch, Region region) {
  g2p_particle_at<R, ADVECT>(A, p, A.x[p], A.x[A.cap + p], A.x[2 * A.cap + p], fetch, region);
}

template <class R, bool ADVECT>
__global__ __launch_bounds__(256) void k_g2p_advect(G2PArgs<R> A) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  g2p_particle<R, ADVECT>(
      A, p,
      [&](int d, int i, int j, int k, R& u) {
        if (!node_active(A.g, A.act, i, j, k)) return false;
        u = (R)A.u[d][node_index(A.g, i, j, k)];
        return true;
      },
      [](int, const int*) -> const double* { return nullptr; });
}

// The same per particle, with particles sorted by bucket: one workgroup per
// bucket stages the 8^3 node region of its particles' stencils (the region
// of k_p2g; inactive nodes hold 0) in LDS once, and each particle gathers
// from there (a node outside the region, which only a particle outside its
// bucket reaches, is read from global memory as above).
template <class R, bool ADVECT>
__global__ __launch_bounds__(256) void k_g2p_bucket(G2PArgs<R> A) {
  const int64_t b = blockIdx.x;
  const uint32_t pbeg = A.bstart[b], pend = A.bstart[b + 1];
  if (pbeg == pend) return;
  const Grid& g = A.g;
  const int bi = (int)(b % g.nb[0]);
  const int bj = (int)((b / g.nb[0]) % g.nb[1]);
  const int bk = (int)(b / ((int64_t)g.nb[0] * g.nb[1]));
  const int o0 = bi * 4 - 2, o1 = bj * 4 - 2, o2 = bk * 4 - 2;
  __shared__ double su[3][kTileNodes];
  // the first particle's position is loaded before the region is staged,
  // and each next one while the current one is processed
  R nx[3] = {0, 0, 0};
  if (pbeg + threadIdx.x < pend) {
    const uint32_t q = pbeg + threadIdx.x;
    nx[0] = A.x[q];
    nx[1] = A.x[A.cap + q];
    nx[2] = A.x[2 * A.cap + q];
  }
  for (int l = threadIdx.x; l < kTileNodes; l += 256) {
    const int i = o0 + (l % kTile), j = o1 + ((l / kTile) % kTile), k = o2 + l / (kTile * kTile);
    const bool on = node_active(g, A.act, i, j, k);
    const int64_t n = on ? node_index(g, i, j, k) : 0;
    for (int d = 0; d < 3; ++d) su[d][l] = on ? A.u[d][n] : 0.0;
  }
  __syncthreads();
  for (uint32_t c0 = pbeg; c0 < pend; c0 += 256) {
    const uint32_t p = c0 + threadIdx.x;
    if (p >= pend) continue;
    const R x0 = nx[0], x1 = nx[1], x2 = nx[2];
    // this thread's next particle (no other thread writes it before then)
    if (p + 256 < pend) {
      nx[0] = A.x[p + 256];
      nx[1] = A.x[A.cap + p + 256];
      nx[2] = A.x[2 * A.cap + p + 256];
    }
    g2p_particle_at<R, ADVECT>(
        A, p, x0, x1, x2,
        [&](int d, int i, int j, int k, R& u) {
          const int li = i - o0, lj = j - o1, lk = k - o2;
          if (li >= 0 && lj >= 0 && lk >= 0 && li < kTile && lj < kTile && lk < kTile) {
            u = (R)su[d][(lk * kTile + lj) * kTile + li];
            return true;
          }
          if (!node_active(g, A.act, i, j, k)) return false;
          u = (R)A.u[d][node_index(g, i, j, k)];
          return true;
        },
        // the whole stencil inside the region (always, for a particle in
        // its bucket): one check per face instead of one per node
        [&](int d, const int* base) -> const double* {
          const int l0 = base[0] - o0, l1 = base[1] - o1, l2 = base[2] - o2;
          const bool in = (unsigned)l0 <= (unsigned)(kTile - 3) &&
                          (unsigned)l1 <= (unsigned)(kTile - 3) &&
                          (unsigned)l2 <= (unsigned)(kTile - 3);
          return in ? &su[d][(l2 * kTile + l1) * kTile + l0] : nullptr;
        });
  }
}

// SceneStepper::advectScene, fluid branch (SceneStepper.cpp:30-31): x += v dt.
// The position before the move is kept in xw: solidProjection weighs the
// corner grid at the position the transfer weights were computed at.
template <class R>
__global__ __launch_bounds__(256) void k_advect(R* x, const R* v, R* xw, int64_t n, int64_t cap,
                                               R dt) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int d = 0; d < 3; ++d) {
    const R x0 = x[d * cap + p];
    xw[d * cap + p] = x0;
    x[d * cap + p] = x0 + v[d * cap + p] * dt;
  }
}

// TwoDScene::solidProjection (TwoDScene.cpp:1776-1822) as its own pass: the
// weights at xw (or at x when no advection has run since the transfer), the
// gradient toward the current x.
template <class R>
__global__ __launch_bounds__(256) void k_solid_projection(G2PArgs<R> A, const R* xw) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  const R* xs = xw ? xw : A.x;
  const R px[3] = {xs[p], xs[A.cap + p], xs[2 * A.cap + p]};
  R nx[3] = {A.x[p], A.x[A.cap + p], A.x[2 * A.cap + p]};
  const R fv[3] = {A.v[p], A.v[A.cap + p], A.v[2 * A.cap + p]};
  solid_project<R>(A, p, px, nx, fv);
  for (int d = 0; d < 3; ++d) A.x[d * A.cap + p] = nx[d];
}

// TwoDScene::getMaxFluidVelocity (TwoDScene.cpp:4180-4188): the largest
// squared norm of a fluid velocity, in FP64, as the bits of a non-negative
// double (their order is the values' order) by atomic max.
template <class R>
__global__ __launch_bounds__(256) void k_max_vel2(const R* v, int64_t n, int64_t cap,
                                                 unsigned long long* out) {
  double m = 0.0;
  for (int64_t p = blockIdx.x * 256ll + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256) {
    const double a = (double)v[p], b = (double)v[cap + p], c = (double)v[2 * cap + p];
    const double q = a * a + b * b + c * c;
    m = q > m ? q : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(m, o, 64);
    m = t > m ? t : m;
  }
  __shared__ double sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) m = sm[w] > m ? sm[w] : m;
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

// TwoDScene::updateParticleBoundingBox's reduction (TwoDScene.cpp:1588-1604):
// the particles' componentwise min and max, as order-preserving 64-bit keys
// (out[0..2] min, out[3..5] max; min slots start at ~0, max slots at 0).
template <class R>
__global__ __launch_bounds__(256) void k_bbox(const R* x, int64_t n, int64_t cap,
                                             unsigned long long* out) {
  unsigned long long lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0, 0, 0};
  for (int64_t p = blockIdx.x * 256ll + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256)
    for (int d = 0; d < 3; ++d) {
      const unsigned long long k = phi_key((double)x[d * cap + p]);
      lo[d] = k < lo[d] ? k : lo[d];
      hi[d] = k > hi[d] ? k : hi[d];
    }
  for (int o = 32; o > 0; o >>= 1)
    for (int d = 0; d < 3; ++d) {
      const unsigned long long a = __shfl_xor(lo[d], o, 64), b = __shfl_xor(hi[d], o, 64);
      lo[d] = a < lo[d] ? a : lo[d];
      hi[d] = b > hi[d] ? b : hi[d];
    }
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; ++d) {
      atomicMin(&out[d], lo[d]);
      atomicMax(&out[3 + d], hi[d]);
    }
}

// ---- seeding ----

__global__ void k_seed_flags(Grid g, int c00, int c01, int c02, int cn0,
                             int cn1, int cn2, double lo0, double lo1,
                             double lo2, double hi0, double hi1, double hi2,
                             uint64_t seed, uint32_t* flag) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ncand = (int64_t)cn0 * cn1 * cn2 * 8;
  if (id >= ncand) return;
  const int sub = (int)(id & 7);
  const int64_t cell = id >> 3;
  const int ii = (int)(cell % cn0), jj = (int)((cell / cn0) % cn1),
            kk = (int)(cell / ((int64_t)cn0 * cn1));
  const int cc[3] = {c00 + ii, c01 + jj, c02 + kk};
  const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
  uint32_t keep = 1;
  for (int d = 0; d < 3; ++d) {
    const int sc = (sub >> d) & 1;
    const double u = uniform01(hash3(seed, (uint64_t)id, (uint64_t)d));
    const double p = ((double)cc[d] + ((double)sc + u) * 0.5) * g.dx;
    if (!(p >= lo[d] && p < hi[d])) keep = 0;
  }
  flag[id] = keep;
}

__global__ void k_seed_write(Grid g, int c00, int c01, int c02, int cn0,
                             int cn1, int cn2, uint64_t seed,
                             const uint32_t* flag_scan, const uint32_t* flag,
                             double* x, double* v, double* C, uint32_t* pid,
                             int64_t cap) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ncand = (int64_t)cn0 * cn1 * cn2 * 8;
  if (id >= ncand || !flag[id]) return;
  const int64_t q = flag_scan[id];
  const int sub = (int)(id & 7);
  const int64_t cell = id >> 3;
  const int ii = (int)(cell % cn0), jj = (int)((cell / cn0) % cn1),
            kk = (int)(cell / ((int64_t)cn0 * cn1));
  const int cc[3] = {c00 + ii, c01 + jj, c02 + kk};
  for (int d = 0; d < 3; ++d) {
    const int sc = (sub >> d) & 1;
    const double u = uniform01(hash3(seed, (uint64_t)id, (uint64_t)d));
    x[d * cap + q] = ((double)cc[d] + ((double)sc + u) * 0.5) * g.dx;
    v[d * cap + q] = 0.0;
  }
  for (int t = 0; t < 9; ++t) C[t * cap + q] = 0.0;
  pid[q] = (uint32_t)q;
}

// Irwin-Hall N(0,1) per world cell and component, cells [-2, n+1]
__global__ void k_noise(Grid g, uint64_t seed, double* raw) {
  const int q0 = g.n[0] + 4, q1 = g.n[1] + 4, q2 = g.n[2] + 4;
  const int64_t qq = (int64_t)q0 * q1 * q2;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 3 * qq) return;
  const int c = (int)(t / qq);
  const int64_t r = t % qq;
  const int i = (int)(r % q0) - 2, j = (int)((r / q0) % q1) - 2,
            k = (int)(r / ((int64_t)q0 * q1)) - 2;
  const uint64_t key = ((uint64_t)(k + (1 << 20)) << 42) |
                       ((uint64_t)(j + (1 << 20)) << 21) |
                       (uint64_t)(i + (1 << 20));
  double s = 0.0;
  for (int u = 0; u < 12; ++u)
    s += uniform01(hash3(seed ^ 0xA076D1F0C3B5E927ull, key,
                         (uint64_t)(c * 16 + u)));
  raw[t] = s - 6.0;
}

// 3x3x3 box filter, world cells [-1, n], scaled to std sigma
__global__ void k_noise_filter(Grid g, const double* raw, double scale,
                               double* fil) {
  const int q0 = g.n[0] + 4, q1 = g.n[1] + 4, q2 = g.n[2] + 4;
  const int m0 = g.n[0] + 2, m1 = g.n[1] + 2, m2 = g.n[2] + 2;
  const int64_t qq = (int64_t)q0 * q1 * q2, mm = (int64_t)m0 * m1 * m2;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 3 * mm) return;
  const int c = (int)(t / mm);
  const int64_t r = t % mm;
  const int i = (int)(r % m0), j = (int)((r / m0) % m1),
            k = (int)(r / ((int64_t)m0 * m1));
  double s = 0.0;
  for (int dk = 0; dk < 3; ++dk)
    for (int dj = 0; dj < 3; ++dj)
      for (int di = 0; di < 3; ++di)
        s += raw[c * qq + ((int64_t)(k + dk) * q1 + (j + dj)) * q0 + (i + di)];
  fil[t] = s * scale;
}

__global__ void k_seed_velocity(Grid g, const double* fil, const double* x,
                                double* v, int64_t n, int64_t cap) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int m0 = g.n[0] + 2, m1 = g.n[1] + 2, m2 = g.n[2] + 2;
  const int64_t mm = (int64_t)m0 * m1 * m2;
  const double px[3] = {x[p], x[cap + p], x[2 * cap + p]};
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
          const int wi = base[0] + c0 - kGhost + 1;
          const int wj = base[1] + c1 - kGhost + 1;
          const int wk = base[2] + c2 - kGhost + 1;
          double f = 0.0;
          if (wi >= 0 && wj >= 0 && wk >= 0 && wi < m0 && wj < m1 && wk < m2)
            f = fil[c * mm + ((int64_t)wk * m1 + wj) * m0 + wi];
          const double ww = w[0][c0] * w[1][c1] * w[2][c2];
          acc += ww * f;
        }
    v[c * cap + p] = acc;
  }
}

inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// The particle planes of one precision: the current set and the sort target.
template <class R>
struct Planes {
  DevBuf<R>*x, *v, *C, *x2, *v2, *C2;
};
inline Planes<double> planes(wc_ctx* c, double) {
  return {&c->x, &c->v, &c->C, &c->x2, &c->v2, &c->C2};
}
inline Planes<float> planes(wc_ctx* c, float) {
  return {&c->fx, &c->fv, &c->fC, &c->fx2, &c->fv2, &c->fC2};
}

template <class R>
void release_planes(Planes<R> P) {
  P.x->release();
  P.v->release();
  P.C->release();
  P.x2->release();
  P.v2->release();
  P.C2->release();
}

void ensure_particles(wc_ctx* ctx, int64_t n) {
  const bool have = ctx->mixed ? ctx->fx.p != nullptr : ctx->x.p != nullptr;
  if (n > ctx->pcap || !have) {
    const int64_t cap = n > 0 ? n : 1;
    release_planes(planes(ctx, 0.0));
    release_planes(planes(ctx, 0.0f));
    ctx->pid.release();
    ctx->pid2.release();
    ctx->pkey.release();
    auto alloc = [&](auto P) {
      P.x->ensure(3 * cap);
      P.v->ensure(3 * cap);
      P.C->ensure(9 * cap);
      P.x2->ensure(3 * cap);
      P.v2->ensure(3 * cap);
      P.C2->ensure(9 * cap);
    };
    if (ctx->mixed)
      alloc(planes(ctx, 0.0f));
    else
      alloc(planes(ctx, 0.0));
    ctx->pid.ensure(cap);
    ctx->pid2.ensure(cap);
    ctx->pkey.ensure(cap);
    ctx->pcap = cap;
  }
  ctx->np = n;
  ctx->sorted = false;
  ctx->xw_valid = false;
}

__global__ void k_narrow(const double* a, float* b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// Upload host planes (FP64) into the current set, rounding to FP32 for the
// mixed form.
template <class R>
void upload(hipStream_t st, DevBuf<R>& dst, const std::vector<double>& h) {
  if constexpr (sizeof(R) == sizeof(double)) {
    WC_HIP(hipMemcpyAsync(dst.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, st));
  } else {
    std::vector<R> t(h.size());
    for (size_t i = 0; i < h.size(); ++i) t[i] = (R)h[i];
    WC_HIP(hipMemcpyAsync(dst.p, t.data(), sizeof(R) * t.size(), hipMemcpyHostToDevice, st));
    WC_HIP(hipStreamSynchronize(st));
  }
}

template <class R>
void download(hipStream_t st, const DevBuf<R>& src, std::vector<double>& h) {
  if constexpr (sizeof(R) == sizeof(double)) {
    WC_HIP(hipMemcpyAsync(h.data(), src.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  } else {
    std::vector<R> t(h.size());
    WC_HIP(hipMemcpyAsync(t.data(), src.p, sizeof(R) * t.size(), hipMemcpyDeviceToHost, st));
    WC_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < h.size(); ++i) h[i] = (double)t[i];
  }
}

template <class R>
void sort_impl(wc_ctx* ctx, Planes<R> P) {
  const Grid& g = ctx->g;
  hipStream_t st = ctx->stream;
  const int64_t n = ctx->np, cap = ctx->pcap;
  ctx->adj.ready = false;
  ctx->bcount.ensure(g.nbucket + 1);
  ctx->bcursor.ensure(g.nbucket + 1);
  WC_HIP(hipMemsetAsync(ctx->bcount.p, 0, sizeof(uint32_t) * (g.nbucket + 1), st));
  WC_HIP(hipMemsetAsync(ctx->bcursor.p, 0, sizeof(uint32_t) * (g.nbucket + 1), st));
  if (n > 0) {
    hipLaunchKernelGGL(k_sort_keys<R>, dim3(blocks(n, 256)), dim3(256), 0, st, g,
                       P.x->p, n, cap, ctx->pkey.p, ctx->bcount.p);
    WC_LAUNCH_CHECK();
  }
  exclusive_scan(ctx, ctx->bcount.p, g.nbucket + 1);  // bcount becomes bstart
  std::swap(ctx->bcount, ctx->bstart);
  ctx->bcount.ensure(g.nbucket + 1);
  if (n > 0) {
    hipLaunchKernelGGL(k_sort_scatter<R>, dim3(blocks(n, 256)), dim3(256), 0, st, n,
                       cap, ctx->pkey.p, ctx->bstart.p, ctx->bcursor.p, P.x->p,
                       P.v->p, P.C->p, ctx->pid.p, P.x2->p, P.v2->p, P.C2->p,
                       ctx->pid2.p);
    WC_LAUNCH_CHECK();
  }
  std::swap(*P.x, *P.x2);
  std::swap(*P.v, *P.v2);
  std::swap(*P.C, *P.C2);
  std::swap(ctx->pid, ctx->pid2);
  ctx->sorted = true;
  ctx->xw_valid = false;
  if (ctx->frz_mode) frozen_after_sort(ctx);
}

template <class R>
void p2g_launch(wc_ctx* ctx, Planes<R> P) {
  const Grid& g = ctx->g;
  P2GArgs<R> A;
  A.g = g;
  A.x = P.x->p;
  A.v = P.v->p;
  A.C = P.C->p;
  A.cap = ctx->pcap;
  A.bstart = ctx->bstart.p;
  for (int d = 0; d < 3; ++d) {
    A.wsum[d] = ctx->mass[d].p;  // weight sums; grid_nodes forms the masses
    A.mom[d] = ctx->mom[d].p;
  }
  A.phikey = ctx->phikey.p;
  A.touched = ctx->touched.p;
  A.fm = (R)ctx->desc.particle_mass;
  A.fvol = (R)ctx->desc.particle_volume;
  A.r0 = (R)(g.dx * 0.883644 > ctx->desc.particle_radius ? g.dx * 0.883644
                                                          : ctx->desc.particle_radius);
  KTimer& tm = ctx->pcg.timer;
  hipEvent_t e0 = tm.on ? tm.mark(ctx->stream) : nullptr;
  if (ctx->deterministic)
    hipLaunchKernelGGL((k_p2g<R, true>), dim3((unsigned)g.nbucket), dim3(kP2GThreads), 0,
                       ctx->stream, A);
  else
    hipLaunchKernelGGL((k_p2g<R, false>), dim3((unsigned)g.nbucket), dim3(kP2GThreads), 0,
                       ctx->stream, A);
  WC_LAUNCH_CHECK();
  if (e0) tm.add(KTimer::P2G, ctx->np, e0, tm.mark(ctx->stream));
}

template <class R>
G2PArgs<R> g2p_args(wc_ctx* ctx, Planes<R> P, double dt) {
  G2PArgs<R> A;
  A.g = ctx->g;
  A.x = P.x->p;
  A.v = P.v->p;
  A.C = P.C->p;
  A.n = ctx->np;
  A.cap = ctx->pcap;
  for (int d = 0; d < 3; ++d) A.u[d] = ctx->unew[d].p;
  A.act = ctx->act.p;
  A.dt = (R)dt;
  A.flip = (R)ctx->desc.flip_coeff;
  A.invD = (R)(4.0 / (ctx->g.dx * ctx->g.dx));
  A.solid_projection = ctx->desc.solid_projection;
  A.bstart = ctx->bstart.p;
  A.F = frozen_view(ctx);
  return A;
}

template <class R, bool ADVECT>
void g2p_launch(wc_ctx* ctx, Planes<R> P, double dt) {
  const G2PArgs<R> A = g2p_args(ctx, P, dt);
  KTimer& tm = ctx->pcg.timer;
  hipEvent_t e0 = tm.on ? tm.mark(ctx->stream) : nullptr;
  // with bstart describing the particle arrays, one workgroup per bucket
  if (ctx->sorted)
    hipLaunchKernelGGL((k_g2p_bucket<R, ADVECT>), dim3((unsigned)ctx->g.nbucket), dim3(256), 0,
                       ctx->stream, A);
  else
    hipLaunchKernelGGL((k_g2p_advect<R, ADVECT>), dim3(blocks(ctx->np, 256)), dim3(256), 0,
                       ctx->stream, A);
  WC_LAUNCH_CHECK();
  if (e0) tm.add(KTimer::G2P, ctx->np, e0, tm.mark(ctx->stream));
}

// the weights-position planes of the separate advection / projection passes
inline DevBuf<double>& xw_planes(wc_ctx* c, double) { return c->xw; }
inline DevBuf<float>& xw_planes(wc_ctx* c, float) { return c->fxw; }

template <class R>
void advect_launch(wc_ctx* ctx, Planes<R> P, double dt) {
  DevBuf<R>& xw = xw_planes(ctx, R());
  xw.ensure(3 * (size_t)ctx->pcap);
  hipLaunchKernelGGL(k_advect<R>, dim3(blocks(ctx->np, 256)), dim3(256), 0, ctx->stream,
                     P.x->p, P.v->p, xw.p, ctx->np, ctx->pcap, (R)dt);
  WC_LAUNCH_CHECK();
  ctx->xw_valid = true;
}

template <class R>
void projection_launch(wc_ctx* ctx, Planes<R> P, double dt) {
  G2PArgs<R> A = g2p_args(ctx, P, dt);
  const R* xw = ctx->xw_valid ? xw_planes(ctx, R()).p : nullptr;
  hipLaunchKernelGGL(k_solid_projection<R>, dim3(blocks(ctx->np, 256)), dim3(256), 0,
                     ctx->stream, A, xw);
  WC_LAUNCH_CHECK();
}

template <class R>
double max_vel_launch(wc_ctx* ctx, Planes<R> P) {
  ctx->vmax_bits.ensure(1);
  WC_HIP(hipMemsetAsync(ctx->vmax_bits.p, 0, sizeof(unsigned long long), ctx->stream));
  const unsigned nb = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, blocks(ctx->np, 256)));
  hipLaunchKernelGGL(k_max_vel2<R>, dim3(nb), dim3(256), 0, ctx->stream, P.v->p, ctx->np,
                     ctx->pcap, ctx->vmax_bits.p);
  WC_LAUNCH_CHECK();
  unsigned long long bits = 0;
  WC_HIP(hipMemcpyAsync(&bits, ctx->vmax_bits.p, sizeof(bits), hipMemcpyDeviceToHost,
                        ctx->stream));
  WC_HIP(hipStreamSynchronize(ctx->stream));
  double m2;
  memcpy(&m2, &bits, sizeof(m2));
  return sqrt(m2);
}

}  // namespace

void particles_set_soa(wc_ctx* ctx, int64_t n, const double* x, const double* v,
                       const double* C, bool ref_layout) {
  ensure_particles(ctx, n);
  ctx->adj.ready = false;
  ctx->adj.have_start = false;
  const int64_t cap = ctx->pcap;
  std::vector<double> hx(3 * (size_t)cap, 0.0), hv(3 * (size_t)cap, 0.0),
      hC(9 * (size_t)cap, 0.0);
  std::vector<uint32_t> hp((size_t)cap);
  for (int64_t p = 0; p < n; ++p) {
    hp[p] = (uint32_t)p;
    for (int d = 0; d < 3; ++d) {
      if (x) hx[d * cap + p] = ref_layout ? x[4 * p + d] : x[d * n + p];
      if (v) hv[d * cap + p] = ref_layout ? v[4 * p + d] : v[d * n + p];
    }
    if (C)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          hC[(r * 3 + c) * cap + p] =
              ref_layout ? C[(int64_t)c * 3 * n + 3 * p + r]  // Eigen 3N x 3
                         : C[(r * 3 + c) * n + p];
  }
  hipStream_t st = ctx->stream;
  auto put = [&](auto P) {
    upload(st, *P.x, hx);
    upload(st, *P.v, hv);
    upload(st, *P.C, hC);
  };
  if (ctx->mixed)
    put(planes(ctx, 0.0f));
  else
    put(planes(ctx, 0.0));
  WC_HIP(hipMemcpyAsync(ctx->pid.p, hp.data(), sizeof(uint32_t) * cap, hipMemcpyHostToDevice, st));
  WC_HIP(hipStreamSynchronize(st));
}

void particles_get_soa(wc_ctx* ctx, double* x, double* v, double* C,
                       bool ref_layout) {
  const int64_t n = ctx->np, cap = ctx->pcap;
  std::vector<double> hx(3 * (size_t)cap), hv(3 * (size_t)cap),
      hC(9 * (size_t)cap);
  std::vector<uint32_t> hp((size_t)cap);
  hipStream_t st = ctx->stream;
  auto get = [&](auto P) {
    download(st, *P.x, hx);
    download(st, *P.v, hv);
    download(st, *P.C, hC);
  };
  if (ctx->mixed)
    get(planes(ctx, 0.0f));
  else
    get(planes(ctx, 0.0));
  WC_HIP(hipMemcpyAsync(hp.data(), ctx->pid.p, sizeof(uint32_t) * cap, hipMemcpyDeviceToHost, st));
  WC_HIP(hipStreamSynchronize(st));
  for (int64_t q = 0; q < n; ++q) {
    const int64_t p = hp[q];
    for (int d = 0; d < 3; ++d) {
      if (x) {
        if (ref_layout) {
          x[4 * p + d] = hx[d * cap + q];
        } else {
          x[d * n + p] = hx[d * cap + q];
        }
      }
      if (v) {
        if (ref_layout) {
          v[4 * p + d] = hv[d * cap + q];
        } else {
          v[d * n + p] = hv[d * cap + q];
        }
      }
    }
    if (ref_layout) {
      if (x) x[4 * p + 3] = 0.0;
      if (v) v[4 * p + 3] = 0.0;
    }
    if (C)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          const double val = hC[(r * 3 + c) * cap + q];
          if (ref_layout)
            C[(int64_t)c * 3 * n + 3 * p + r] = val;
          else
            C[(r * 3 + c) * n + p] = val;
        }
  }
}

int64_t particles_seed_block(wc_ctx* ctx, const double lo[3], const double hi[3],
                             uint64_t seed, double sigma) {
  const Grid& g = ctx->g;
  hipStream_t st = ctx->stream;
  int c0[3], cn[3];
  for (int d = 0; d < 3; ++d) {
    c0[d] = (int)floor(lo[d] / g.dx);
    const int c1 = (int)ceil(hi[d] / g.dx);
    cn[d] = c1 > c0[d] ? c1 - c0[d] : 0;
  }
  const int64_t ncand = (int64_t)cn[0] * cn[1] * cn[2] * 8;
  if (ncand >= 0xFFFFFFFFll) throw HipError{hipErrorInvalidValue, "seed block too large", __FILE__, __LINE__};
  DevBuf<uint32_t> flag, scan;
  flag.ensure(ncand + 1);
  scan.ensure(ncand + 1);
  hipLaunchKernelGGL(k_seed_flags, dim3(blocks(ncand, 256)), dim3(256), 0, st, g,
                     c0[0], c0[1], c0[2], cn[0], cn[1], cn[2], lo[0], lo[1],
                     lo[2], hi[0], hi[1], hi[2], seed, flag.p);
  WC_LAUNCH_CHECK();
  WC_HIP(hipMemcpyAsync(scan.p, flag.p, sizeof(uint32_t) * ncand,
                        hipMemcpyDeviceToDevice, st));
  const uint32_t n = exclusive_scan(ctx, scan.p, ncand);
  ensure_particles(ctx, n);
  ctx->adj.ready = false;
  ctx->adj.have_start = false;
  // Seeds are generated in FP64; the mixed form rounds them once to FP32.
  const int64_t cap = ctx->pcap;
  DevBuf<double> tx, tv, tC;
  double *sx = ctx->x.p, *sv = ctx->v.p, *sC = ctx->C.p;
  if (ctx->mixed) {
    tx.ensure(3 * cap);
    tv.ensure(3 * cap);
    tC.ensure(9 * cap);
    sx = tx.p;
    sv = tv.p;
    sC = tC.p;
  }
  hipLaunchKernelGGL(k_seed_write, dim3(blocks(ncand, 256)), dim3(256), 0, st, g,
                     c0[0], c0[1], c0[2], cn[0], cn[1], cn[2], seed, scan.p,
                     flag.p, sx, sv, sC, ctx->pid.p, cap);
  WC_LAUNCH_CHECK();
  if (sigma > 0.0) {
    const int64_t qq = (int64_t)(g.n[0] + 4) * (g.n[1] + 4) * (g.n[2] + 4);
    const int64_t mm = (int64_t)(g.n[0] + 2) * (g.n[1] + 2) * (g.n[2] + 2);
    DevBuf<double> raw, fil;
    raw.ensure(3 * qq);
    fil.ensure(3 * mm);
    hipLaunchKernelGGL(k_noise, dim3(blocks(3 * qq, 256)), dim3(256), 0, st, g,
                       seed, raw.p);
    const double scale = sigma / sqrt(27.0);
    hipLaunchKernelGGL(k_noise_filter, dim3(blocks(3 * mm, 256)), dim3(256), 0,
                       st, g, raw.p, scale, fil.p);
    hipLaunchKernelGGL(k_seed_velocity, dim3(blocks(n, 256)), dim3(256), 0, st,
                       g, fil.p, sx, sv, (int64_t)n, cap);
    WC_LAUNCH_CHECK();
    WC_HIP(hipStreamSynchronize(st));
    raw.release();
    fil.release();
  }
  if (ctx->mixed) {
    hipLaunchKernelGGL(k_narrow, dim3(blocks(3 * cap, 256)), dim3(256), 0, st, sx,
                       ctx->fx.p, 3 * cap);
    hipLaunchKernelGGL(k_narrow, dim3(blocks(3 * cap, 256)), dim3(256), 0, st, sv,
                       ctx->fv.p, 3 * cap);
    hipLaunchKernelGGL(k_narrow, dim3(blocks(9 * cap, 256)), dim3(256), 0, st, sC,
                       ctx->fC.p, 9 * cap);
    WC_LAUNCH_CHECK();
  }
  WC_HIP(hipStreamSynchronize(st));
  tx.release();
  tv.release();
  tC.release();
  flag.release();
  scan.release();
  return n;
}

void particles_sort(wc_ctx* ctx) {
  if (ctx->mixed)
    sort_impl(ctx, planes(ctx, 0.0f));
  else
    sort_impl(ctx, planes(ctx, 0.0));
}

void particles_p2g(wc_ctx* ctx) {
  // the workgroup-per-bucket P2G reads the bucket offsets of the last sort
  if (!ctx->sorted) particles_sort(ctx);
  const Grid& g = ctx->g;
  hipStream_t st = ctx->stream;
  const int64_t nn = g.nnode;
  for (int d = 0; d < 3; ++d) {
    WC_HIP(hipMemsetAsync(ctx->mass[d].p, 0, sizeof(double) * nn, st));
    WC_HIP(hipMemsetAsync(ctx->mom[d].p, 0, sizeof(double) * nn, st));
  }
  WC_HIP(hipMemsetAsync(ctx->phikey.p, 0xFF, sizeof(unsigned long long) * nn, st));
  WC_HIP(hipMemsetAsync(ctx->touched.p, 0, g.nbucket, st));
  if (ctx->np == 0) return;
  if (ctx->mixed)
    p2g_launch(ctx, planes(ctx, 0.0f));
  else
    p2g_launch(ctx, planes(ctx, 0.0));
}

void particles_g2p_advect(wc_ctx* ctx, double dt) {
  if (ctx->np == 0) return;
  if (ctx->mixed)
    g2p_launch<float, true>(ctx, planes(ctx, 0.0f), dt);
  else
    g2p_launch<double, true>(ctx, planes(ctx, 0.0), dt);
  ctx->xw_valid = false;
}

void particles_g2p(wc_ctx* ctx) {
  if (ctx->np == 0) return;
  if (ctx->mixed)
    g2p_launch<float, false>(ctx, planes(ctx, 0.0f), 0.0);
  else
    g2p_launch<double, false>(ctx, planes(ctx, 0.0), 0.0);
  ctx->xw_valid = false;
}

void particles_advect(wc_ctx* ctx, double dt) {
  if (ctx->np == 0) return;
  if (ctx->mixed)
    advect_launch(ctx, planes(ctx, 0.0f), dt);
  else
    advect_launch(ctx, planes(ctx, 0.0), dt);
}

void particles_solid_projection(wc_ctx* ctx, double dt) {
  if (ctx->np == 0) return;
  if (ctx->mixed)
    projection_launch(ctx, planes(ctx, 0.0f), dt);
  else
    projection_launch(ctx, planes(ctx, 0.0), dt);
}

void particles_bbox(wc_ctx* ctx, double lo[3], double hi[3]) {
  for (int d = 0; d < 3; ++d) {
    lo[d] = 1e+20;  // the reduction's identities (TwoDScene.cpp:1589-1590)
    hi[d] = -1e+20;
  }
  if (ctx->np == 0) return;
  ctx->vmax_bits.ensure(6);
  unsigned long long init[6] = {~0ull, ~0ull, ~0ull, 0, 0, 0};
  WC_HIP(hipMemcpyAsync(ctx->vmax_bits.p, init, sizeof(init), hipMemcpyHostToDevice,
                        ctx->stream));
  const unsigned nb = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, blocks(ctx->np, 256)));
  if (ctx->mixed)
    hipLaunchKernelGGL(k_bbox<float>, dim3(nb), dim3(256), 0, ctx->stream, ctx->fx.p, ctx->np,
                       ctx->pcap, ctx->vmax_bits.p);
  else
    hipLaunchKernelGGL(k_bbox<double>, dim3(nb), dim3(256), 0, ctx->stream, ctx->x.p, ctx->np,
                       ctx->pcap, ctx->vmax_bits.p);
  WC_LAUNCH_CHECK();
  unsigned long long k[6];
  WC_HIP(hipMemcpyAsync(k, ctx->vmax_bits.p, sizeof(k), hipMemcpyDeviceToHost, ctx->stream));
  WC_HIP(hipStreamSynchronize(ctx->stream));
  for (int d = 0; d < 3; ++d) {
    lo[d] = key_value(k[d]);
    hi[d] = key_value(k[3 + d]);
  }
}

double particles_max_velocity(wc_ctx* ctx) {
  if (ctx->np == 0) return 0.0;
  return ctx->mixed ? max_vel_launch(ctx, planes(ctx, 0.0f)) : max_vel_launch(ctx, planes(ctx, 0.0));
}

}  // namespace wc
