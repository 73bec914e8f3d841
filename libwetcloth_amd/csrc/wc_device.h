// Device helpers shared by the gfx950 kernels: the dense MAC grid in 4x4x4
// bucket-major order, the quadratic B-spline, fraction_inside, the box
// container's distance field and the counter-based RNG.
//
// Every floating-point expression here is written operation for operation
// like the reference function it restates, and the library is compiled with
// -ffp-contract=off, so the results are bitwise those of the FP64 CPU form.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wc {

// Two buckets of ghost cells below the container: the B2 stencils of
// particles in the container reach one cell below it, into the bucket just
// below, and the one-bucket dilation of the activated set
// (expandFluidNodesMarked, TwoDScene.cpp:5021-5064) reaches one bucket
// further.  The reference's grid extends three buckets past the particles'
// bounding box (TwoDScene.cpp:1657-1690), so that dilation is never clipped
// there; with two ghost buckets it is not clipped here either, and the
// activated buckets, and with them the pressure DOFs, are the reference's.
constexpr int kGhost = 8;
constexpr int kBucket = 4;  // cells per bucket edge (every reference scene)

struct Grid {
  int n[3];     // container cells per axis
  double dx;
  int nx[3];    // dense nodes per axis
  int nb[3];    // buckets per axis
  int64_t nnode;
  int64_t nbucket;
};

inline Grid make_grid(int ni, int nj, int nk, double dx) {
  Grid g;
  g.n[0] = ni;
  g.n[1] = nj;
  g.n[2] = nk;
  g.dx = dx;
  for (int d = 0; d < 3; ++d) {
    // two buckets below (kGhost) and two above the container
    g.nb[d] = (g.n[d] + kBucket - 1) / kBucket + 4;
    g.nx[d] = g.nb[d] * kBucket;
  }
  g.nbucket = (int64_t)g.nb[0] * g.nb[1] * g.nb[2];
  g.nnode = g.nbucket * 64;
  return g;
}

__host__ __device__ inline int64_t bucket_index(const Grid& g, int bi, int bj,
                                                int bk) {
  return ((int64_t)bk * g.nb[1] + bj) * g.nb[0] + bi;
}

// TwoDScene.cpp:3754-3761 local numbering k*n^2 + j*n + i inside a bucket,
// buckets in Sorter::bucket_index order (Sorter.h:66-68).
__host__ __device__ inline int64_t node_index(const Grid& g, int i, int j,
                                              int k) {
  return (bucket_index(g, i >> 2, j >> 2, k >> 2) << 6) |
         ((int64_t)(k & 3) << 4) | ((j & 3) << 2) | (i & 3);
}

__host__ __device__ inline void node_coords(const Grid& g, int64_t n, int& i,
                                            int& j, int& k) {
  const int64_t b = n >> 6;
  const int l = (int)(n & 63);
  const int bi = (int)(b % g.nb[0]);
  const int bj = (int)((b / g.nb[0]) % g.nb[1]);
  const int bk = (int)(b / ((int64_t)g.nb[0] * g.nb[1]));
  i = bi * 4 + (l & 3);
  j = bj * 4 + ((l >> 2) & 3);
  k = bk * 4 + (l >> 4);
}

__host__ __device__ inline bool in_range(const Grid& g, int i, int j, int k) {
  return i >= 0 && j >= 0 && k >= 0 && i < g.nx[0] && j < g.nx[1] &&
         k < g.nx[2];
}

__device__ inline bool node_active(const Grid& g, const uint8_t* act, int i,
                                   int j, int k) {
  if (!in_range(g, i, j, k)) return false;
  return act[bucket_index(g, i >> 2, j >> 2, k >> 2)] != 0;
}

__host__ __device__ inline double node_pos(const Grid& g, int i, double off) {
  return ((double)(i - kGhost) + off) * g.dx;
}

// MathUtilities.h:416-424
__host__ __device__ inline double quad_kernel(double x) {
  const double ax = fabs(x);
  if (ax < 0.5) return 0.75 - ax * ax;
  if (ax < 1.5) return 0.5 * (1.5 - ax) * (1.5 - ax);
  return 0.0;
}

// MathUtilities.h:286-293
__host__ __device__ inline double fraction_inside2(double l, double r) {
  if (l < 0 && r < 0) return 1;
  if (l < 0 && r >= 0) return l / (l - r);
  if (l >= 0 && r < 0) return r / (r - l);
  return 0;
}

__host__ __device__ inline void rotate4(double* a) {
  const double t = a[0];
  a[0] = a[1];
  a[1] = a[2];
  a[2] = a[3];
  a[3] = t;
}

// MathUtilities.h:301-375 (corners walked bl, br, tr, tl)
__host__ __device__ inline double fraction_inside4(double bl, double br,
                                                   double tl, double tr) {
  const int inside = (bl < 0) + (tl < 0) + (br < 0) + (tr < 0);
  double c[4] = {bl, br, tr, tl};
  if (inside == 4) return 1.0;
  if (inside == 3) {
    while (c[0] < 0) rotate4(c);
    const double s0 = 1.0 - fraction_inside2(c[0], c[3]);
    const double s1 = 1.0 - fraction_inside2(c[0], c[1]);
    return 1.0 - 0.5 * s0 * s1;
  }
  if (inside == 2) {
    while (c[0] >= 0 || !(c[1] < 0 || c[2] < 0)) rotate4(c);
    if (c[1] < 0) {
      const double sl = fraction_inside2(c[0], c[3]);
      const double sr = fraction_inside2(c[1], c[2]);
      return 0.5 * (sl + sr);
    }
    const double mid = 0.25 * (c[0] + c[1] + c[2] + c[3]);
    if (mid < 0) {
      double area = 0.0;
      const double s1 = 1.0 - fraction_inside2(c[0], c[3]);
      const double s3 = 1.0 - fraction_inside2(c[2], c[3]);
      area += 0.5 * s1 * s3;
      const double s2 = 1 - fraction_inside2(c[2], c[1]);
      const double s0 = 1 - fraction_inside2(c[0], c[1]);
      area += 0.5 * s0 * s2;
      return 1.0 - area;
    }
    double area = 0.0;
    const double s0 = fraction_inside2(c[0], c[1]);
    const double s1 = fraction_inside2(c[0], c[3]);
    area += 0.5 * s0 * s1;
    const double s2 = fraction_inside2(c[2], c[1]);
    const double s3 = fraction_inside2(c[2], c[3]);
    area += 0.5 * s2 * s3;
    return area;
  }
  if (inside == 1) {
    while (c[0] >= 0) rotate4(c);
    const double s0 = fraction_inside2(c[0], c[3]);
    const double s1 = fraction_inside2(c[0], c[1]);
    return 0.5 * s0 * s1;
  }
  return 0;
}

// DistanceFields.cpp:68-78 box_phi (zero corner radius) of the container
// [0, n*dx], negated (inside=1, DistanceFields.cpp:486): positive inside.
__host__ __device__ inline double container_phi(const Grid& g, double p0,
                                                double p1, double p2) {
  const double p[3] = {p0, p1, p2};
  double dd[3], da[3];
  for (int d = 0; d < 3; ++d) {
    const double lo = 0.0, hi = (double)g.n[d] * g.dx;
    const double c = 0.5 * (lo + hi);
    const double e = 0.5 * (hi - lo);
    dd[d] = fabs(p[d] - c) - e;
    da[d] = dd[d] > 0.0 ? dd[d] : 0.0;
  }
  double m = dd[0] > dd[1] ? dd[0] : dd[1];
  m = m > dd[2] ? m : dd[2];
  if (m > 0.0) m = 0.0;
  const double phi = m + sqrt(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]);
  return -phi;
}

// Three-node stencil base on one axis for a grid offset by off cells
// (TwoDScene.cpp:3700-3713), clamped into the dense grid.
__host__ __device__ inline int stencil_base(const Grid& g, int axis, double x,
                                            double off) {
  const double t = x / g.dx + ((double)kGhost - off);
  const double fl = floor(t);
  const double fr = t - fl;
  int b = (int)fl + (fr > 0.5 ? 0 : -1);
  if (b < 0) b = 0;
  if (b > g.nx[axis] - 3) b = g.nx[axis] - 3;
  return b;
}

__host__ __device__ inline void stencil_weights(const Grid& g, double x,
                                                int base, double off,
                                                double w[3]) {
  for (int a = 0; a < 3; ++a)
    w[a] = quad_kernel((x - node_pos(g, base + a, off)) / g.dx);
}

// grid offsets (in cells) of the u, v, w face grids (TwoDScene.cpp:5076-5081)
__host__ __device__ inline double face_off(int d, int a) {
  return (d == a) ? 0.0 : 0.5;
}

// The transfer helpers above in a chosen arithmetic type R (float for the
// mixed-precision form).  With R = double they perform exactly the
// operations of the double versions.
template <class R>
__host__ __device__ inline R node_pos_r(const Grid& g, int i, R off) {
  return ((R)(i - kGhost) + off) * (R)g.dx;
}
template <class R>
__host__ __device__ inline R quad_kernel_r(R x) {
  const R ax = x < (R)0 ? -x : x;
  if (ax < (R)0.5) return (R)0.75 - ax * ax;
  if (ax < (R)1.5) return (R)0.5 * ((R)1.5 - ax) * ((R)1.5 - ax);
  return (R)0;
}
template <class R>
__host__ __device__ inline int stencil_base_r(const Grid& g, int axis, R x, R off) {
  const R t = x / (R)g.dx + ((R)kGhost - off);
  const R fl = floor(t);
  const R fr = t - fl;
  int b = (int)fl + (fr > (R)0.5 ? 0 : -1);
  if (b < 0) b = 0;
  if (b > g.nx[axis] - 3) b = g.nx[axis] - 3;
  return b;
}
template <class R>
__host__ __device__ inline void stencil_weights_r(const Grid& g, R x, int base, R off,
                                                  R w[3]) {
  for (int a = 0; a < 3; ++a)
    w[a] = quad_kernel_r<R>((x - node_pos_r<R>(g, base + a, off)) / (R)g.dx);
}
template <class R>
__host__ __device__ inline R face_off_r(int d, int a) {
  return (d == a) ? (R)0 : (R)0.5;
}
template <class R>
__host__ __device__ inline R container_phi_r(const Grid& g, R p0, R p1, R p2) {
  const R p[3] = {p0, p1, p2};
  R dd[3], da[3];
  for (int d = 0; d < 3; ++d) {
    const R lo = (R)0, hi = (R)g.n[d] * (R)g.dx;
    const R c = (R)0.5 * (lo + hi);
    const R e = (R)0.5 * (hi - lo);
    const R q = p[d] - c;
    dd[d] = (q < (R)0 ? -q : q) - e;
    da[d] = dd[d] > (R)0 ? dd[d] : (R)0;
  }
  R m = dd[0] > dd[1] ? dd[0] : dd[1];
  m = m > dd[2] ? m : dd[2];
  if (m > (R)0) m = (R)0;
  const R phi = m + sqrt(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]);
  return -phi;
}

// The discrete decisions of a substep that the adjoint holds fixed, as the
// kernels see them: mode 0 computes every decision, 1 computes and records
// them, 2 replays recorded ones.  liquid: per node, the cell is liquid
// (phi < 0); flags: per node, bit d = face d has mass (M > 1e-20), bit 3 + d
// = theta of face d clamped at the theta criterion; proj: per caller index,
// the solid projection moved the particle; pid: the caller index of each
// particle slot; dep (mode 2): a count of the decisions the run would have
// taken differently from the record.
struct FrozenView {
  int mode = 0;
  uint8_t* liquid = nullptr;
  uint8_t* flags = nullptr;
  uint8_t* proj = nullptr;
  const uint32_t* pid = nullptr;
  unsigned long long* dep = nullptr;
};
__device__ inline void fz_depart(const FrozenView& F) {
  if (F.dep) atomicAdd(F.dep, 1ull);
}
__device__ inline bool fz_liquid(const FrozenView& F, const double* lphi, int64_t n) {
  return F.mode == 2 ? F.liquid[n] != 0 : lphi[n] < 0.0;
}
// A replayed face with mass that has lost all of it (a departure) is taken
// as massless rather than dividing by zero.
__device__ inline bool fz_mass(const FrozenView& F, int d, int64_t n, double M) {
  return F.mode == 2 ? ((F.flags[n] >> d) & 1) != 0 && M > 0.0 : M > 1e-20;
}
constexpr double kThetaCriterion = 0.01;  // Pressure.cpp:20
// theta of a face from the two cells' phis and liquid flags (fraction_inside,
// MathUtilities.h:286-293, clamped at the criterion); mode 2 takes the clamp
// decision from flags (and counts a departure if count is set), mode 1
// records it there
__device__ inline double fz_theta(const FrozenView& F, int d, int64_t fn, bool ll, bool lr,
                                  double phil, double phir, uint8_t* clamp_bits,
                                  bool count = false) {
  double th;
  if (ll && lr)
    th = 1.0;
  else if (ll)
    th = phil / (phil - phir);
  else if (lr)
    th = phir / (phir - phil);
  else
    th = 0.0;
  bool clamp;
  if (F.mode == 2) {
    clamp = ((F.flags[fn] >> (3 + d)) & 1) != 0;
    if (count && clamp != (th < kThetaCriterion)) fz_depart(F);
  } else {
    clamp = th < kThetaCriterion;
  }
  if (clamp_bits && clamp) *clamp_bits |= (uint8_t)(8u << d);
  return clamp ? kThetaCriterion : th;
}

// Counter-based RNG (splitmix64 finaliser chained over the counter words).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t hash3(uint64_t seed, uint64_t a,
                                          uint64_t b) {
  return mix64(mix64(mix64(seed) ^ a) ^ b);
}
__host__ __device__ inline double uniform01(uint64_t h) {
  return (double)(h >> 11) * 0x1.0p-53;
}

// Order-preserving map of a double onto uint64 (for atomic min of phi).
__device__ inline unsigned long long phi_key(double v) {
  unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
// phi_from_key on the host
inline double key_value(unsigned long long k) {
  const unsigned long long b =
      (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  double v;
  __builtin_memcpy(&v, &b, sizeof(v));
  return v;
}
__device__ inline double phi_from_key(unsigned long long k) {
  const unsigned long long b =
      (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __longlong_as_double((long long)b);
}

}  // namespace wc
