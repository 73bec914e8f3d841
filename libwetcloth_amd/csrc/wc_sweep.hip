// Pipelined MIC(0) triangular solves and the box-resident PCG vectors of a
// structured pressure system.
//
// solve_lower / solve_lower_transpose_in_place (PCGSolver.h:211-239) of the
// 7-point factor run as one persistent kernel per sweep instead of one launch
// per i+j+k hyperplane.  The DOF bounding box is cut into 16x16 tiles of
// k-pencils; a workgroup walks its tile's local hyperplanes i'+j'+k = w, one
// step per w, passing the newest value of each pencil to its +x/+y
// neighbours through LDS.  Values on a tile's low faces come from the
// upstream tiles (a-1, b) and (a, b-1), which run ahead.  The output array is
// filled with a NaN sentinel before the sweep and every cell is written
// exactly once with a single device-coherent (agent-scope atomic) store, so
// a consumer knows a value is final when it no longer reads the sentinel.
// A dedicated face wave per workgroup reads the upstream face values a batch
// of steps ahead and re-reads only those that were not final yet.  Tiles are
// claimed with a ticket counter in (a+b, a) order, so a tile only waits for
// tiles already claimed by running workgroups (no deadlock), and every wait
// is bounded.
//
// While these sweeps are in use the PCG vectors live in the same box layout
// (cells that are not DOFs hold zero), so the sweeps read and write them in
// place and the SpMV is a 7-point stencil without column indices.
//
// Every sum keeps the reference's order: a sweep cell applies its <= 3
// neighbour terms in ascending DOF order (the column loop's order) and a
// SpMV row adds its entries in ascending column order (lower neighbours,
// diagonal, upper neighbours), so results are bitwise those of the
// sequential solve and of multiply(FixedSparseMatrix, ...).
#include <math.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "wc_ctx.h"

namespace wc {
namespace {

constexpr int kT = 16;           // tile edge (pencils)
constexpr int kThreads = kT * kT;
constexpr int kChunk = 8;        // cell records prefetched ahead per pencil
constexpr long long kSpinLimit = 1ll << 22;  // ~seconds; a correct run waits us
constexpr unsigned long long kSentinel = 0xFFFFFFFFFFFFFFFFull;

// code bits per cell
constexpr uint32_t C_DOF = 1u;
constexpr int C_LOWER = 1;   // bits 1..3: coupling to -x, -y, -z
constexpr int C_UPPER = 4;   // bits 4..6: coupling to +x, +y, +z
constexpr int C_LORD = 7;    // bits 7..9: order of the lower terms (0..5)
constexpr int C_UORD = 10;   // bits 10..12: order of the upper terms
constexpr uint32_t C_DIAG = 1u << 13;  // the row has a diagonal entry
constexpr int C_ALO = 14;    // bits 14..16: A(me, -x/-y/-z) present
constexpr int C_AUP = 17;    // bits 17..19: A(me, +x/+y/+z) present

// SpMV flags per slot (16 bits): the row is a DOF, which of its entries
// below the diagonal exist (A(me, k_d) = A(k_d, me), read from the
// neighbour's +e_d plane), and the orders of its lower and upper terms
constexpr uint16_t M_DOF = 1u;
constexpr int M_LO = 1;     // bits 1..3: A(me, -x/-y/-z) present
constexpr int M_LORD = 4;   // bits 4..6: order of the lower terms
constexpr int M_UORD = 7;   // bits 7..9: order of the upper terms

// Sweep code word per slot, with each sweep's neighbour order applied: term
// q of a sweep is the q-th neighbour in ascending DOF order (the order the
// column loops of the reference apply them), present or not.
constexpr uint32_t S_DOF = 1u;
constexpr int S_FPRES = 1;   // bits 1..3: forward term q present
constexpr int S_FSRC = 4;    // bits 4..9: forward term q axis (0 x, 1 y, 2 z)
constexpr int S_BPRES = 10;  // bits 10..12: backward term q present
constexpr int S_BSRC = 13;   // bits 13..18: backward term q axis
constexpr int S_FX = 19, S_FY = 20;  // a forward term comes from -x / -y
constexpr int S_BX = 21, S_BY = 22;  // a backward term comes from +x / +y

// the six orders of three axes, 2 bits per axis, 6 bits per order:
// 0 xyz, 1 xzy, 2 yxz, 3 yzx, 4 zxy, 5 zyx
constexpr uint64_t kPerms =
    (uint64_t)(0 | 1 << 2 | 2 << 4) | (uint64_t)(0 | 2 << 2 | 1 << 4) << 6 |
    (uint64_t)(1 | 0 << 2 | 2 << 4) << 12 | (uint64_t)(1 | 2 << 2 | 0 << 4) << 18 |
    (uint64_t)(2 | 0 << 2 | 1 << 4) << 24 | (uint64_t)(2 | 1 << 2 | 0 << 4) << 30;
__host__ __device__ inline int perm_axis(int p, int q) {
  return (int)((kPerms >> (6 * p + 2 * q)) & 3);
}

__device__ inline int perm_of(int64_t k0, int64_t k1, int64_t k2) {
  // index of the permutation that sorts (k0, k1, k2) ascending
  if (k0 <= k1 && k1 <= k2) return 0;
  if (k0 <= k2 && k2 <= k1) return 1;
  if (k1 <= k0 && k0 <= k2) return 2;
  if (k1 <= k2 && k2 <= k0) return 3;
  if (k2 <= k0 && k0 <= k1) return 4;
  return 5;
}

__device__ inline double load_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_coherent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline float load_coherent(const float* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_coherent(float* p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// the all-ones NaN that marks an output not yet written
__device__ inline bool is_sentinel(double v) {
  return __double_as_longlong(v) == (long long)kSentinel;
}
__device__ inline bool is_sentinel(float v) { return __float_as_uint(v) == 0xFFFFFFFFu; }
__device__ inline int load_flag(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_flag(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// one box cell's factor entries for one sweep direction, in that sweep's
// term order (two 16-byte loads in FP64, one in FP32)
template <class V>
struct __attribute__((aligned(16))) CellT {
  V l0, l1, l2, inv;
};
using Cell = CellT<double>;
static_assert(sizeof(Cell) == 32, "Cell is two dwordx4");
static_assert(sizeof(CellT<float>) == 16, "an FP32 cell is one dwordx4");
constexpr int kCoef = 7;  // per box cell: A diagonal, A(-x,-y,-z), A(+x,+y,+z)

// one box cell's inputs of the factor pass (six 16-byte loads): the lower
// triangle of A as factor_modified_incomplete_cholesky0 copies it
// (PCGSolver.h:110-121), seen from this cell's row of L
struct __attribute__((aligned(16))) FacIn {
  double adiag;      // A(me, me), 0 if the row has no diagonal entry
  double lo[3];      // per forward term q: A(k, me) from the row of that lower neighbour k
  double oth[3][2];  // per term: row k's other entries above its diagonal, ascending column
  uint32_t fcode;    // F_* bits below (per term q)
  uint32_t code;     // the cell's S_* bits
};
static_assert(sizeof(FacIn) == 96, "FacIn is six dwordx4");
constexpr int F_KNULL = 0;  // bits 0..2: adiag(k) == 0 for term q (column k is skipped)
constexpr int F_NOTH = 3;   // bits 3..8: number of oth entries per term (2 bits each)

// Skewed tile layout of the box.  Cell (px, py, kk) belongs to tile
// (a, b) = (px / 16, py / 16) and lane l = (py % 16) * 16 + px % 16, and is
// stored at the tile's step w = kk + px % 16 + py % 16, the step at which the
// forward sweep computes it:
//     slot = ((b * ntile_a + a) * nsteps + w) * 256 + l,  nsteps = Bz + 30.
// At every step of a sweep (forward, or backward with w reversed) the 256
// lanes of a tile therefore touch 256 consecutive slots.  Slots whose kk lies
// outside [0, Bz) are padding and hold no DOF.
struct Skew {
  int ntile_a, ntile_b, nsteps, Bz;
  // slot offset from a lane on a tile's low x (y) face to the cell before it
  // in the tile below: (15, lj, kk) of tile a-1 (resp. (li, 15, kk) of b-1)
  int64_t dx, dy;
};

__host__ __device__ inline int64_t skew_slot(const Skew& K, int a, int b, int li, int lj,
                                             int w) {
  return (((int64_t)b * K.ntile_a + a) * K.nsteps + w) * kThreads + lj * kT + li;
}

struct BoxArgs {
  Grid g;
  int o[3];        // box origin (dense node coordinates)
  Skew K;
  int64_t nslot;
  const uint8_t* act;
  const uint32_t* isdof;
  const uint32_t* dofscan;
  const uint32_t* rowlen;  // ELL-7 rows of the assembled matrix
  const uint32_t* col;
  const double* invdiag;
  const double* lv;
  const double* val;
  int32_t* pdof;
  uint32_t* code;   // per-axis coupling bits and orders (C_*)
  uint32_t* scode;  // sweep code (S_*)
  uint32_t* opf;    // operand offsets of the forward / backward sweep
  uint32_t* opb;
  uint16_t* sflag;  // SpMV flags (M_*)
  Cell* cf;      // forward sweep (lower terms), zeroed here, filled by the factor pass
  Cell* cb;      // backward sweep (upper terms), likewise
  double* coef;  // kCoef planes of A entries: coef[e * nslot + cell]
  FacIn* fin;    // factor pass inputs
};

// Bounding box of the DOF nodes: each block strides over the nodes, reduces
// its threads' bounds through shuffles and LDS, and issues one set of six
// atomics (the grid is bounded, so a few thousand atomics in all).
constexpr int kBoxBlocks = 2048;
__global__ __launch_bounds__(256) void k_dof_box(Grid g, const uint32_t* isdof, int* bb) {
  __shared__ int sh[6][4];
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
  for (int64_t n = blockIdx.x * 256ll + threadIdx.x; n < g.nnode; n += (int64_t)gridDim.x * 256) {
    if (!isdof[n]) continue;
    int c[3];
    node_coords(g, n, c[0], c[1], c[2]);
    for (int d = 0; d < 3; ++d) {
      lo[d] = min(lo[d], c[d]);
      hi[d] = max(hi[d], c[d]);
    }
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int d = 0; d < 3; ++d) {
      lo[d] = min(lo[d], __shfl_xor(lo[d], o, 64));
      hi[d] = max(hi[d], __shfl_xor(hi[d], o, 64));
    }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; ++d) {
      sh[d][wv] = lo[d];
      sh[3 + d][wv] = hi[d];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; ++d) {
      int l = sh[d][0], h = sh[3 + d][0];
      for (int w = 1; w < 4; ++w) {
        l = min(l, sh[d][w]);
        h = max(h, sh[3 + d][w]);
      }
      if (h >= 0) {
        atomicMin(&bb[d], l);
        atomicMax(&bb[3 + d], h);
      }
    }
  }
}

// The factor's and the matrix's entries around each box cell, in pencil
// layout.  The row of a DOF holds its lower neighbours, its diagonal and its
// upper neighbours; L(me, lower) sits in the lower neighbour's row and
// L(upper, me) in this row (lv is indexed like val).
// The sweeps' exchange planes: 4 planes (step % 4) of 17 x 17 values; cell
// [1 + lj][1 + li] holds local pencil (li, lj), row 0 and column 0 hold the
// upstream y and x faces, and [0][0] holds +0.
constexpr int kPlane = (kT + 1) * (kT + 1);
constexpr int kPlanes = 4;

// One sweep's operands of a cell, for local pencil (li, lj) of that sweep:
// per ordered term q, a signed 10-bit offset from the pencil's own position
// in the previous step's plane to the term's operand: -1 for the -x
// neighbour (or x face), -(kT + 1) for the -y neighbour (or y face), 0 for
// the pencil's own previous value, and [0][0] (+0) for an absent term.
__host__ __device__ inline uint32_t operand_word(uint32_t sc, int presb, int srcb, int li,
                                                 int lj) {
  const int pos = (lj + 1) * (kT + 1) + li + 1;
  uint32_t word = 0;
  for (int q = 0; q < 3; ++q) {
    int d = -pos;
    if (sc & (1u << (presb + q))) {
      const uint32_t a = (sc >> (srcb + 2 * q)) & 3u;
      d = a == 0 ? -1 : (a == 1 ? -(kT + 1) : 0);
    }
    word |= ((uint32_t)d & 0x3FFu) << (10 * q);
  }
  return word;
}

// One DOF's row of the assembled matrix, summarised for the fill: written
// by k_row_sums (in node order, so the ELL rows are read coalesced) at the
// DOF's slot of the skewed layout, and read by k_fill_box at that slot and
// at the slots of the cell's six neighbours (fixed offsets within a tile).
// A slot without a DOF holds zeros (id1 = 0).
struct __attribute__((aligned(16))) RowSum {
  uint32_t id1;       // DOF id + 1, or 0
  uint32_t bits;      // R_DIAG, R_LO + d, R_UP + d
  double diag;        // A(me, me)
  double lo[3];       // A(me, me - e_d)
  double up[3];       // A(me, me + e_d)
  uint32_t upcol[3];  // the column (DOF id) of up[d]
  uint32_t pad;
};
static_assert(sizeof(RowSum) == 80, "RowSum is five dwordx4");
constexpr uint32_t R_DIAG = 1u;
constexpr int R_LO = 1;  // bits 1..3: the row holds A(me, me - e_d)
constexpr int R_UP = 4;  // bits 4..6: the row holds A(me, me + e_d)

__global__ __launch_bounds__(256) void k_row_sums(BoxArgs A, RowSum* S) {
  const Grid& g = A.g;
  for (int64_t n = blockIdx.x * 256ll + threadIdx.x; n < g.nnode; n += (int64_t)gridDim.x * 256) {
    if (!A.isdof[n]) continue;
    int c[3];
    node_coords(g, n, c[0], c[1], c[2]);
    const int x = c[0] - A.o[0], y = c[1] - A.o[1], z = c[2] - A.o[2];
    const int li = x % kT, lj = y % kT;
    const int64_t t = skew_slot(A.K, x / kT, y / kT, li, lj, z + li + lj);
    const uint32_t me = A.dofscan[n];
    const int64_t row = 7 * (int64_t)me;
    const uint32_t len = A.rowlen[me];
    uint32_t mc[7];
    double mv[7];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      mc[s] = (uint32_t)s < len ? A.col[row + s] : 0xFFFFFFFFu;
      mv[s] = A.val[row + s];
    }
    RowSum r;
    r.id1 = me + 1;
    r.bits = 0;
    r.diag = 0;
    r.pad = 0;
#pragma unroll
    for (int s = 0; s < 7; ++s)
      if (mc[s] == me) {
        r.bits |= R_DIAG;
        r.diag = mv[s];
      }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      r.lo[d] = r.up[d] = 0;
      r.upcol[d] = 0;
#pragma unroll
      for (int dir = 0; dir < 2; ++dir) {
        int q[3] = {c[0], c[1], c[2]};
        q[d] += dir ? 1 : -1;
        if (!in_range(g, q[0], q[1], q[2])) continue;
        const int64_t nn = node_index(g, q[0], q[1], q[2]);
        if (!A.isdof[nn]) continue;
        const uint32_t other = A.dofscan[nn];
#pragma unroll
        for (int s = 0; s < 7; ++s)
          if (mc[s] == other) {
            if (dir) {
              r.up[d] = mv[s];
              r.upcol[d] = other;
              r.bits |= 1u << (R_UP + d);
            } else {
              r.lo[d] = mv[s];
              r.bits |= 1u << (R_LO + d);
            }
          }
      }
    }
    S[t] = r;
  }
}

__global__ void k_fill_box(BoxArgs A, const RowSum* S) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= A.nslot) return;
  const int lane = (int)(t % kThreads);
  const int64_t q = t / kThreads;
  const int w = (int)(q % A.K.nsteps);
  const int tile = (int)(q / A.K.nsteps);
  const int li = lane % kT, lj = lane / kT;
  const int z = w - li - lj;
  const int ta = tile % A.K.ntile_a, tb = tile / A.K.ntile_a;
  uint32_t code = 0, fcode = 0;
  int32_t me = -1;
  double cf[kCoef] = {0, 0, 0, 0, 0, 0, 0};
  FacIn f;
  f.adiag = 0;
  for (int d = 0; d < 3; ++d) f.lo[d] = f.oth[d][0] = f.oth[d][1] = 0;
  const RowSum own = S[t];  // padding slots hold zeros
  if (own.id1) {
    me = (int32_t)(own.id1 - 1);
    code |= C_DOF;
    if (own.bits & R_DIAG) {
      code |= C_DIAG;
      cf[0] = own.diag;
      f.adiag = own.diag;
    }
    int64_t lkey[3], ukey[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lkey[d] = ukey[d] = 0x7FFFFFFFFFFFll;
      // the slots of the -e_d and +e_d neighbours (-1 outside the box, where
      // no cell is a DOF): the previous / next step of the lane one to the
      // left (below), or across the tile face
      int64_t tl = -1, tu = -1;
      if (d == 0) {
        tl = li > 0 ? t - kThreads - 1 : (ta > 0 ? t + A.K.dx : -1);
        tu = li < kT - 1 ? t + kThreads + 1 : (ta + 1 < A.K.ntile_a ? t - A.K.dx : -1);
      } else if (d == 1) {
        tl = lj > 0 ? t - kThreads - kT : (tb > 0 ? t + A.K.dy : -1);
        tu = lj < kT - 1 ? t + kThreads + kT : (tb + 1 < A.K.ntile_b ? t - A.K.dy : -1);
      } else {
        tl = z > 0 ? t - kThreads : -1;
        tu = z + 1 < A.K.Bz ? t + kThreads : -1;
      }
      // lower neighbour k_d: its row holds (k_d, me) -> column k_d of L;
      // this row holds A(me, k_d) for the SpMV
      if (tl >= 0) {
        const RowSum L = S[tl];
        if (L.id1) {
          const uint32_t other = L.id1 - 1;
          lkey[d] = other;
          if (L.bits & (1u << (R_UP + d))) {  // its +e_d entry is (k_d, me)
            f.lo[d] = L.up[d];
            code |= 1u << (C_LOWER + d);
          }
          const double kdiag = (L.bits & R_DIAG) ? L.diag : 0.0;
          // row k_d's other entries above its diagonal, ascending column
          uint32_t oc[2];
          double ov[2];
          int noth = 0;
#pragma unroll
          for (int e = 0; e < 3; ++e)
            if (e != d && (L.bits & (1u << (R_UP + e)))) {
              oc[noth] = L.upcol[e];
              ov[noth] = L.up[e];
              ++noth;
            }
          if (noth == 2 && oc[1] < oc[0]) {
            f.oth[d][0] = ov[1];
            f.oth[d][1] = ov[0];
          } else {
            if (noth > 0) f.oth[d][0] = ov[0];
            if (noth > 1) f.oth[d][1] = ov[1];
          }
          if (kdiag == 0) fcode |= 1u << (F_KNULL + d);
          fcode |= (uint32_t)noth << (F_NOTH + 2 * d);
          if (own.bits & (1u << (R_LO + d))) {
            cf[1 + d] = own.lo[d];
            code |= 1u << (C_ALO + d);
          }
        }
      }
      // upper neighbour: this row holds A(me, other); L(other, me) is
      // written here by the factor pass of the neighbour
      if (tu >= 0) {
        const uint32_t uid1 = S[tu].id1;
        if (uid1) {
          ukey[d] = uid1 - 1;
          if (own.bits & (1u << (R_UP + d))) {
            cf[4 + d] = own.up[d];
            code |= (1u << (C_UPPER + d)) | (1u << (C_AUP + d));
          }
        }
      }
    }
    code |= (uint32_t)perm_of(lkey[0], lkey[1], lkey[2]) << C_LORD;
    code |= (uint32_t)perm_of(ukey[0], ukey[1], ukey[2]) << C_UORD;
  }
  // the sweeps' and the factor's term order
  uint32_t sc = 0;
  FacIn fo;
  fo.adiag = f.adiag;
  uint32_t fco = 0;
  for (int q = 0; q < 3; ++q) fo.lo[q] = fo.oth[q][0] = fo.oth[q][1] = 0;
  if (code & C_DOF) {
    sc |= S_DOF;
    const int lord = (code >> C_LORD) & 7, uord = (code >> C_UORD) & 7;
    for (int q = 0; q < 3; ++q) {
      const int ax = perm_axis(lord, q);
      sc |= (uint32_t)ax << (S_FSRC + 2 * q);
      if (code & (1u << (C_LOWER + ax))) sc |= 1u << (S_FPRES + q);
      fo.lo[q] = f.lo[ax];
      fo.oth[q][0] = f.oth[ax][0];
      fo.oth[q][1] = f.oth[ax][1];
      if (fcode & (1u << (F_KNULL + ax))) fco |= 1u << (F_KNULL + q);
      fco |= ((fcode >> (F_NOTH + 2 * ax)) & 3u) << (F_NOTH + 2 * q);
      const int bx = perm_axis(uord, q);
      sc |= (uint32_t)bx << (S_BSRC + 2 * q);
      if (code & (1u << (C_UPPER + bx))) sc |= 1u << (S_BPRES + q);
    }
    if (code & (1u << (C_LOWER + 0))) sc |= 1u << S_FX;
    if (code & (1u << (C_LOWER + 1))) sc |= 1u << S_FY;
    if (code & (1u << (C_UPPER + 0))) sc |= 1u << S_BX;
    if (code & (1u << (C_UPPER + 1))) sc |= 1u << S_BY;
  }
  A.pdof[t] = me;
  A.code[t] = code;
  A.scode[t] = sc;
  uint16_t mf = 0;
  if (code & C_DOF)
    mf = (uint16_t)(M_DOF | (((code >> C_ALO) & 7u) << M_LO) |
                    (((code >> C_LORD) & 7u) << M_LORD) | (((code >> C_UORD) & 7u) << M_UORD));
  A.sflag[t] = mf;
  A.opf[t] = operand_word(sc, S_FPRES, S_FSRC, li, lj);
  A.opb[t] = operand_word(sc, S_BPRES, S_BSRC, kT - 1 - li, kT - 1 - lj);
  fo.fcode = fco;
  fo.code = sc;
  A.fin[t] = fo;
  const Cell zero = {0, 0, 0, 0};
  A.cf[t] = zero;
  A.cb[t] = zero;
  for (int e = 0; e < kCoef; ++e) A.coef[e * A.nslot + t] = cf[e];
}

// DOF-ordered vector -> box (zero where the cell is not a DOF)
__global__ void k_box_gather(int64_t ncell, const int32_t* pdof, const double* src,
                             double* dst) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  const int32_t d = pdof[t];
  dst[t] = d >= 0 ? src[d] : 0.0;
}

// box -> DOF-ordered vector
__global__ void k_box_scatter(int64_t ncell, const int32_t* pdof, const double* src,
                              double* dst) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  const int32_t d = pdof[t];
  if (d >= 0) dst[d] = src[t];
}

// The slots of the six neighbours of slot t (valid where the coupling to that
// neighbour is present, which implies the neighbour is a DOF in the box).
struct Nbr {
  int64_t lo[3], up[3];
};
__device__ inline Nbr neighbours(const Skew& K, int64_t t) {
  const int lane = (int)(t % kThreads);
  const int li = lane % kT, lj = lane / kT;
  Nbr n;
  n.lo[0] = li > 0 ? t - kThreads - 1 : t + K.dx;
  n.up[0] = li < kT - 1 ? t + kThreads + 1 : t - K.dx;
  n.lo[1] = lj > 0 ? t - kThreads - kT : t + K.dy;
  n.up[1] = lj < kT - 1 ? t + kThreads + kT : t - K.dy;
  n.lo[2] = t - kThreads;
  n.up[2] = t + kThreads;
  return n;
}

// Row t of A x on box vectors: multiply(FixedSparseMatrix, x, y)
// (SparseMatrix.h:298-322) with the row's entries added in ascending column
// order, in FP64 whatever the storage type of A (CV) and x (V).  The entries
// are stored in kCoef planes of nslot values, and an absent entry (and every
// entry of a cell that is not a DOF) is stored as +0.  So the row adds all
// seven products in order: an absent entry's product is +-0, and adding it
// is an exact identity (a sum that starts at +0 never becomes -0).  Every
// load is then independent of the code word, which only orders the terms;
// the x index of an absent neighbour is clamped into the box.  Slots are
// addressed with 32-bit indices (every plan holds fewer than 2^32 slots).
template <class CV, class V>
__device__ inline double box_row(const Skew& K, uint32_t t, uint32_t c,
                                 const CV* __restrict__ coef, uint32_t nslot,
                                 const V* __restrict__ x) {
  const uint32_t lane = t & (kThreads - 1);
  const uint32_t li = lane & (kT - 1), lj = lane / kT;
  const uint32_t dx = (uint32_t)(int32_t)K.dx, dy = (uint32_t)(int32_t)K.dy;
  const uint32_t last = nslot - 1;
  auto clampn = [&](uint32_t n) { return n < last ? n : last; };
  const uint32_t n_lo[3] = {clampn(li > 0 ? t - kThreads - 1 : t + dx),
                            clampn(lj > 0 ? t - kThreads - kT : t + dy),
                            clampn(t - kThreads)};
  const uint32_t n_up[3] = {clampn(li < kT - 1 ? t + kThreads + 1 : t - dx),
                            clampn(lj < kT - 1 ? t + kThreads + kT : t - dy),
                            clampn(t + kThreads)};
  double a[kCoef], xv[kCoef];
#pragma unroll
  for (int e = 0; e < kCoef; ++e) a[e] = (double)coef[(uint32_t)e * nslot + t];
  xv[0] = (double)x[t];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    xv[1 + d] = (double)x[n_lo[d]];
    xv[4 + d] = (double)x[n_up[d]];
  }
  const int lor = (int)((c >> C_LORD) & 7), upr = (int)((c >> C_UORD) & 7);
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int ax = perm_axis(lor, q);
    const double ca = ax == 0 ? a[1] : (ax == 1 ? a[2] : a[3]);
    const double cx = ax == 0 ? xv[1] : (ax == 1 ? xv[2] : xv[3]);
    acc += ca * cx;
  }
  acc += a[0] * xv[0];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int ax = perm_axis(upr, q);
    const double ca = ax == 0 ? a[4] : (ax == 1 ? a[5] : a[6]);
    const double cx = ax == 0 ? xv[4] : (ax == 1 ? xv[5] : xv[6]);
    acc += ca * cx;
  }
  return acc;
}

// A row's products in ascending column order: the lower terms in the row's
// lower order, the diagonal, the upper terms in its upper order (FP64);
// +0 for a slot that is not a DOF.  (Entries and values by axis, as scalars:
// arrays picked by a computed axis end up in scratch memory.)
struct RowTerms {
  double al0, al1, al2, xl0, xl1, xl2;  // -x, -y, -z neighbours
  double a0, x0;                        // diagonal
  double au0, au1, au2, xu0, xu1, xu2;  // +x, +y, +z neighbours
};
__device__ inline double pick3(int ax, double v0, double v1, double v2) {
  return ax == 0 ? v0 : (ax == 1 ? v1 : v2);
}
__device__ inline double row_sum(uint32_t f, const RowTerms& R) {
  const int lor = (int)((f >> M_LORD) & 7), upr = (int)((f >> M_UORD) & 7);
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int ax = perm_axis(lor, q);
    acc += pick3(ax, R.al0, R.al1, R.al2) * pick3(ax, R.xl0, R.xl1, R.xl2);
  }
  acc += R.a0 * R.x0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int ax = perm_axis(upr, q);
    acc += pick3(ax, R.au0, R.au1, R.au2) * pick3(ax, R.xu0, R.xu1, R.xu2);
  }
  return (f & M_DOF) ? acc : 0.0;
}

// y = A x (FP64, or FP32 storage with FP64 row sums).  With DOT, also the
// FP64 block partials of x . y, with the thread-to-slot mapping and the
// reduction tree of the solver's dot-product kernel, so the sum is the one
// that kernel would compute from the stored y.
// The same row from four planes: the diagonal and the +x/+y/+z entries; each
// entry below the diagonal, A(me, k_d), is the lower neighbour's +e_d entry
// A(k_d, me) (the matrix is symmetric bit for bit: both rows take the term of
// their shared face).  The products are added in the same order, so y is the
// seven-plane row's.  The loads do not branch: a slot that is not a DOF, or
// an entry below the diagonal that does not exist, reads slot 0 instead (one
// cached line for the whole launch) and its product is replaced by +0, which
// leaves the sum as skipping it would (a sum that starts at +0 never
// becomes -0); a slot that is not a DOF gives +0.  xself: x at the slot (+0
// where it is not a DOF, as the vectors hold there), for the dot.
template <class V, class CV>
__device__ inline double box_row_sym(const Skew& K, uint32_t t, uint32_t f,
                                     const CV* __restrict__ coef, uint32_t nslot,
                                     const V* __restrict__ x, double& xself) {
  const bool dof = (f & M_DOF) != 0;
  const uint32_t lane = t & (kThreads - 1);
  const uint32_t li = lane & (kT - 1), lj = lane / kT;
  const uint32_t dx = (uint32_t)(int32_t)K.dx, dy = (uint32_t)(int32_t)K.dy;
  const uint32_t last = nslot - 1;
  auto clampn = [&](uint32_t n) { return n < last ? n : last; };
  const uint32_t self = dof ? t : 0u;
  // lower neighbours: where they exist (flags) they are DOFs inside the box
  const uint32_t lo_raw[3] = {li > 0 ? t - kThreads - 1 : t + dx,
                              lj > 0 ? t - kThreads - kT : t + dy, t - kThreads};
  const uint32_t up_raw[3] = {li < kT - 1 ? t + kThreads + 1 : t - dx,
                              lj < kT - 1 ? t + kThreads + kT : t - dy, t + kThreads};
  double alo[3], xlo[3], aup[3], xup[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const bool lo = dof && ((f >> (M_LO + d)) & 1u);
    const uint32_t nl = lo ? lo_raw[d] : 0u;
    const uint32_t nu = dof ? clampn(up_raw[d]) : 0u;
    const double cl = (double)coef[(uint32_t)(4 + d) * nslot + nl];
    const double vl = (double)x[nl];
    alo[d] = lo ? cl : 0.0;
    xlo[d] = lo ? vl : 0.0;
    aup[d] = (double)coef[(uint32_t)(4 + d) * nslot + self];
    xup[d] = (double)x[nu];
  }
  const double a0 = (double)coef[self], x0 = (double)x[self];
  xself = dof ? x0 : 0.0;
  const RowTerms T{alo[0], alo[1], alo[2], xlo[0], xlo[1], xlo[2], a0, x0,
                   aup[0], aup[1], aup[2], xup[0], xup[1], xup[2]};
  return row_sum(f, T);
}

// kSpmvThreads per block (twice the reduction kernels' 256: the row's loads
// are latency-bound, so more waves in flight), one dot partial per block.
constexpr int kSpmvThreads = 512;
template <class V, bool DOT>
__global__ __launch_bounds__(kSpmvThreads) void k_box_spmv_sym(uint32_t nslot, Skew K,
                                                      const uint16_t* __restrict__ sflag,
                                                      const V* __restrict__ coef,
                                                      const V* __restrict__ x,
                                                      V* __restrict__ y, double* partial) {
  __shared__ double sh[kSpmvThreads];
  double acc = 0.0;
  for (uint32_t t = blockIdx.x * (uint32_t)kSpmvThreads + threadIdx.x; t < nslot;
       t += gridDim.x * (uint32_t)kSpmvThreads) {
    double xs;
    const V v = (V)box_row_sym(K, t, (uint32_t)sflag[t], coef, nslot, x, xs);
    if (DOT) acc += xs * (double)v;
    y[t] = v;
  }
  if (DOT) {
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kSpmvThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
  }
}

// A tile's x faces as compact columns: f15[t >> 4] = x[t] at the slots
// with li = 15 and f0[t >> 4] = x[t] at li = 0, a step's 16 values of each
// column side by side.  The -x neighbour of lane (li = 0, lj) at step w is
// slot (tile - 1, w + 15, li = 15, lj), so its column entry is (t + dx) >> 4
// (the clamp keeps li = 15); the +x neighbour of (li = 15, lj) is
// (tile + 1, w - 15, li = 0, lj), entry (t - dx) >> 4.  A column is spread
// over every cache line of its step; the compact copy is one line.
template <class V>
__global__ void k_face_cols(int64_t n16, const V* __restrict__ x, V* __restrict__ f15,
                            V* __restrict__ f0) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x) {
    f15[i] = x[i * 16 + 15];
    if (f0) f0[i] = x[i * 16];
  }
}

template <class V>
static void face_cols(hipStream_t st, int64_t ncell, const V* x, V* f15, V* f0) {
  const int64_t n16 = ncell / 16;
  const int64_t nb = std::min<int64_t>(4096, (n16 + 255) / 256);
  hipLaunchKernelGGL(k_face_cols<V>, dim3((unsigned)nb), dim3(256), 0, st, n16, x, f15, f0);
  WC_LAUNCH_CHECK();
}

// The same product through a window: one 256-thread block per (tile, run
// of consecutive steps), walking the run's steps.  Every neighbour of a slot
// in the tile lies in the step before or after it (lane -1 / -16 / same, or
// +1 / +16 / same), so the block keeps x of steps w - 1 .. w + 1 and the
// three upper coefficient planes of step w - 1 (which hold the lower
// couplings of step w) in LDS rings, and reads from global memory only its
// own slots of each step and, on the tile's faces, the neighbours in the
// next tile (the x faces from the compact columns above).  Each slot's
// values are read once per run instead of by up to seven rows.  A step's
// loads are issued one step ahead, into registers.
// The products are row_sum's, so y is k_box_spmv_sym's bit for bit; with
// DOT, one partial of x . y per block (blocks in (tile, run) order), summed
// by wave shuffles.
template <class V>
struct WinRow {
  uint16_t f;
  V a0, up[3], xn;          // diagonal, upper planes; x of the next step
  V cl[2], xl[2], xu[2];    // face neighbours in the next tile (x, y)
};

template <class V, bool DOT>
__global__ __launch_bounds__(256) void k_box_spmv_win(uint32_t nslot, Skew K, int run, int nrun,
                                                     const uint16_t* __restrict__ sflag,
                                                     const V* __restrict__ coef,
                                                     const V* __restrict__ x,
                                                     const V* __restrict__ f15,
                                                     const V* __restrict__ f0,
                                                     const V* __restrict__ c15,
                                                     V* __restrict__ y, double* partial) {
  __shared__ V xs[4][kThreads];     // x of step w at [w & 3]
  __shared__ V us[3][3][kThreads];  // upper planes of step w at [w % 3]
  __shared__ double sh[kThreads / 64];  // the waves' x . y sums
  static_assert(kThreads / 64 == 4, "four waves per block");
  const int lane = threadIdx.x, li = lane & (kT - 1), lj = lane / kT;
  const int nsteps = K.nsteps;
  const int tile = blockIdx.x / nrun;
  const int w0 = (blockIdx.x % nrun) * run;
  const int w1 = w0 + run < nsteps ? w0 + run : nsteps;
  const uint32_t tb = (uint32_t)tile * (uint32_t)nsteps * kThreads + (uint32_t)lane;
  const uint32_t dx = (uint32_t)(int32_t)K.dx, dy = (uint32_t)(int32_t)K.dy;
  const uint32_t last = nslot - 1;
  auto clampn = [&](uint32_t n) { return n < last ? n : last; };
  auto ldx = [&](int w) -> V { return w >= 0 && w < nsteps ? x[tb + (uint32_t)w * kThreads] : (V)0; };
  // the loads of step w (w0 <= w < w1); the face slots are clamped into the
  // box and used only where the neighbour exists (then they are in it); the
  // x faces come from the compact columns of x and of the -x coupling plane
  auto fetch = [&](int w, WinRow<V>& R) {
    const uint32_t t = tb + (uint32_t)w * kThreads;
    R.f = sflag[t];
    R.a0 = coef[t];
#pragma unroll
    for (int d = 0; d < 3; ++d) R.up[d] = coef[(uint32_t)(4 + d) * nslot + t];
    R.xn = ldx(w + 1);
    if (li == 0) {
      const uint32_t n = clampn(t + dx) >> 4;
      R.cl[0] = c15[n];
      R.xl[0] = f15[n];
    }
    if (lj == 0) {
      const uint32_t n = clampn(t + dy);
      R.cl[1] = coef[5u * nslot + n];
      R.xl[1] = x[n];
    }
    if (li == kT - 1) R.xu[0] = f0[clampn(t - dx) >> 4];
    if (lj == kT - 1) R.xu[1] = x[clampn(t - dy)];
  };
  xs[(w0 + 3) & 3][lane] = ldx(w0 - 1);
  xs[w0 & 3][lane] = ldx(w0);
#pragma unroll
  for (int d = 0; d < 3; ++d)
    us[(w0 + 2) % 3][d][lane] =
        w0 > 0 ? coef[(uint32_t)(4 + d) * nslot + tb + (uint32_t)(w0 - 1) * kThreads] : (V)0;
  WinRow<V> R;
  R.cl[0] = R.cl[1] = R.xl[0] = R.xl[1] = R.xu[0] = R.xu[1] = (V)0;
  fetch(w0, R);
  double acc = 0.0;
  for (int w = w0; w < w1; ++w) {
    const WinRow<V> C = R;
#pragma unroll
    for (int d = 0; d < 3; ++d) us[w % 3][d][lane] = C.up[d];
    xs[(w + 1) & 3][lane] = C.xn;
    if (w + 1 < w1) fetch(w + 1, R);
    __syncthreads();  // step w + 1's x and step w's planes are in place
    const uint32_t f = C.f;
    const bool dof = (f & M_DOF) != 0;
    const int pm = (w + 2) % 3, xm = (w + 3) & 3, xp = (w + 1) & 3;
    const bool lox = dof && ((f >> M_LO) & 1u), loy = dof && ((f >> (M_LO + 1)) & 1u),
               loz = dof && ((f >> (M_LO + 2)) & 1u);
    RowTerms T;
    T.al0 = lox ? (double)(li > 0 ? us[pm][0][lane - 1] : C.cl[0]) : 0.0;
    T.xl0 = lox ? (double)(li > 0 ? xs[xm][lane - 1] : C.xl[0]) : 0.0;
    T.al1 = loy ? (double)(lj > 0 ? us[pm][1][lane - kT] : C.cl[1]) : 0.0;
    T.xl1 = loy ? (double)(lj > 0 ? xs[xm][lane - kT] : C.xl[1]) : 0.0;
    T.al2 = loz ? (double)us[pm][2][lane] : 0.0;
    T.xl2 = loz ? (double)xs[xm][lane] : 0.0;
    T.a0 = (double)C.a0;
    T.x0 = (double)xs[w & 3][lane];
    T.au0 = (double)C.up[0];
    T.au1 = (double)C.up[1];
    T.au2 = (double)C.up[2];
    T.xu0 = (double)(li < kT - 1 ? xs[xp][lane + 1] : C.xu[0]);
    T.xu1 = (double)(lj < kT - 1 ? xs[xp][lane + kT] : C.xu[1]);
    T.xu2 = (double)xs[xp][lane];
    const V v = (V)row_sum(f, T);
    if (DOT) acc += (dof ? T.x0 : 0.0) * (double)v;
    y[tb + (uint32_t)w * kThreads] = v;
  }
  if (DOT) {
    // a butterfly over the wave's 64 lanes, then the 4 waves in order
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((lane & 63) == 0) sh[lane >> 6] = acc;
    __syncthreads();
    if (lane == 0) partial[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  }
}

// The refinement residual of the mixed-precision solve, r = b - A x in FP64
// (FP64 A, b and x), stored rounded to FP32, with |r|_inf block partials.
__global__ __launch_bounds__(256) void k_box_residual(uint32_t nslot, Skew K,
                                                      const uint32_t* code,
                                                      const double* coef, const double* b,
                                                      const double* x, float* r,
                                                      double* partial) {
  __shared__ double sh[256];
  double m = 0.0;
  for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < nslot; t += gridDim.x * 256u) {
    const double v = b[t] - box_row(K, t, code[t], coef, nslot, x);
    r[t] = (float)v;
    m = fmax(m, fabs(v));
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// FP64 -> FP32 copies of the factor records and matrix entries
__global__ void k_narrow_box(const double* a, float* b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// The slot of the upper neighbour of slot t along axis d.
__device__ inline int64_t upper_slot(const Skew& K, int64_t t, int d) {
  const int lane = (int)(t % kThreads);
  const int li = lane % kT, lj = lane / kT;
  if (d == 0) return li < kT - 1 ? t + kThreads + 1 : t - K.dx;
  if (d == 1) return lj < kT - 1 ? t + kThreads + kT : t - K.dy;
  return t + kThreads;
}

// The backward sweep's record of cell k, in its term order: L(u, k) for the
// upper neighbour u of each present term, read from u's forward record at
// the position of u's term that points back to k, and invdiag(k).
__global__ void k_backward_records(int64_t nslot, Skew K, const uint32_t* scode,
                                   const Cell* cf, Cell* cb) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nslot) return;
  const uint32_t c = scode[t];
  double v[3] = {0, 0, 0};
  Cell r = {0, 0, 0, 0};
  if (c & S_DOF) {
    for (int q = 0; q < 3; ++q) {
      if (!(c & (1u << (S_BPRES + q)))) continue;
      const int d = (c >> (S_BSRC + 2 * q)) & 3;
      const int64_t u = upper_slot(K, t, d);
      const uint32_t uc = scode[u];
      const int qq = (int)((uc >> S_FSRC) & 3u) == d ? 0 : ((int)((uc >> (S_FSRC + 2)) & 3u) == d ? 1 : 2);
      v[q] = reinterpret_cast<const double*>(&cf[u])[qq];
    }
    r.l0 = v[0];
    r.l1 = v[1];
    r.l2 = v[2];
    r.inv = cf[t].inv;
  }
  cb[t] = r;
}

enum { K_FWD = 0, K_BWD = 1, K_FAC = 2 };

// V is the sweep's value type: double, or float for the FP32 sweeps of the
// mixed-precision solve (the factor pass is always FP64).
template <class V>
struct SweepArgs {
  Skew K;
  int ntiles;
  const uint32_t* order;   // ticket -> a | b << 16
  const CellT<V>* cells;   // forward: lower terms; backward: upper terms
  const uint32_t* code;
  const V* in;             // forward: r; backward: the forward result
  V* out;                  // the sweep result, or the factor's invdiag
  V* reset;                // forward sweep: set to the sentinel step by step
  const V* dot_r;          // backward sweep: r (the forward sweep's input) ...
  double* dot_out;         // ... and its per-tile partials of z . r
  int* ticket;
  int* next_ticket;        // sweeps: zeroed by the block that takes ticket 0
  int* abort_flag;
  // factor pass
  const FacIn* fin;
  Cell* cf;
  double tau, mdr;
};

// inputs of one cell of a sweep
template <class V>
struct CellIn {
  CellT<V> c;
  V rhs;
  uint32_t code;
};

template <int KIND, class V>
struct PassIn {
  using T = CellIn<V>;
  static constexpr int chunk = kChunk;  // steps prefetched ahead
};
template <class V>
struct PassIn<K_FAC, V> {
  using T = FacIn;
  static constexpr int chunk = 4;
};

// inputs of the kk-th cell of a pencil; kk is clamped into the pencil so the
// loads are always issued to a valid address (steps outside it discard them)
// base[i] addressed as a 32-bit byte offset from a uniform base (the plan
// keeps every array below 4 GiB), so no 64-bit address arithmetic is needed
template <class T>
__device__ inline const T* at(const T* base, uint32_t i) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                    i * (uint32_t)sizeof(T));
}
template <class T>
__device__ inline T* at(T* base, uint32_t i) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(T));
}

template <int KIND, class V>
__device__ inline typename PassIn<KIND, V>::T load_in(const SweepArgs<V>& S, uint32_t idx) {
  if constexpr (KIND == K_FAC) {
    return *at(S.fin, idx);
  } else {
    CellIn<V> v;
    v.c = *at(S.cells, idx);
    v.rhs = *at(S.in, idx);
    v.code = *at(S.code, idx);
    return v;
  }
}
template <class V>
__device__ inline uint32_t code_of(const CellIn<V>& v) {
  return v.code;
}
__device__ inline uint32_t code_of(const FacIn& v) { return v.code; }

// the value at p once it is no longer the sentinel (rare path: the
// prefetched copy was read before its tile had written it)
template <class V>
__device__ __attribute__((always_inline)) inline V wait_value(const V* p, int* abort_flag) {
  long long spins = 0;
  V v = load_coherent(p);
  while (is_sentinel(v)) {
    if (++spins > kSpinLimit || load_flag(abort_flag)) {
      store_flag(abort_flag, 1);
      return (V)0;
    }
    __builtin_amdgcn_s_sleep(1);
    v = load_coherent(p);
  }
  return v;
}

// solve_lower / solve_lower_transpose_in_place for one cell: the rhs minus
// the present neighbour terms in ascending DOF order, times invdiag.  v_q is
// the neighbour value of term q in that order, or +0 for an absent term,
// whose factor entry is +0 as well (records are zeroed and only present
// terms are written), so an absent term subtracts +0 * +0 = +0, an exact
// identity.  Cells that are not DOFs carry rhs = inv = 0 and no terms, so
// their value is 0.
template <class V>
__device__ inline V sweep_cell_ordered(const CellIn<V>& ci, V v0, V v1, V v2) {
  V acc = ci.rhs;
  acc -= ci.c.l0 * v0;
  acc -= ci.c.l1 * v1;
  acc -= ci.c.l2 * v2;
  return acc * ci.c.inv;
}

// factor_modified_incomplete_cholesky0 (PCGSolver.h:136-205) for one cell of
// a 7-point system, where no update ever fills in: each lower neighbour
// column k (ascending DOF order, skipped when adiag(k) == 0) scales
// L(me, k) = A(k, me) invdiag(k) and modifies this diagonal by -L^2 and by
// -(tau L) missing, missing being column k's other scaled entries summed in
// ascending row order.  Returns invdiag(me) and writes the scaled entries
// into this cell's forward record (the backward records are gathered from
// the forward ones after the pass, so every record has a single writer).
__device__ inline double factor_cell(const SweepArgs<double>& S, const FacIn& f, uint32_t code,
                                     double ix, double iy, double iprev, uint32_t idx) {
  double L0 = 0, L1 = 0, L2 = 0;
  double d = f.adiag;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!(code & (1u << (S_FPRES + q)))) continue;
    const uint32_t ax = (code >> (S_FSRC + 2 * q)) & 3u;
    const double ik = ax == 0 ? ix : (ax == 1 ? iy : iprev);
    const double a = f.lo[q];
    double m = a;  // column k skipped (adiag(k) == 0): its entries stay unscaled
    if (!(f.fcode & (1u << (F_KNULL + q)))) {
      m = a * ik;
      const int noth = (f.fcode >> (F_NOTH + 2 * q)) & 3;
      double missing = 0;
      if (noth > 0) missing += f.oth[q][0] * ik;
      if (noth > 1) missing += f.oth[q][1] * ik;
      d -= m * m;
      d -= S.tau * m * missing;
    }
    if (q == 0) L0 = m;
    else if (q == 1) L1 = m;
    else L2 = m;
  }
  double inv;
  if (f.adiag == 0)
    inv = d;  // null row: left as accumulated (PCGSolver.h:141)
  else if (d < S.mdr * f.adiag)
    inv = 1 / sqrt(f.adiag);
  else
    inv = 1 / sqrt(d);
  Cell c;
  c.l0 = L0;
  c.l1 = L1;
  c.l2 = L2;
  c.inv = inv;
  *at(S.cf, idx) = c;  // the backward records are gathered from these afterwards
  return inv;
}

// One pipelined pass.  K_BWD mirrors every axis: local coordinates run from
// the high corner of the box, upstream tiles are (a+1, b) and (a, b+1).
//
// A workgroup is four compute waves (one lane per pencil of the 16x16 tile)
// and one face wave.  The compute lanes exchange each step's values through
// the LDS array lz; the values beside the tile's low x and y faces, which
// the upstream tiles produce, reach them through the LDS rings hx / hy,
// filled by the face wave.  So the compute waves never issue a
// device-coherent load (which would sit in their in-order load counter
// ahead of the record prefetches, and takes a round trip to memory).  The
// face wave works in batches of kFB steps: once per batch it takes the face
// values of the next kFB steps, read one batch earlier (re-reading, with a
// bounded wait, a value that was not final yet), publishes them into the
// rings and reads the batch after.
constexpr int kFB = 2;             // steps per face batch
constexpr int kRing = 2 * kFB;     // ring slots (a power of two)

// The forward and backward sweeps stage each step's records (factor
// entries, rhs, code) through an LDS ring filled by two loader waves with
// direct global-to-LDS loads (one wave-wide 16-byte load copies 1 KB), kAhead
// steps ahead.  The compute waves read a step's records from LDS one step
// early, so their only global-memory operations are the output stores, which
// nothing waits for; the loaders' in-order load counters hold nothing but
// their own ring loads, and each waits for exactly the step it publishes.
template <class V>
struct ApplyRing {
  static constexpr int kAhead = sizeof(V) == 8 ? 7 : 14;   // steps loaded ahead
  static constexpr int kSlots = sizeof(V) == 8 ? 8 : 16;   // > kAhead, a power of two
  static constexpr int kCellOps = kThreads * (int)sizeof(CellT<V>) / 1024;
  static constexpr int kRhsOps = kThreads * (int)sizeof(V) / 1024;
  static constexpr int kCodeOps = kThreads * 4 / 1024;
  // rhs loader, per step: rhs and code loads, and then (forward) the
  // sentinel stores or (backward) the loads of r for the z . r partials
  static constexpr int kStepOps = kRhsOps + kCodeOps;
  static constexpr int kBwdOps = kStepOps + kRhsOps;
  static_assert(kCellOps * kAhead <= 63 && kStepOps * kAhead + kRhsOps * (kAhead - 1) <= 63 &&
                    kBwdOps * kAhead <= 63,
                "a loader's loads and stores in flight fit its 6-bit counter");
};
template <int KIND>
struct SweepBlock {
  static constexpr int value = KIND == K_FAC ? kThreads + 64 : kThreads + 192;
};

// s_waitcnt with only the vector-memory count set (expcnt and lgkmcnt at
// their maxima, so they are not waited for); gfx9 encoding
template <int N>
__device__ inline void wait_vmcnt() {
  static_assert(N >= 0 && N <= 63, "vmcnt is 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

template <int KIND, class V>
__global__ __launch_bounds__(SweepBlock<KIND>::value) void k_sweep(SweepArgs<V> S) {
  constexpr bool BWD = KIND == K_BWD;
  constexpr bool FAC = KIND == K_FAC;
  constexpr int kCh = PassIn<KIND, V>::chunk;
  using In = typename PassIn<KIND, V>::T;
  using Ring = ApplyRing<V>;
  constexpr int kR = FAC ? 1 : Ring::kSlots;
  // One LDS array, indexed by element: the step-parity exchange planes
  // [parity][lj][li] (row stride kT + 1), the face rings [step % kRing][16]
  // for the upstream x face (indexed by lj) and y face (by li), and one +0,
  // the operand of an absent term.
  // (The factor pass uses this layout; the sweeps use kPlanes exchange
  // planes with halos in the same array.)
  constexpr uint32_t kLz = 0, kHx = 2 * kT * (kT + 1), kHy = kHx + kRing * kT,
                     kZero = kHy + kRing * kT;
  constexpr uint32_t kLdsN = FAC ? kZero + 1 : (uint32_t)(kPlanes * kPlane);
  __shared__ V lds[kLdsN];
  // record rings of the sweeps: [step % kR][position of the slot in the step]
  __shared__ CellT<V> rc[kR][kThreads];
  __shared__ V rr[kR][kThreads];
  __shared__ uint32_t rk[kR][kThreads];
  // backward sweep: r of each step, for the z . r partials
  __shared__ V rq[BWD ? kR : 1][kThreads];
  __shared__ double sdot[kThreads / 64];
  auto LZ = [](int par, int lj, int li) -> uint32_t {
    return kLz + (uint32_t)((par * kT + lj) * (kT + 1) + li);
  };
  auto HX = [](int w, int lj) -> uint32_t { return kHx + (uint32_t)((w & (kRing - 1)) * kT + lj); };
  auto HY = [](int w, int li) -> uint32_t { return kHy + (uint32_t)((w & (kRing - 1)) * kT + li); };
  __shared__ int s_tile;
  if constexpr (FAC) {
    if (threadIdx.x == 0) lds[kZero] = (V)0;
  } else {
    if (threadIdx.x < kPlanes) lds[threadIdx.x * kPlane] = (V)0;  // the +0 cells
  }
  // plane of step w, its cell of local pencil (li, lj), and its halos
  auto PL = [](int w) -> uint32_t { return (uint32_t)(w & (kPlanes - 1)) * kPlane; };
  // wave-uniform (scalar) branches: waves 0-3 compute, 4 reads the faces,
  // 5 and 6 (sweeps only) load the record rings
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const bool face_wave = wave == kThreads / 64;
  const int Bz = S.K.Bz;
  const int nsteps = S.K.nsteps;
  for (;;) {
    if (threadIdx.x == 0) {
      const int t = atomicAdd(S.ticket, 1);
      s_tile = t < S.ntiles ? (int)S.order[t] : -1;
      // the other pass's counter is not used while this one runs
      if constexpr (!FAC)
        if (t == 0) *S.next_ticket = 0;
    }
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    if (tile < 0) return;
    // logical tile (a, b) counted from the sweep's start corner
    const int ta = tile & 0xFFFF, tb = tile >> 16;
    const int a = BWD ? S.K.ntile_a - 1 - ta : ta;
    const int b = BWD ? S.K.ntile_b - 1 - tb : tb;
    const bool has_a = BWD ? a + 1 < S.K.ntile_a : a > 0;
    const bool has_b = BWD ? b + 1 < S.K.ntile_b : b > 0;
    const int sgn = BWD ? -kThreads : kThreads;
    // The pencil of local lane (li, lj) computes its cell kk at step
    // w = kk + li + lj, stored at slot base + sgn * w (the backward sweep runs
    // the skewed steps in reverse).  kk is clamped into the pencil for loads,
    // so every load address is valid (steps outside it discard the value).
    auto base_of = [&](int li, int lj) -> uint32_t {
      const int lip = BWD ? kT - 1 - li : li, ljp = BWD ? kT - 1 - lj : lj;
      return (uint32_t)skew_slot(S.K, a, b, lip, ljp, BWD ? nsteps - 1 : 0);
    };
    auto in_index = [&](uint32_t base, int off, int kk) -> uint32_t {
      const int kc = kk < 0 ? 0 : (kk >= Bz ? Bz - 1 : kk);
      return base + (uint32_t)(sgn * (kc + off));
    };
    if constexpr (!FAC) {
      if (wave > kThreads / 64) {
        // loader waves.  The 256 slots of step w are consecutive, starting
        // at the slot of physical lane 0; steps past the end reload the
        // last step (their ring slots are never read), so every loader
        // issues the same number of loads for every step and its counter
        // wait stays exact.
        const bool cells = wave == kThreads / 64 + 1;
        const int lane = threadIdx.x & 63;
        auto block = [&](int w) -> uint32_t {
          const int wc = w < nsteps ? w : nsteps - 1;
          return (uint32_t)skew_slot(S.K, a, b, 0, 0, BWD ? nsteps - 1 - wc : wc);
        };
        auto issue = [&](int w) {
          const uint32_t s0 = block(w);
          const int r = w & (kR - 1);
          if (cells) {
            const char* g = reinterpret_cast<const char*>(S.cells) + (size_t)s0 * sizeof(CellT<V>);
#pragma unroll
            for (int k = 0; k < Ring::kCellOps; ++k)
              __builtin_amdgcn_global_load_lds(
                  g + k * 1024 + lane * 16,
                  (__attribute__((address_space(3))) void*)(reinterpret_cast<char*>(&rc[r][0]) + k * 1024),
                  16, 0, 0);
          } else {
            const char* g = reinterpret_cast<const char*>(S.in) + (size_t)s0 * sizeof(V);
#pragma unroll
            for (int k = 0; k < Ring::kRhsOps; ++k)
              __builtin_amdgcn_global_load_lds(
                  g + k * 1024 + lane * 16,
                  (__attribute__((address_space(3))) void*)(reinterpret_cast<char*>(&rr[r][0]) + k * 1024),
                  16, 0, 0);
            const char* gk = reinterpret_cast<const char*>(S.code) + (size_t)s0 * 4;
#pragma unroll
            for (int k = 0; k < Ring::kCodeOps; ++k)
              __builtin_amdgcn_global_load_lds(
                  gk + k * 1024 + lane * 16,
                  (__attribute__((address_space(3))) void*)(reinterpret_cast<char*>(&rk[r][0]) + k * 1024),
                  16, 0, 0);
            if constexpr (BWD) {
              const char* gq = reinterpret_cast<const char*>(S.dot_r) + (size_t)s0 * sizeof(V);
#pragma unroll
              for (int k = 0; k < Ring::kRhsOps; ++k)
                __builtin_amdgcn_global_load_lds(
                    gq + k * 1024 + lane * 16,
                    (__attribute__((address_space(3))) void*)(reinterpret_cast<char*>(&rq[r][0]) + k * 1024),
                    16, 0, 0);
            }
          }
        };
        // The forward pass's rhs loader also sets the step's slots of
        // S.reset (the backward pass's output) to the sentinel, one 16-byte
        // store per lane and KB.
        // (the sentinel words are kept in four registers of their own, so a
        // store never waits for an earlier one to have read its data)
        typedef unsigned int U4 __attribute__((ext_vector_type(4)));
        U4 ones = {~0u, ~0u, ~0u, ~0u};
        asm volatile("" : "+v"(ones));
        auto clear = [&](int w) {
          char* g = reinterpret_cast<char*>(S.reset) + (size_t)block(w) * sizeof(V);
#pragma unroll
          for (int k = 0; k < Ring::kRhsOps; ++k)
            *reinterpret_cast<U4*>(g + k * 1024 + lane * 16) = ones;
        };
        for (int t = 0; t < Ring::kAhead; ++t) issue(t);
        for (int w = 0; w < nsteps; ++w) {
          // steps w and w + 1 complete: the loads of the kAhead - 2 later
          // steps may still be in flight, and (rhs loader) the stores of
          // the min(w, kAhead - 1) steps before w, issued after step w + 1's
          // loads (the counter is in order; early steps wait for more)
          if (cells)
            wait_vmcnt<Ring::kCellOps * (Ring::kAhead - 2)>();
          else if (BWD)
            wait_vmcnt<Ring::kBwdOps * (Ring::kAhead - 2)>();
          else if (w < Ring::kAhead - 1)
            wait_vmcnt<Ring::kStepOps * (Ring::kAhead - 2)>();
          else
            wait_vmcnt<Ring::kStepOps * (Ring::kAhead - 2) +
                       Ring::kRhsOps * (Ring::kAhead - 1)>();
          __builtin_amdgcn_s_barrier();
          issue(w + Ring::kAhead);
          if (!cells && !BWD) clear(w);
        }
        wait_vmcnt<0>();
        __builtin_amdgcn_s_barrier();
        continue;
      }
    }
    if (face_wave) {
      // lanes 0..15: x face of local row lj = lane (the pencil (0, lj));
      // lanes 16..31: y face of column li = lane - 16 (the pencil (li, 0));
      // lanes 32..63 repeat lanes 0..31 and publish nothing
      const int fl = (threadIdx.x - kThreads) & 31;
      const bool xf = fl < kT;
      const int li = xf ? 0 : fl - kT, lj = xf ? fl : 0;
      const bool pub = threadIdx.x - kThreads < 32;
      const bool on = pub && (xf ? has_a : has_b);
      const uint32_t base = base_of(li, lj);
      const int off = li + lj;
      const int dstep = on ? (int)(xf ? (BWD ? -S.K.dx : S.K.dx) : (BWD ? -S.K.dy : S.K.dy)) : 0;
      // the upstream value beside this pencil's cell of step w (own slot if
      // there is no upstream tile: read, never waited for, never used)
      auto face_addr = [&](int w) -> uint32_t { return in_index(base, off, w - off) + (uint32_t)dstep; };
      V fr[kFB];
      auto fetch = [&](int w0) {
#pragma unroll
        for (int q = 0; q < kFB; ++q) fr[q] = load_coherent(at(S.out, face_addr(w0 + q)));
      };
      // steps w0 .. w0 + kFB - 1; a value is waited for only where the
      // pencil's cell of that step exists (elsewhere the compute lane
      // discards it and 0 is published)
      auto publish = [&](int w0) {
#pragma unroll
        for (int q = 0; q < kFB; ++q) {
          const int w = w0 + q;
          const int kk = w - off;
          const bool need = on && kk >= 0 && kk < Bz;
          V v = fr[q];
          if (need && is_sentinel(v)) v = wait_value(at(S.out, face_addr(w)), S.abort_flag);
          if (!need) v = (V)0;
          if (pub) {
            if constexpr (FAC) {
              if (xf)
                lds[HX(w, lj)] = v;
              else
                lds[HY(w, li)] = v;
            } else {
              // the halo of the plane the step's operands are read from
              lds[PL(w - 1) + (xf ? (uint32_t)(lj + 1) * (kT + 1) : (uint32_t)(li + 1))] = v;
            }
          }
        }
      };
      fetch(0);
      publish(0);
      fetch(kFB);
      for (int w0 = 0; w0 < nsteps; w0 += kFB) {
#pragma clang loop unroll(full)
        for (int q = 0; q < kFB; ++q) {
          const int w = w0 + q;
          if (w < nsteps) {
            __syncthreads();
            // during the batch's last step: publish the next batch (read
            // one batch ago) and read the one after
            if (q == kFB - 1 && w + 1 < nsteps) {
              publish(w + 1);
              fetch(w + 1 + kFB);
            }
          }
        }
      }
      __syncthreads();
      continue;
    }
    const int li = threadIdx.x % kT, lj = threadIdx.x / kT;
    const uint32_t base = base_of(li, lj);
    const int off = li + lj;  // this pencil starts at step off
    const int lxi = li > 0 ? li - 1 : 0;  // LDS neighbours (read unconditionally)
    const int lyj = lj > 0 ? lj - 1 : 0;
    if constexpr (!FAC) {
      // The step's records come from the ring, read one step early: the
      // factor entries, the rhs and the operand word (three offsets from
      // this pencil's position in the previous plane).  Padding slots hold
      // zero records and operand words that all point at +0, so they compute
      // +0 like any other slot.
      const int pos = BWD ? kThreads - 1 - (int)threadIdx.x : (int)threadIdx.x;
      const uint32_t own = (uint32_t)((lj + 1) * (kT + 1) + li + 1);
      CellT<V> c;
      V rhs, rv = 0;
      uint32_t op = 0;
      auto take = [&](int w) {
        const int r = w & (kR - 1);
        c = rc[r][pos];
        rhs = rr[r][pos];
        op = rk[r][pos];
        if constexpr (BWD) rv = rq[r][pos];
      };
      // backward sweep: this lane's z . r over its pencil, in FP64 (padding
      // slots have z = r = 0 and add +0)
      double dot = 0.0;
      uint32_t src[3];
      auto sources = [&](int w) {
        const uint32_t at0 = PL(w - 1) + own;
#pragma unroll
        for (int q = 0; q < 3; ++q)
          src[q] = at0 + (uint32_t)__builtin_amdgcn_sbfe((int)op, 10 * q, 10);
      };
      for (int w = 0; w < nsteps; ++w) {
        __syncthreads();  // ring slots of steps w and w + 1 are complete
        if (w == 0) {
          take(0);
          sources(0);
        }
        const V v0 = lds[src[0]], v1 = lds[src[1]], v2 = lds[src[2]];
        CellIn<V> ci;
        ci.c = c;
        ci.rhs = rhs;
        const V z = sweep_cell_ordered<V>(ci, v0, v1, v2);
        // every step writes the lane's slot of that step (+0 on the padding
        // slots), so no slot keeps the sentinel after the pass
        store_coherent(at(S.out, base + (uint32_t)(sgn * w)), z);
        lds[PL(w) + own] = z;
        if constexpr (BWD) dot += (double)z * (double)rv;
        if (w + 1 < nsteps) {
          take(w + 1);
          sources(w + 1);
        }
      }
      if constexpr (BWD) {
        // the tile's partial: a fixed tree over the 64 lanes of each wave,
        // then the four waves in order
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o, 64);
        if ((threadIdx.x & 63) == 0) sdot[threadIdx.x >> 6] = dot;
      }
      __syncthreads();
      if constexpr (BWD) {
        if (threadIdx.x == 0) {
          double t = sdot[0];
#pragma unroll
          for (int q = 1; q < kThreads / 64; ++q) t += sdot[q];
          S.dot_out[b * S.K.ntile_a + a] = t;
        }
      }
    } else {
    V zprev = 0;
    // Record prefetch ring of the factor pass: slot s holds the inputs of
    // step w0 + s, read kCh steps ahead.  The step loop is unrolled by kCh
    // and a slot is refilled only after its step has used it, so every load
    // writes the registers it is consumed from and no wait on a load in
    // flight is needed before its own step.  (The scheduling barriers keep
    // the loads in consumption order: the counter wait before a slot's first
    // use then leaves the loads of the later slots in flight.)
    In win[kCh];
#pragma unroll
    for (int s = 0; s < kCh; ++s) {
      win[s] = load_in<KIND, V>(S, in_index(base, off, s - off));
      __builtin_amdgcn_sched_barrier(0);
    }
    for (int w0 = 0; w0 < nsteps; w0 += kCh) {
#pragma clang loop unroll(full)
      for (int s = 0; s < kCh; ++s) {
        const int w = w0 + s;
        if (w < nsteps) {
          const int kk = w - off;
          const bool active = kk >= 0 && kk < Bz;
          const In ci = win[s];
          const uint32_t code = active ? code_of(ci) : 0u;
          const uint32_t idx = base + (uint32_t)(sgn * w);
          __syncthreads();
          const V lx = lds[LZ((w - 1) & 1, lj, lxi)];
          const V ly = lds[LZ((w - 1) & 1, lyj, li)];
          const V fx = lds[HX(w, lj)];
          const V fy = lds[HY(w, li)];
          const V zx = li > 0 ? lx : fx;
          const V zy = lj > 0 ? ly : fy;
          V z = 0.0;
          if (code & S_DOF) z = factor_cell(S, ci, code, zx, zy, zprev, idx);
          store_coherent(at(S.out, idx), z);
          zprev = z;
          lds[LZ(w & 1, lj, li)] = z;
        }
        win[s] = load_in<KIND, V>(S, in_index(base, off, w + kCh - off));
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();
    }
  }
}

// ---- Sub-tile sweeps: one wave per 8x8 quarter of a tile ----
//
// The same passes over the same skewed layout, records and operand words,
// with the work unit cut to an 8x8 quarter (sub-tile) of a 16x16 tile, so
// that one wave holds every pencil of its unit.  Its lanes then exchange
// each step's values through LDS without a workgroup barrier (the LDS
// processes one wave's operations in order, so a read issued after a write
// sees it), and the only barriers are those that hand over a batch of kNB
// steps of records and face values from the loader and face waves.  There
// are twice as many hops between units as between tiles, each costing the
// face read-ahead and a round trip; the shorter step is what the change
// buys.
//
// A sub-tile (sa, sb) of tile (a, b) runs every step w of the tile, 0 ..
// nsteps - 1, for its 64 pencils (li, lj) = (8 sa + ui, 8 sb + uj): the
// steps outside its pencils' cells are padding, which it writes as +0 like
// every other slot, so each slot of the box is still written once per pass.
// Units are claimed by ticket in (A + B, A) order over the sub-tile grid
// (A, B) = (2 a + sa, 2 b + sb), and a unit waits only for the units at
// (A - 1, B) and (A, B - 1): inside a tile through the slot offsets -257 and
// -272 (one step back and one pencil over), across tiles through K.dx and
// K.dy, mirrored in the backward pass.
constexpr int kS = 8;             // sub-tile edge (pencils)
constexpr int kNB = 4;            // steps per batch
constexpr int kSubPlanes = 16;    // exchange planes (step % 16), > 2 kNB - 1
constexpr int kFaceWaves = 1;     // face waves (one read in flight each)
constexpr int kSubThreads = 64 * (2 + kFaceWaves);  // compute, face(s), loader

template <class V>
struct SubRing {
  static constexpr int kCellOps = kNB * 64 * (int)sizeof(CellT<V>) / 1024;
  static constexpr int kRhsOps = kNB * 64 * (int)sizeof(V) / 1024;
  static constexpr int kOpOps = kNB * 64 * 4 / 1024;
  static constexpr int kFwdLoads = kCellOps + kRhsOps + kOpOps;
  static constexpr int kBwdLoads = kFwdLoads + kRhsOps;  // and r for z . r
  // ring depth (batches, a power of two): 64 KB of factor records
  static constexpr int kBatches = 64 * 1024 / (kNB * 64 * (int)sizeof(CellT<V>));
  // batches loaded ahead: as many as the 6-bit load counter and the ring allow
  static constexpr int ahead() {
    int a = kBatches - 1;
    while ((a - 1) * kFwdLoads + a * kRhsOps > 63 || (a - 1) * kBwdLoads > 63) --a;
    return a;
  }
  static constexpr int kAhead = ahead();
  static_assert(kAhead >= 2 && kAhead < kBatches, "the ring holds the batches in flight");
};

// s_waitcnt lgkmcnt(0) alone (gfx9 encoding)
__device__ inline void wait_lgkmcnt0() {
  __builtin_amdgcn_s_waitcnt(15 | (7 << 4) | (3 << 14));
}

template <int KIND, class V>
__global__ __launch_bounds__(kSubThreads) void k_sweep_sub(SweepArgs<V> S) {
  static_assert(KIND == K_FWD || KIND == K_BWD, "sweeps only");
  constexpr bool BWD = KIND == K_BWD;
  using Ring = SubRing<V>;
  constexpr int kRB = Ring::kBatches;
  constexpr int kLA = Ring::kAhead;
  // exchange planes in the tile's 17 x 17 geometry (the operand words'),
  // each with a +0 at [0][0] that no unit writes
  __shared__ V pl[kSubPlanes * kPlane];
  // record rings [batch % kRB][step in batch][physical position in the unit]
  __shared__ CellT<V> rc[kRB][kNB][64];
  __shared__ V rr[kRB][kNB][64];
  __shared__ uint32_t rk[kRB][kNB][64];
  __shared__ V rq[BWD ? kRB : 1][kNB][64];
  __shared__ int s_unit;
  if (threadIdx.x < kSubPlanes) pl[threadIdx.x * kPlane] = (V)0;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int Bz = S.K.Bz;
  const int nsteps = S.K.nsteps;
  const int nb = (nsteps + kNB - 1) / kNB;
  const int nsa = 2 * S.K.ntile_a;
  auto PL = [](int w) -> uint32_t { return (uint32_t)(w & (kSubPlanes - 1)) * kPlane; };
  const int sgn = BWD ? -kThreads : kThreads;
  // The next unit (A, B), counted from the sweep's start corner, or -1.
  // Each role runs its own loop over the units (so the compute wave's code
  // holds none of the loader's direct-to-LDS loads, and the compiler adds
  // no wait for them there); the loops meet at the same barriers.
  auto claim = [&]() -> int {
    if (threadIdx.x == 0) {
      const int t = atomicAdd(S.ticket, 1);
      s_unit = t < S.ntiles ? (int)S.order[t] : -1;
      if (t == 0) *S.next_ticket = 0;  // the other pass's counter
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(s_unit);
  };
  // a unit's tile and quarter in the sweep's frame and the physical ones
  struct Unit {
    int A, B, sa, sb, a, b, psa, psb;
  };
  auto decode = [&](int unit) -> Unit {
    Unit u;
    u.A = unit & 0xFFFF;
    u.B = unit >> 16;
    u.sa = u.A & 1;
    u.sb = u.B & 1;
    u.a = BWD ? S.K.ntile_a - 1 - (u.A >> 1) : u.A >> 1;
    u.b = BWD ? S.K.ntile_b - 1 - (u.B >> 1) : u.B >> 1;
    u.psa = BWD ? 1 - u.sa : u.sa;
    u.psb = BWD ? 1 - u.sb : u.sb;
    return u;
  };
  auto base_of = [&](const Unit& u, int li, int lj) -> uint32_t {
    const int lip = BWD ? kT - 1 - li : li, ljp = BWD ? kT - 1 - lj : lj;
    return (uint32_t)skew_slot(S.K, u.a, u.b, lip, ljp, BWD ? nsteps - 1 : 0);
  };
  if (wave == 2) {
    // loader: the unit's 64 slots of step w are 8 rows of 8 consecutive
    // slots (physical order p = row * 8 + column); steps past the end
    // reload the last one, so every batch issues the same operations and
    // the counter waits stay exact
    typedef __attribute__((address_space(3))) void* LdsPtr;
    typedef unsigned int U4 __attribute__((ext_vector_type(4)));
    U4 ones = {~0u, ~0u, ~0u, ~0u};
    asm volatile("" : "+v"(ones));
    for (;;) {
      const int unit = claim();
      if (unit < 0) return;
      const Unit u = decode(unit);
      const uint32_t tile0 = (uint32_t)skew_slot(S.K, u.a, u.b, 0, 0, 0);
      auto slot = [&](int w, int p) -> uint32_t {
        const int wc = w < nsteps ? w : nsteps - 1;
        const int wp = BWD ? nsteps - 1 - wc : wc;
        return tile0 + (uint32_t)wp * kThreads +
               (uint32_t)((kS * u.psb + (p >> 3)) * kT + kS * u.psa + (p & 7));
      };
      // 16 bytes per lane from 'src' (element size E) into 'dst' for the
      // kNB steps of batch n: each wave-wide load covers 1024 / (64 E) steps
      auto stream = [&](const void* src, void* dst, int n, auto E) {
        constexpr int per = 1024 / (64 * (int)decltype(E)::value);  // steps per load
        constexpr int lanes = 64 / per;                              // lanes per step
#pragma unroll
        for (int j = 0; j < kNB / per; ++j) {
          const int st = j * per + lane / lanes;
          const int p = (lane % lanes) * (16 / (int)decltype(E)::value);
          const char* g = reinterpret_cast<const char*>(src) +
                          (size_t)slot(n * kNB + st, p) * decltype(E)::value;
          __builtin_amdgcn_global_load_lds(g, (LdsPtr)(reinterpret_cast<char*>(dst) + j * 1024),
                                           16, 0, 0);
        }
      };
      using EV = std::integral_constant<int, (int)sizeof(V)>;
      using E4 = std::integral_constant<int, 4>;
      auto issue = [&](int n) {
        const int r = n & (kRB - 1);
        // records: 16 bytes per lane, sizeof(CellT) / 16 pieces per slot
        constexpr int kPieces = (int)sizeof(CellT<V>) / 16;
#pragma unroll
        for (int st = 0; st < kNB; ++st)
#pragma unroll
          for (int h = 0; h < kPieces; ++h) {
            const int p = h * (64 / kPieces) + lane / kPieces;
            const char* g = reinterpret_cast<const char*>(S.cells) +
                            (size_t)slot(n * kNB + st, p) * sizeof(CellT<V>) +
                            (lane % kPieces) * 16;
            __builtin_amdgcn_global_load_lds(
                g, (LdsPtr)(reinterpret_cast<char*>(&rc[r][st][0]) + h * 1024), 16, 0, 0);
          }
        stream(S.in, &rr[r][0][0], n, EV{});
        stream(S.code, &rk[r][0][0], n, E4{});
        if constexpr (BWD) stream(S.dot_r, &rq[r][0][0], n, EV{});
      };
      // forward pass: the backward output's slots of batch n to the sentinel
      auto clear = [&](int n) {
        constexpr int per = 1024 / (64 * (int)sizeof(V));
        constexpr int lanes = 64 / per;
#pragma unroll
        for (int j = 0; j < kNB / per; ++j) {
          const int st = j * per + lane / lanes;
          const int p = (lane % lanes) * (16 / (int)sizeof(V));
          *reinterpret_cast<U4*>(reinterpret_cast<char*>(S.reset) +
                                 (size_t)slot(n * kNB + st, p) * sizeof(V)) = ones;
        }
      };
      for (int n = 0; n < kLA; ++n) issue(n);
      for (int n = 0; n < nb; ++n) {
        // batch n complete; the loads of the kLA - 1 batches after it (and
        // in the forward pass the clears of the kLA batches before it,
        // issued after its loads) may still be in flight
        if (BWD)
          wait_vmcnt<(kLA - 1) * Ring::kBwdLoads>();
        else if (n < kLA)
          wait_vmcnt<(kLA - 1) * Ring::kFwdLoads>();
        else
          wait_vmcnt<(kLA - 1) * Ring::kFwdLoads + kLA * Ring::kRhsOps>();
        __builtin_amdgcn_s_barrier();
        issue(n + kLA);
        if constexpr (!BWD) clear(n);
      }
      wait_vmcnt<0>();
      __builtin_amdgcn_s_barrier();
    }
  }
  if (wave == 1 || wave >= 3) {
    // face waves (wave 1, and waves 3 .. with kFaceWaves > 1): face wave p
    // publishes the batches n = p (mod kFaceWaves), each read with one
    // read in flight on the wave's own counter, issued when the wave
    // publishes its previous batch, so a read has kFaceWaves batches to
    // return.  A batch's face values are the x face of rows m (pencils
    // (8 sa, 8 sb + m)) and the y face of columns m (pencils (8 sa + m,
    // 8 sb)) at each of its steps; upstream values come from the unit
    // before in x (y): the other quarter of this tile, one step back and
    // one pencil over, or the tile before.
    const int par = wave == 1 ? 0 : wave - 2;
    // item k of a lane is value g = 64 k + lane of the batch: face g /
    // (kNB kS) (x, then y), step (g / kS) % kNB, row or column g % kS
    constexpr int kPer = kNB * kS;
    constexpr int kItems = 2 * kPer / 64;
    static_assert(2 * kPer % 64 == 0, "a batch's face values fill whole waves");
    for (;;) {
      const int unit = claim();
      if (unit < 0) return;
      const Unit u = decode(unit);
      const int64_t dxs =
          u.sa ? (BWD ? kThreads + 1 : -(kThreads + 1)) : (BWD ? -S.K.dx : S.K.dx);
      const int64_t dys =
          u.sb ? (BWD ? kThreads + kT : -(kThreads + kT)) : (BWD ? -S.K.dy : S.K.dy);
      uint32_t ibase[kItems], idel[kItems], ihalo[kItems];
      int ioff[kItems], istep[kItems];
      bool iup[kItems];
#pragma unroll
      for (int k = 0; k < kItems; ++k) {
        const int g = 64 * k + lane;
        const bool xf = g < kPer;
        const int m = g % kS;
        istep[k] = (g / kS) % kNB;
        const int li = xf ? kS * u.sa : kS * u.sa + m, lj = xf ? kS * u.sb + m : kS * u.sb;
        iup[k] = xf ? u.A > 0 : u.B > 0;
        ibase[k] = base_of(u, li, lj);
        ioff[k] = li + lj;
        idel[k] = iup[k] ? (uint32_t)(xf ? dxs : dys) : 0u;
        // the halo cell beside the pencil in the exchange plane
        ihalo[k] = (uint32_t)(xf ? (lj + 1) * (kT + 1) + li : lj * (kT + 1) + li + 1);
      }
      // the pencil's cell of step w (clamped into the pencil, so the
      // address is always a slot of the box) and the upstream slot beside it
      auto addr = [&](int k, int w) -> uint32_t {
        const int kk = w - ioff[k];
        const int kc = kk < 0 ? 0 : (kk >= Bz ? Bz - 1 : kk);
        return ibase[k] + (uint32_t)(sgn * (kc + ioff[k])) + idel[k];
      };
      auto need = [&](int k, int w) -> bool {
        const int kk = w - ioff[k];
        return iup[k] && w < nsteps && kk >= 0 && kk < Bz;
      };
      V fv[kItems];
#pragma unroll
      for (int k = 0; k < kItems; ++k) fv[k] = (V)0;
      // (only where the value is needed: a unit's face reads are a fair
      // share of the sweep's coherent traffic)
      auto fetch = [&](int n) {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
          fv[k] = need(k, n * kNB + istep[k]) ? load_coherent(at(S.out, addr(k, n * kNB + istep[k])))
                                              : (V)0;
      };
      // batch n's values into the halos of the planes its steps read;
      // a value still the sentinel is read again (bounded wait)
      auto publish = [&](int n) {
        bool nd[kItems];
        bool bad = false;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
          nd[k] = need(k, n * kNB + istep[k]);
          bad = bad || (nd[k] && is_sentinel(fv[k]));
        }
        // (the first check stays outside the retry loop: with the loop's
        // reloads on its path the compiler waits for every read in flight)
        if (__any(bad)) {
          long long spins = 0;
          for (;;) {
            if (++spins > kSpinLimit || load_flag(S.abort_flag)) {
              store_flag(S.abort_flag, 1);
#pragma unroll
              for (int k = 0; k < kItems; ++k) fv[k] = (V)0;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            bool again = false;
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
              if (nd[k] && is_sentinel(fv[k]))
                fv[k] = load_coherent(at(S.out, addr(k, n * kNB + istep[k])));
              again = again || (nd[k] && is_sentinel(fv[k]));
            }
            if (!__any(again)) break;
          }
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
          const int w = n * kNB + istep[k];
          if (w < nsteps) pl[PL(w - 1) + ihalo[k]] = nd[k] ? fv[k] : (V)0;
        }
      };
      // backward pass: the z . r partial of compute lane 'lane', summed
      // here so the compute wave's steps stay short: r of batch n is copied
      // from the ring during batch n, and z of batch n read from the
      // exchange planes during batch n + 1 (the planes keep a step for 16
      // steps), in step order, as the compute lane would sum them
      const int cli = kS * u.sa + (lane & 7), clj = kS * u.sb + (lane >> 3);
      const uint32_t cown = (uint32_t)((clj + 1) * (kT + 1) + cli + 1);
      const int cpos = BWD ? 63 - lane : lane;
      double dot = 0.0;
      V rsave[kNB];
      auto take_r = [&](int n) {
#pragma unroll
        for (int st = 0; st < kNB; ++st) rsave[st] = rq[n & (kRB - 1)][st][cpos];
      };
      auto add_z = [&](int n) {
#pragma unroll
        for (int st = 0; st < kNB; ++st) {
          const int w = n * kNB + st;
          if (w < nsteps) dot += (double)pl[PL(w) + cown] * (double)rsave[st];
        }
      };
      // batch 0 is due at the first barrier: face wave 0 reads and
      // publishes it now; then each face wave has its next batch in flight
      if (par == 0) {
        fetch(0);
        publish(0);
      }
      {
        const int first = par == 0 ? kFaceWaves : par;
        if (first < nb) fetch(first);
      }
      for (int n = 0; n < nb; ++n) {
        __syncthreads();  // batch n handed over
        const int due = n + 1;  // published during batch n
        if (due < nb && due % kFaceWaves == par) {
          publish(due);
          if (due + kFaceWaves < nb) fetch(due + kFaceWaves);
        }
        if constexpr (BWD) {
          if (par == 0) {
            if (n > 0) add_z(n - 1);
            take_r(n);
          }
        }
      }
      __syncthreads();
      if (BWD && par == 0) {
        add_z(nb - 1);
        // the unit's partial: a fixed tree over its 64 lanes, stored at the
        // physical quarter's index
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o, 64);
        if (lane == 0) S.dot_out[(2 * u.b + u.psb) * nsa + 2 * u.a + u.psa] = dot;
      }
    }
  }
  // compute wave: lane (ui, uj) = (lane % 8, lane / 8) is pencil
  // (li, lj) = (8 sa + ui, 8 sb + uj) of the tile
  const int pos = BWD ? 63 - lane : lane;
  for (;;) {
    const int unit = claim();
    if (unit < 0) return;
    const Unit u = decode(unit);
    const int li = kS * u.sa + (lane & 7), lj = kS * u.sb + (lane >> 3);
    const uint32_t base = base_of(u, li, lj);
    const uint32_t own = (uint32_t)((lj + 1) * (kT + 1) + li + 1);
    for (int n = 0; n < nb; ++n) {
      // this wave's plane writes are complete (the face wave sums z . r
      // from them), and batch n's records and face values are in place
      wait_lgkmcnt0();
      __syncthreads();
      const int r = n & (kRB - 1);
      // the batch's records and operand addresses first (the ring is
      // complete for the whole batch), so a step waits on one LDS round
      // trip only: its operands, written by the step before
      CellIn<V> ci[kNB];
      uint32_t src[kNB][3];
#pragma unroll
      for (int st = 0; st < kNB; ++st) {
        ci[st].c = rc[r][st][pos];
        ci[st].rhs = rr[r][st][pos];
        const uint32_t op = rk[r][st][pos];
        const uint32_t at0 = PL(n * kNB + st - 1) + own;
#pragma unroll
        for (int q = 0; q < 3; ++q)
          src[st][q] = at0 + (uint32_t)__builtin_amdgcn_sbfe((int)op, 10 * q, 10);
      }
#pragma unroll
      for (int st = 0; st < kNB; ++st) {
        const int w = n * kNB + st;
        if (w >= nsteps) break;
        const V z = sweep_cell_ordered<V>(ci[st], pl[src[st][0]], pl[src[st][1]], pl[src[st][2]]);
        store_coherent(at(S.out, base + (uint32_t)(sgn * w)), z);
        pl[PL(w) + own] = z;
      }
    }
    wait_lgkmcnt0();
    __syncthreads();
  }
}

inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

Skew skew_of(const SweepPlan& P) {
  Skew K;
  K.ntile_a = P.ntile_a;
  K.ntile_b = P.ntile_b;
  K.nsteps = P.nsteps;
  K.Bz = P.B[2];
  K.dx = P.kdx;
  K.dy = P.kdy;
  return K;
}

}  // namespace

void sweep_prepare(hipStream_t st, SweepPlan& P, const Grid& g, const uint8_t* act,
                   const uint32_t* isdof, const uint32_t* dofscan,
                   const uint32_t* rowlen, const uint32_t* col, const double* val,
                   int64_t ndof) {
  P.valid = false;
  if (ndof == 0) return;
  P.flags.ensure(8);
  int init[6] = {1 << 30, 1 << 30, 1 << 30, -1, -1, -1};
  WC_HIP(hipMemcpyAsync(P.flags.p, init, sizeof(init), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_dof_box, dim3((unsigned)std::min<int64_t>(kBoxBlocks, blocks(g.nnode, 256))),
                     dim3(256), 0, st, g, isdof, P.flags.p);
  WC_LAUNCH_CHECK();
  int bb[6];
  WC_HIP(hipMemcpyAsync(bb, P.flags.p, sizeof(bb), hipMemcpyDeviceToHost, st));
  WC_HIP(hipStreamSynchronize(st));
  for (int d = 0; d < 3; ++d) {
    P.o[d] = bb[d];
    int ext = bb[3 + d] - bb[d] + 1;
    if (d < 2) ext = (ext + kT - 1) / kT * kT;
    P.B[d] = ext;
  }
  P.ntile_a = P.B[0] / kT;
  P.ntile_b = P.B[1] / kT;
  P.ntiles = P.ntile_a * P.ntile_b;
  P.nsteps = P.B[2] + 2 * (kT - 1);
  // slots of the skewed tile layout (every vector and record of the plan)
  P.ncell = (int64_t)P.ntiles * P.nsteps * kThreads;
  P.kdx = -(int64_t)P.nsteps * kThreads + (kT - 1) * kThreads + (kT - 1);
  P.kdy = -(int64_t)P.ntile_a * P.nsteps * kThreads + (kT - 1) * kThreads + (kT - 1) * kT;
  // the sweep kernels address every array with 32-bit byte offsets; larger
  // systems are solved with the level-scheduled kernels instead
  if (P.ncell * (int64_t)sizeof(FacIn) >= (int64_t(1) << 32)) return;
  P.pdof.ensure(P.ncell);
  P.code.ensure(P.ncell);
  P.scode.ensure(P.ncell);
  P.opf.ensure(P.ncell);
  P.opb.ensure(P.ncell);
  P.sflag.ensure(P.ncell);
  P.cf.ensure(P.ncell * sizeof(Cell) / sizeof(double));
  P.cb.ensure(P.ncell * sizeof(Cell) / sizeof(double));
  P.coef.ensure(P.ncell * kCoef);
  P.fin.ensure(P.ncell * sizeof(FacIn) / sizeof(double));
  P.zf.ensure(P.ncell);
  P.zb.ensure(P.ncell);
  P.rb.ensure(P.ncell);
  P.sb.ensure(P.ncell);
  P.xb.ensure(P.ncell);
  // tickets in (a+b, a) order
  std::vector<uint32_t> order;
  order.reserve(P.ntiles);
  for (int s = 0; s <= P.ntile_a + P.ntile_b - 2; ++s)
    for (int a = 0; a < P.ntile_a; ++a) {
      const int b = s - a;
      if (b >= 0 && b < P.ntile_b) order.push_back((uint32_t)a | ((uint32_t)b << 16));
    }
  P.order.ensure(P.ntiles);
  WC_HIP(hipMemcpyAsync(P.order.p, order.data(), sizeof(uint32_t) * P.ntiles,
                        hipMemcpyHostToDevice, st));
  // sub-tile tickets in (A+B, A) order over the 2 ntile_a x 2 ntile_b units
  std::vector<uint32_t> order_sub;
  const int nua = 2 * P.ntile_a, nub = 2 * P.ntile_b;
  order_sub.reserve(4 * P.ntiles);
  for (int s = 0; s <= nua + nub - 2; ++s)
    for (int A = 0; A < nua; ++A) {
      const int B = s - A;
      if (B >= 0 && B < nub) order_sub.push_back((uint32_t)A | ((uint32_t)B << 16));
    }
  P.order_sub.ensure(4 * P.ntiles);
  WC_HIP(hipMemcpyAsync(P.order_sub.p, order_sub.data(), sizeof(uint32_t) * 4 * P.ntiles,
                        hipMemcpyHostToDevice, st));
  P.progress.ensure(8);
  P.tdot.ensure(4 * P.ntiles);
  // tickets and the abort flag start at zero (the abort flag is cleared
  // again by each factor, not by the applies: once set it stays set)
  WC_HIP(hipMemsetAsync(P.progress.p, 0, sizeof(int) * 8, st));
  P.ticks_zero = true;
  BoxArgs A;
  A.g = g;
  for (int d = 0; d < 3; ++d) A.o[d] = P.o[d];
  A.K = skew_of(P);
  A.nslot = P.ncell;
  A.act = act;
  A.isdof = isdof;
  A.dofscan = dofscan;
  A.rowlen = rowlen;
  A.col = col;
  A.val = val;
  A.pdof = P.pdof.p;
  A.code = P.code.p;
  A.scode = P.scode.p;
  A.opf = P.opf.p;
  A.opb = P.opb.p;
  A.sflag = P.sflag.p;
  A.cf = reinterpret_cast<Cell*>(P.cf.p);
  A.cb = reinterpret_cast<Cell*>(P.cb.p);
  A.coef = P.coef.p;
  A.fin = reinterpret_cast<FacIn*>(P.fin.p);
  // row summaries at the DOFs' slots (zeros elsewhere), then the fill
  P.rsum.ensure(P.ncell * sizeof(RowSum) / sizeof(double));
  RowSum* rs = reinterpret_cast<RowSum*>(P.rsum.p);
  WC_HIP(hipMemsetAsync(rs, 0, sizeof(RowSum) * P.ncell, st));
  hipLaunchKernelGGL(k_row_sums, dim3((unsigned)std::min<int64_t>(4096, blocks(g.nnode, 256))),
                     dim3(256), 0, st, A, rs);
  WC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fill_box, dim3(blocks(P.ncell, 256)), dim3(256), 0, st, A,
                     (const RowSum*)rs);
  WC_LAUNCH_CHECK();
  // the windowed SpMV's x-face columns: of the -x coupling plane once, of
  // its operand with every operand
  P.cf15.ensure(P.ncell / 16);
  P.xf15.ensure(P.ncell / 16);
  P.xf0.ensure(P.ncell / 16);
  face_cols<double>(st, P.ncell, P.coef.p + 4 * P.ncell, P.cf15.p, nullptr);
  WC_HIP(hipStreamSynchronize(st));  // 'order' and 'order_sub' must outlive their copies
  P.valid = true;
}

template <class V>
static SweepArgs<V> sweep_args(SweepPlan& P) {
  SweepArgs<V> S = {};
  S.K = skew_of(P);
  S.ntiles = P.ntiles;
  S.order = P.order.p;
  S.code = P.scode.p;
  S.abort_flag = P.progress.p + 2;
  return S;
}

void sweep_factor(hipStream_t st, SweepPlan& P, double tau, double mdr) {
  int* tick = P.progress.p;
  WC_HIP(hipMemsetAsync(tick, 0, sizeof(int) * 8, st));
  WC_HIP(hipMemsetAsync(P.zf.p, 0xFF, sizeof(double) * P.ncell, st));
  SweepArgs<double> S = sweep_args<double>(P);
  S.fin = reinterpret_cast<const FacIn*>(P.fin.p);
  S.cf = reinterpret_cast<Cell*>(P.cf.p);
  S.tau = tau;
  S.mdr = mdr;
  S.out = P.zf.p;
  S.ticket = tick + 3;
  P.ticks_zero = true;   // counters 0 and 1 are not touched by the factor
  hipLaunchKernelGGL((k_sweep<K_FAC, double>), dim3(P.ntiles), dim3(SweepBlock<K_FAC>::value), 0, st, S);
  WC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_backward_records, dim3(blocks(P.ncell, 256)), dim3(256), 0, st,
                     P.ncell, skew_of(P), P.scode.p, reinterpret_cast<const Cell*>(P.cf.p),
                     reinterpret_cast<Cell*>(P.cb.p));
  WC_LAUNCH_CHECK();
}

void sweep_narrow(hipStream_t st, SweepPlan& P) {
  const int64_t nrec = P.ncell * 4;
  P.cf32.ensure(nrec);
  P.cb32.ensure(nrec);
  P.coef32.ensure(P.ncell * kCoef);
  P.zf32.ensure(P.ncell);
  P.zb32.ensure(P.ncell);
  P.rb32.ensure(P.ncell);
  P.sb32.ensure(P.ncell);
  P.bb.ensure(P.ncell);
  hipLaunchKernelGGL(k_narrow_box, dim3(blocks(nrec, 256)), dim3(256), 0, st, P.cf.p,
                     P.cf32.p, nrec);
  hipLaunchKernelGGL(k_narrow_box, dim3(blocks(nrec, 256)), dim3(256), 0, st, P.cb.p,
                     P.cb32.p, nrec);
  hipLaunchKernelGGL(k_narrow_box, dim3(blocks(P.ncell * kCoef, 256)), dim3(256), 0, st,
                     P.coef.p, P.coef32.p, P.ncell * kCoef);
  WC_LAUNCH_CHECK();
  P.cf15_32.ensure(P.ncell / 16);
  P.xf15_32.ensure(P.ncell / 16);
  P.xf0_32.ensure(P.ncell / 16);
  face_cols<float>(st, P.ncell, P.coef32.p + 4 * P.ncell, P.cf15_32.p, nullptr);
}

// all-ones (the sentinel) over n 16-byte words
__global__ void k_fill_sentinel(uint4* p, int64_t n) {
  const uint4 ones = make_uint4(~0u, ~0u, ~0u, ~0u);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = ones;
}

template <class V>
static void apply_impl(hipStream_t st, SweepPlan& P, const V* r, const V* cf, const V* cb,
                       V* zf, V* zb, KTimer* timer, int64_t ndof, bool cleared) {
  // Two ticket counters; the outputs start as the sentinel.  Each pass zeroes
  // the other pass's counter, and the forward pass's rhs loader sets every
  // slot of zb to the sentinel as it goes, so an apply clears only zf (just
  // before the forward pass, which leaves it cache-resident for the face
  // reads) and, after the plan or the factor, the counters.  The clear is a
  // kernel of its own: a runtime fill waited ~15 us longer to start.
  int* tick = P.progress.p;
  if (!P.ticks_zero) WC_HIP(hipMemsetAsync(tick, 0, sizeof(int) * 2, st));
  static_assert(kThreads % 16 == 0, "slots come in whole 16-byte words");
  const int64_t words = (int64_t)sizeof(V) * P.ncell / 16;
  if (!cleared)  // else the PCG update before this apply has filled zf
    hipLaunchKernelGGL(k_fill_sentinel,
                       dim3((unsigned)std::min<int64_t>((words + 255) / 256, 4096)), dim3(256), 0,
                       st, reinterpret_cast<uint4*>(zf), words);
  SweepArgs<V> S = sweep_args<V>(P);
  S.cells = reinterpret_cast<const CellT<V>*>(cf);
  S.in = r;
  S.out = zf;
  S.reset = zb;
  S.code = P.opf.p;  // the sweeps read operand words, not the code
  S.ticket = tick;
  S.next_ticket = tick + 1;
  const bool sub = P.sub;
  if (sub) {
    S.ntiles = 4 * P.ntiles;  // tickets: the sub-tile units
    S.order = P.order_sub.p;
  }
  P.ndot = S.ntiles;
  hipEvent_t e0 = timer ? timer->mark(st) : nullptr;
  if (sub)
    hipLaunchKernelGGL((k_sweep_sub<K_FWD, V>), dim3(S.ntiles), dim3(kSubThreads), 0, st, S);
  else
    hipLaunchKernelGGL((k_sweep<K_FWD, V>), dim3(P.ntiles), dim3(SweepBlock<K_FWD>::value), 0, st, S);
  WC_LAUNCH_CHECK();
  hipEvent_t e1 = timer ? timer->mark(st) : nullptr;
  S.cells = reinterpret_cast<const CellT<V>*>(cb);
  S.in = zf;
  S.out = zb;
  S.reset = nullptr;
  S.dot_r = r;
  S.dot_out = P.tdot.p;
  S.code = P.opb.p;
  S.ticket = tick + 1;
  S.next_ticket = tick;
  hipEvent_t e2 = timer ? timer->mark(st) : nullptr;
  if (sub)
    hipLaunchKernelGGL((k_sweep_sub<K_BWD, V>), dim3(S.ntiles), dim3(kSubThreads), 0, st, S);
  else
    hipLaunchKernelGGL((k_sweep<K_BWD, V>), dim3(P.ntiles), dim3(SweepBlock<K_BWD>::value), 0, st, S);
  WC_LAUNCH_CHECK();
  P.ticks_zero = true;
  if (timer) {
    timer->add(KTimer::SWEEP_FWD, ndof, e0, e1);
    timer->add(KTimer::SWEEP_BWD, ndof, e2, timer->mark(st));
  }
}

void sweep_apply(hipStream_t st, SweepPlan& P, const double* r, KTimer* timer,
                 int64_t ndof, bool cleared) {
  apply_impl<double>(st, P, r, P.cf.p, P.cb.p, P.zf.p, P.zb.p, timer, ndof, cleared);
}

void sweep_apply32(hipStream_t st, SweepPlan& P, const float* r, KTimer* timer,
                   int64_t ndof, bool cleared) {
  apply_impl<float>(st, P, r, P.cf32.p, P.cb32.p, P.zf32.p, P.zb32.p, timer, ndof, cleared);
}

static unsigned spmv_blocks(const SweepPlan& P) {
  return (unsigned)std::min<int64_t>(1024, blocks(P.ncell, kSpmvThreads));
}

// faces_ready: the columns of x are already in f15 / f0 (the s update
// writes them); else they are extracted first
template <class V, bool DOT>
static DotParts spmv_win(hipStream_t st, SweepPlan& P, const V* coef, const V* c15, V* f15,
                         V* f0, const V* x, V* y, bool faces_ready) {
  const int nrun = (P.nsteps + P.spmv_run - 1) / P.spmv_run;
  const int nb = P.ntiles * nrun;
  if (DOT) P.spart.ensure(nb);
  if (!faces_ready) face_cols<V>(st, P.ncell, x, f15, f0);
  hipLaunchKernelGGL((k_box_spmv_win<V, DOT>), dim3(nb), dim3(kThreads), 0, st,
                     (uint32_t)P.ncell, skew_of(P), P.spmv_run, nrun, P.sflag.p, coef, x, f15,
                     f0, c15, y, DOT ? P.spart.p : nullptr);
  WC_LAUNCH_CHECK();
  return DotParts{P.spart.p, nb};
}

void box_spmv(hipStream_t st, SweepPlan& P, const double* x, double* y) {
  if (P.spmv_win) {
    spmv_win<double, false>(st, P, P.coef.p, P.cf15.p, P.xf15.p, P.xf0.p, x, y, false);
    return;
  }
  hipLaunchKernelGGL((k_box_spmv_sym<double, false>), dim3(spmv_blocks(P)), dim3(kSpmvThreads), 0, st,
                     (uint32_t)P.ncell, skew_of(P), P.sflag.p, P.coef.p, x, y, nullptr);
  WC_LAUNCH_CHECK();
}

DotParts box_spmv_dot(hipStream_t st, SweepPlan& P, const double* x, double* y,
                      double* partial, int nblocks, bool faces_ready) {
  if (P.spmv_win)
    return spmv_win<double, true>(st, P, P.coef.p, P.cf15.p, P.xf15.p, P.xf0.p, x, y,
                                  faces_ready);
  hipLaunchKernelGGL((k_box_spmv_sym<double, true>), dim3(nblocks), dim3(kSpmvThreads), 0, st,
                     (uint32_t)P.ncell, skew_of(P), P.sflag.p, P.coef.p, x, y, partial);
  WC_LAUNCH_CHECK();
  return DotParts{partial, nblocks};
}

DotParts box_spmv32_dot(hipStream_t st, SweepPlan& P, const float* x, float* y,
                        double* partial, int nblocks, bool faces_ready) {
  if (P.spmv_win)
    return spmv_win<float, true>(st, P, P.coef32.p, P.cf15_32.p, P.xf15_32.p, P.xf0_32.p, x, y,
                                 faces_ready);
  hipLaunchKernelGGL((k_box_spmv_sym<float, true>), dim3(nblocks), dim3(kSpmvThreads), 0, st,
                     (uint32_t)P.ncell, skew_of(P), P.sflag.p, P.coef32.p, x, y, partial);
  WC_LAUNCH_CHECK();
  return DotParts{partial, nblocks};
}

void box_residual(hipStream_t st, const SweepPlan& P, const double* b, const double* x,
                  float* r, double* partial, int nblocks) {
  hipLaunchKernelGGL(k_box_residual, dim3(nblocks), dim3(256), 0, st, (uint32_t)P.ncell, skew_of(P),
                     P.code.p, P.coef.p, b, x, r, partial);
  WC_LAUNCH_CHECK();
}

void box_gather(hipStream_t st, const SweepPlan& P, const double* dof, double* box) {
  hipLaunchKernelGGL(k_box_gather, dim3(blocks(P.ncell, 256)), dim3(256), 0, st,
                     P.ncell, P.pdof.p, dof, box);
  WC_LAUNCH_CHECK();
}

void box_scatter(hipStream_t st, const SweepPlan& P, const double* box, double* dof) {
  hipLaunchKernelGGL(k_box_scatter, dim3(blocks(P.ncell, 256)), dim3(256), 0, st,
                     P.ncell, P.pdof.p, box, dof);
  WC_LAUNCH_CHECK();
}

bool sweep_aborted(hipStream_t st, SweepPlan& P) {
  int v = 0;
  WC_HIP(hipMemcpyAsync(&v, P.progress.p + 2, sizeof(int), hipMemcpyDeviceToHost, st));
  WC_HIP(hipStreamSynchronize(st));
  if (v != 0) {
    // reported once: the applies do not clear the flag themselves
    WC_HIP(hipMemsetAsync(P.progress.p + 2, 0, sizeof(int), st));
    WC_HIP(hipStreamSynchronize(st));
  }
  return v != 0;
}
void SweepPlan::release() {
  pdof.release();
  code.release();
  scode.release();
  opf.release();
  opb.release();
  sflag.release();
  cf.release();
  cb.release();
  coef.release();
  fin.release();
  rsum.release();
  rb.release();
  sb.release();
  xb.release();
  zf.release();
  zb.release();
  cf32.release();
  cb32.release();
  coef32.release();
  zf32.release();
  zb32.release();
  rb32.release();
  sb32.release();
  bb.release();
  order.release();
  order_sub.release();
  spart.release();
  cf15.release();
  xf15.release();
  xf0.release();
  cf15_32.release();
  xf15_32.release();
  xf0_32.release();
  progress.release();
  tdot.release();
  flags.release();
  valid = false;
}

}  // namespace wc
