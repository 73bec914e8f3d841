// The AMG preconditioner of the reference's AMGPCGSolveSparse on the device
// (AlgebraicMultigrid.h:34-245, GeometricLevelGen.h:125-249).
//
// Levels: the DOFs of a level sit in cells (i, j, k); a coarse DOF is the
// 2x2x2 block (i/2, j/2, k/2) of its children, coarse DOFs are numbered in
// the order their first child appears among the fine DOFs, P = 1 and
// R = 0.125 P^T, and A_c = 0.5 R (A P), while a level holds more than 4096
// DOFs.  Every level's rows keep the 7-point shape (aggregates of face
// neighbours only couple to face-neighbour aggregates), so a coarse row has
// at most 7 entries, kept in ascending column order.
//
// Each value is formed with the reference's operations in its order:
// temp(k, J) adds row k's entries with parent J in column order, A_c(I, J)
// adds 0.5 * 0.125 * temp(k, J) over the children k of I in ascending order
// (SparseMatrix.h:327-348); the smoother sums a row's off-diagonal products
// in column order and divides by the diagonal (AlgebraicMultigrid.h:117-177);
// the residual subtracts the row's products one by one, and R sums its
// 0.125 r over the children in ascending order (:96-105).  So one V-cycle
// gives the reference's z bit for bit; only the PCG's inner products, which
// the reference sums with Eigen's blocked dot, differ by rounding.
//
// The setup (the level graph and the Galerkin products) runs on the device as
// well: a parent per fine DOF from an atomic min of the fine ids per coarse
// cell and a scan of first appearances; children per coarse DOF by their
// position in the 2x2x2 block, sorted.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "wc_ctx.h"

namespace wc {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kCoarsestThreads = 1024;
constexpr int kCoarsestMax = 4096;  // DOFs the single-workgroup coarse solve holds in LDS
constexpr int kCoarsestRows = kCoarsestMax / kCoarsestThreads;  // rows per thread
constexpr int kSlots = 7;           // entries of a colour-ordered row

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

__device__ inline void rows(const DevCsr& A, int64_t i, uint32_t& b, uint32_t& e) {
  if (A.rowptr) {
    b = A.rowptr[i];
    e = A.rowptr[i + 1];
  } else {
    b = (uint32_t)(i * (int64_t)A.stride);
    e = b + A.len[i];
  }
}

// the coarse cell of DOF f, or -1 if its cell lies outside the grid
__device__ inline int64_t cell_key(const int32_t* ijk, int64_t f, int nni, int nnj, int nnk) {
  const int32_t i = ijk[3 * f], j = ijk[3 * f + 1], k = ijk[3 * f + 2];
  if (i < 0 || j < 0 || k < 0 || i / 2 >= nni || j / 2 >= nnj || k / 2 >= nnk) return -1;
  return (int64_t)(i / 2) + (int64_t)(j / 2) * nni + (int64_t)(k / 2) * nni * nnj;
}

// GeometricLevelGen.h:173-184
__global__ void k_amg_red(int64_t n, const int32_t* ijk, uint8_t* red) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  red[i] = ((ijk[3 * i] + ijk[3 * i + 1] + ijk[3 * i + 2]) % 2) == 1;
}

__global__ void k_amg_cmin(int64_t n, const int32_t* ijk, int nni, int nnj, int nnk,
                           uint32_t* cmin, int* err) {
  const int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (f >= n) return;
  const int64_t key = cell_key(ijk, f, nni, nnj, nnk);
  if (key < 0) {
    atomicOr(err, 2);  // a DOF outside the grid (the build then fails)
    return;
  }
  atomicMin(&cmin[key], (uint32_t)f);
}

__global__ void k_amg_first(int64_t n, const int32_t* ijk, int nni, int nnj, int nnk,
                            const uint32_t* cmin, uint32_t* flag) {
  const int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (f >= n) return;
  const int64_t key = cell_key(ijk, f, nni, nnj, nnk);
  flag[f] = key >= 0 && cmin[key] == (uint32_t)f ? 1u : 0u;
}

// parent of each fine DOF, coordinates of each coarse DOF, and each coarse
// DOF's children by their position in the block (GeometricLevelGen.h:138-156)
__global__ void k_amg_parent(int64_t n, const int32_t* ijk, int nni, int nnj, int nnk,
                             const uint32_t* cmin, const uint32_t* first, uint32_t* parent,
                             int32_t* cijk, uint32_t* kids, int* err) {
  const int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (f >= n) return;
  const int64_t key = cell_key(ijk, f, nni, nnj, nnk);
  if (key < 0) {
    parent[f] = 0;
    return;
  }
  const uint32_t m = cmin[key];
  const uint32_t I = first[m];
  parent[f] = I;
  const int i = ijk[3 * f], j = ijk[3 * f + 1], k = ijk[3 * f + 2];
  if (m == (uint32_t)f) {
    cijk[3 * (int64_t)I] = i / 2;
    cijk[3 * (int64_t)I + 1] = j / 2;
    cijk[3 * (int64_t)I + 2] = k / 2;
  }
  // a child's slot is its position in the 2x2x2 block; two DOFs of one cell
  // would claim the same slot (the caller's cells must be distinct)
  const uint32_t was =
      atomicCAS(&kids[8 * (int64_t)I + (i & 1) + 2 * (j & 1) + 4 * (k & 1)], kNone, (uint32_t)f);
  if (was != kNone) atomicOr(err, 8);
}

// A_c = 0.5 R (A P) row by row (SparseMatrix.h:327-348 twice); the children
// are sorted ascending in place (the restriction sums them in that order)
__global__ void k_amg_galerkin(int64_t nc, uint32_t* kids, DevCsr A, const uint32_t* parent,
                               uint32_t* ccol, double* cval, uint32_t* clen, int* err) {
  const int64_t I = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (I >= nc) return;
  uint32_t kd[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) kd[q] = kids[8 * I + q];
  // insertion sort of 8 (kNone last)
#pragma unroll
  for (int a = 1; a < 8; ++a) {
    const uint32_t v = kd[a];
    int b = a - 1;
    while (b >= 0 && kd[b] > v) {
      kd[b + 1] = kd[b];
      --b;
    }
    kd[b + 1] = v;
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) kids[8 * I + q] = kd[q];
  uint32_t jc[7];
  double jv[7];
  int nj = 0;
  for (int q = 0; q < 8; ++q) {
    const uint32_t k = kd[q];
    if (k == kNone) break;
    // temp(k, J) = sum over row k (column order) of 1 * A_kj * 1, J = parent(j)
    uint32_t tc[7];
    double tv[7];
    int nt = 0;
    uint32_t b, e;
    rows(A, k, b, e);
    for (uint32_t s = b; s < e; ++s) {
      const uint32_t J = parent[A.col[s]];
      const double v = 1.0 * A.val[s] * 1.0;
      int t = 0;
      while (t < nt && tc[t] != J) ++t;
      if (t < nt) {
        tv[t] += v;
      } else if (nt < 7) {
        tc[nt] = J;
        tv[nt] = v;
        ++nt;
      } else {
        atomicOr(err, 1);  // more than 7 coarse neighbours: not a 7-point system
      }
    }
    // c(I, J) += 0.5 * 0.125 * temp(k, J); temp's row in ascending J
    for (int t = 0; t < nt; ++t) {
      const double v = 0.5 * 0.125 * tv[t];
      int u = 0;
      while (u < nj && jc[u] != tc[t]) ++u;
      if (u < nj) {
        jv[u] += v;
      } else if (nj < 7) {
        jc[nj] = tc[t];
        jv[nj] = v;
        ++nj;
      } else {
        atomicOr(err, 1);
      }
    }
  }
  // ascending columns
  for (int a = 1; a < nj; ++a) {
    const uint32_t c = jc[a];
    const double v = jv[a];
    int b = a - 1;
    while (b >= 0 && jc[b] > c) {
      jc[b + 1] = jc[b];
      jv[b + 1] = jv[b];
      --b;
    }
    jc[b + 1] = c;
    jv[b + 1] = v;
  }
  for (int a = 0; a < 7; ++a) {
    ccol[7 * I + a] = a < nj ? jc[a] : 0u;
    cval[7 * I + a] = a < nj ? jv[a] : 0.0;
  }
  clen[I] = (uint32_t)nj;
}

// one colour of RBGS_with_pattern (AlgebraicMultigrid.h:117-177)
__device__ inline double rbgs_row(const DevCsr& A, int64_t i, const double* b, const double* x,
                                  int64_t n) {
  uint32_t rb, re;
  rows(A, i, rb, re);
  double sum = 0, diag = 0;
  for (uint32_t s = rb; s < re; ++s) {
    const uint32_t c = A.col[s];
    if (c != (uint32_t)i && c < (uint64_t)n)
      sum += A.val[s] * x[c];
    else
      diag = A.val[s];
  }
  return diag != 0 ? (b[i] - sum) / diag : 0.0;
}

__global__ void k_amg_rbgs(DevCsr A, const uint8_t* red, uint8_t color, const double* b,
                           double* x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= A.n || red[i] != color) return;
  x[i] = rbgs_row(A, i, b, x, A.n);
}

// the coarsest level's 200 sweeps in one workgroup, x in LDS
__global__ __launch_bounds__(kCoarsestThreads) void k_amg_coarsest(DevCsr A, const uint8_t* red,
                                                                  const double* b, double* x,
                                                                  int iters) {
  __shared__ double xs[kCoarsestMax];
  const int n = (int)A.n;
  for (int i = threadIdx.x; i < n; i += kCoarsestThreads) xs[i] = x[i];
  __syncthreads();
  for (int it = 0; it < iters; ++it)
    for (int color = 1; color >= 0; --color) {
      for (int i = threadIdx.x; i < n; i += kCoarsestThreads)
        if (red[i] == color) xs[i] = rbgs_row(A, i, b, xs, n);
      __syncthreads();
    }
  for (int i = threadIdx.x; i < n; i += kCoarsestThreads) x[i] = xs[i];
}

// ---- the colour-ordered rows ----

__global__ void k_amg_redflag(int64_t n, const uint8_t* red, uint32_t* flag) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) flag[i] = red[i];
}

// row i to position p (red rows first); its entries to the planes.  scan
// holds the exclusive scan of the red flags, and scan[n] their count.
__global__ void k_amg_compact(DevCsr A, const uint8_t* red, const uint32_t* scan,
                              uint32_t* crow, uint32_t* ccol, double* cval, uint8_t* clen,
                              int* err) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t n = A.n;
  if (i >= n) return;
  const int64_t nred = scan[n];
  const int64_t p = red[i] ? (int64_t)scan[i] : nred + (i - (int64_t)scan[i]);
  uint32_t rb, re;
  rows(A, i, rb, re);
  if (re - rb > (uint32_t)kSlots) {
    atomicOr(err, 4);  // a longer row: the generic kernels serve this hierarchy
    return;
  }
  crow[p] = (uint32_t)i;
  clen[p] = (uint8_t)(re - rb);
  for (uint32_t s = rb; s < re; ++s) {
    ccol[(int64_t)(s - rb) * n + p] = A.col[s];
    cval[(int64_t)(s - rb) * n + p] = A.val[s];
  }
}

// rbgs_row on a colour-ordered row: the same operations in the same order
__device__ inline double rbgs_crow(const uint32_t* ccol, const double* cval, int len, int64_t n,
                                   int64_t p, uint32_t i, double bi, const double* x) {
  double sum = 0, diag = 0;
  for (int s = 0; s < len; ++s) {
    const uint32_t c = ccol[(int64_t)s * n + p];
    const double v = cval[(int64_t)s * n + p];
    if (c != i && c < (uint64_t)n)
      sum += v * x[c];
    else
      diag = v;
  }
  return diag != 0 ? (bi - sum) / diag : 0.0;
}

// one colour: positions [p0, p1)
__global__ void k_amg_rbgs_c(int64_t n, int64_t p0, int64_t p1, const uint32_t* crow,
                             const uint32_t* ccol, const double* cval, const uint8_t* clen,
                             const double* b, double* x) {
  const int64_t p = p0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= p1) return;
  const uint32_t i = crow[p];
  x[i] = rbgs_crow(ccol, cval, clen[p], n, p, i, b[i], x);
}

// r = b - A x on the colour-ordered rows (as k_amg_resid)
__global__ void k_amg_resid_c(int64_t n, const uint32_t* crow, const uint32_t* ccol,
                              const double* cval, const uint8_t* clen, const double* b,
                              const double* x, double* r) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t i = crow[p];
  const int len = clen[p];
  double v = b[i];
  for (int s = 0; s < len; ++s) v -= cval[(int64_t)s * n + p] * x[ccol[(int64_t)s * n + p]];
  r[i] = v;
}

// y = A x on the colour-ordered rows (multiply, SparseMatrix.h:298-322)
__global__ void k_amg_spmv_c(int64_t n, const uint32_t* crow, const uint32_t* ccol,
                             const double* cval, const uint8_t* clen, const double* x,
                             double* y) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int len = clen[p];
  double acc = 0;
  for (int s = 0; s < len; ++s) acc += cval[(int64_t)s * n + p] * x[ccol[(int64_t)s * n + p]];
  y[crow[p]] = acc;
}

// the coarsest level's sweeps with each thread's rows in registers: thread t
// owns positions t + q * 1024; x in LDS
__global__ __launch_bounds__(kCoarsestThreads) void k_amg_coarsest_c(
    int n, int nred, const uint32_t* crow, const uint32_t* ccol, const double* cval,
    const uint8_t* clen, const double* b, double* x, int iters) {
  __shared__ double xs[kCoarsestMax];
  uint32_t row[kCoarsestRows];
  int len[kCoarsestRows];
  double bi[kCoarsestRows];
  uint32_t col[kCoarsestRows][kSlots];
  double val[kCoarsestRows][kSlots];
#pragma unroll
  for (int q = 0; q < kCoarsestRows; ++q) {
    const int p = (int)threadIdx.x + q * kCoarsestThreads;
    len[q] = 0;
    row[q] = 0;
    bi[q] = 0;
    if (p < n) {
      row[q] = crow[p];
      len[q] = clen[p];
      bi[q] = b[row[q]];
    }
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      col[q][s] = s < len[q] ? ccol[(int64_t)s * n + p] : 0u;
      val[q][s] = s < len[q] ? cval[(int64_t)s * n + p] : 0.0;
    }
  }
  for (int i = threadIdx.x; i < n; i += kCoarsestThreads) xs[i] = x[i];
  __syncthreads();
  for (int it = 0; it < iters; ++it)
    for (int color = 1; color >= 0; --color) {
#pragma unroll
      for (int q = 0; q < kCoarsestRows; ++q) {
        const int p = (int)threadIdx.x + q * kCoarsestThreads;
        if (p < n && (p < nred) == (color == 1)) {
          double sum = 0, diag = 0;
#pragma unroll
          for (int s = 0; s < kSlots; ++s)
            if (s < len[q]) {
              if (col[q][s] != row[q] && col[q][s] < (uint32_t)n)
                sum += val[q][s] * xs[col[q][s]];
              else
                diag = val[q][s];
            }
          xs[row[q]] = diag != 0 ? (bi[q] - sum) / diag : 0.0;
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < n; i += kCoarsestThreads) x[i] = xs[i];
}

// r = b - A x, subtracting the row's products one by one (SparseMatrix.h:404-418)
__global__ void k_amg_resid(DevCsr A, const double* b, const double* x, double* r) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  uint32_t rb, re;
  rows(A, i, rb, re);
  double v = b[i];
  for (uint32_t s = rb; s < re; ++s) v -= A.val[s] * x[A.col[s]];
  r[i] = v;
}

// b_c = R r: 0.125 r over the children in ascending order, from 0
// (multiply, SparseMatrix.h:298-322)
__global__ void k_amg_restrict(int64_t nc, const uint32_t* kids, const double* r, double* bc) {
  const int64_t I = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (I >= nc) return;
  double acc = 0;
  for (int q = 0; q < 8; ++q) {
    const uint32_t k = kids[8 * I + q];
    if (k == kNone) break;
    acc += 0.125 * r[k];
  }
  bc[I] = acc;
}

// x += P x_c (prolongatoin, AlgebraicMultigrid.h:106-116: xx = P x_c from 0,
// then x += 1 * xx)
__global__ void k_amg_prolong(int64_t n, const uint32_t* parent, const double* xc, double* x) {
  const int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (f >= n) return;
  x[f] += 1.0 * (0.0 + 1.0 * xc[parent[f]]);
}

void smooth(hipStream_t st, const AmgHier& H, const AmgLevel& L, const double* b, double* x,
            int iters) {
  for (int it = 0; it < iters; ++it)
    for (int color = 1; color >= 0; --color) {
      if (H.compact) {
        const int64_t p0 = color ? 0 : L.nred, p1 = color ? L.nred : L.n;
        if (p1 > p0)
          hipLaunchKernelGGL(k_amg_rbgs_c, dim3(nblk(p1 - p0, 256)), dim3(256), 0, st, L.n, p0,
                             p1, L.crow.p, L.ccol.p, L.cval.p, L.clen.p, b, x);
      } else {
        hipLaunchKernelGGL(k_amg_rbgs, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.A, L.red.p,
                           (uint8_t)color, b, x);
      }
    }
  WC_LAUNCH_CHECK();
}

}  // namespace

void AmgHier::release() {
  for (int l = 0; l < kAmgMaxLevels; ++l) {
    AmgLevel& L = this->L[l];
    L.col.release();
    L.len.release();
    L.val.release();
    L.ijk.release();
    L.red.release();
    L.parent.release();
    L.kids.release();
    L.x.release();
    L.b.release();
    L.r.release();
    L.crow.release();
    L.ccol.release();
    L.cval.release();
    L.clen.release();
    L.n = 0;
    L.nred = 0;
  }
  nl = 0;
  compact = false;
  cmin.release();
  flag.release();
  scan_tmp.release();
  nred.release();
  err.release();
}

// GeometricLevelGen.h:186-249
void amg_build(hipStream_t st, AmgHier& H, const DevCsr& A, const int32_t* ijk, int ni, int nj,
               int nk) {
  H.nl = 1;
  AmgLevel& L0 = H.L[0];
  L0.n = A.n;
  L0.A = A;
  L0.ni = ni;
  L0.nj = nj;
  L0.nk = nk;
  L0.ijk.ensure(3 * (size_t)A.n);
  WC_HIP(hipMemcpyAsync(L0.ijk.p, ijk, sizeof(int32_t) * 3 * (size_t)A.n,
                        hipMemcpyDeviceToDevice, st));
  H.err.ensure(1);
  H.nred.ensure(kAmgMaxLevels);
  WC_HIP(hipMemsetAsync(H.err.p, 0, sizeof(int), st));
  while (H.L[H.nl - 1].n > 4096 && H.nl < kAmgMaxLevels) {
    AmgLevel& F = H.L[H.nl - 1];
    AmgLevel& C = H.L[H.nl];
    const int nni = (F.ni + 1) / 2, nnj = (F.nj + 1) / 2, nnk = (F.nk + 1) / 2;
    const int64_t ncell = (int64_t)nni * nnj * nnk;
    H.cmin.ensure((size_t)ncell);
    WC_HIP(hipMemsetAsync(H.cmin.p, 0xFF, sizeof(uint32_t) * (size_t)ncell, st));
    hipLaunchKernelGGL(k_amg_cmin, dim3(nblk(F.n, 256)), dim3(256), 0, st, F.n, F.ijk.p, nni,
                       nnj, nnk, H.cmin.p, H.err.p);
    H.flag.ensure((size_t)F.n + 1);
    hipLaunchKernelGGL(k_amg_first, dim3(nblk(F.n, 256)), dim3(256), 0, st, F.n, F.ijk.p, nni,
                       nnj, nnk, H.cmin.p, H.flag.p);
    WC_LAUNCH_CHECK();
    WC_HIP(hipMemsetAsync(H.flag.p + F.n, 0, sizeof(uint32_t), st));
    // (a DOF outside the grid sets err and is skipped by every kernel; the
    // build fails at its end)
    const int64_t nc = exclusive_scan(st, H.scan_tmp, H.flag.p, F.n + 1);  // synchronizes
    if (nc == 0 || nc >= F.n) break;  // no reduction (the reference would not stop)
    F.parent.ensure((size_t)F.n);
    C.n = nc;
    C.ni = nni;
    C.nj = nnj;
    C.nk = nnk;
    C.ijk.ensure(3 * (size_t)nc);
    C.kids.ensure(8 * (size_t)nc);
    WC_HIP(hipMemsetAsync(C.kids.p, 0xFF, sizeof(uint32_t) * 8 * (size_t)nc, st));
    hipLaunchKernelGGL(k_amg_parent, dim3(nblk(F.n, 256)), dim3(256), 0, st, F.n, F.ijk.p, nni,
                       nnj, nnk, H.cmin.p, H.flag.p, F.parent.p, C.ijk.p, C.kids.p, H.err.p);
    C.col.ensure(7 * (size_t)nc);
    C.val.ensure(7 * (size_t)nc);
    C.len.ensure((size_t)nc);
    hipLaunchKernelGGL(k_amg_galerkin, dim3(nblk(nc, 256)), dim3(256), 0, st, nc, C.kids.p, F.A,
                       F.parent.p, C.col.p, C.val.p, C.len.p, H.err.p);
    WC_LAUNCH_CHECK();
    DevCsr Ac;
    Ac.n = nc;
    Ac.stride = 7;
    Ac.len = C.len.p;
    Ac.col = C.col.p;
    Ac.val = C.val.p;
    Ac.nnz_slots = 7 * nc;
    C.A = Ac;
    H.nl++;
  }
  for (int l = 0; l < H.nl; ++l) {
    AmgLevel& L = H.L[l];
    L.red.ensure((size_t)L.n);
    hipLaunchKernelGGL(k_amg_red, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.n, L.ijk.p, L.red.p);
    if (l > 0) {
      L.x.ensure((size_t)L.n);
      L.b.ensure((size_t)L.n);
    }
    if (l + 1 < H.nl) L.r.ensure((size_t)L.n);
    // colour-ordered rows: red positions by a scan of the red flags
    H.flag.ensure((size_t)L.n + 1);
    hipLaunchKernelGGL(k_amg_redflag, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.n, L.red.p,
                       H.flag.p);
    WC_LAUNCH_CHECK();
    WC_HIP(hipMemsetAsync(H.flag.p + L.n, 0, sizeof(uint32_t), st));
    // the count lands in flag[n]; kept per level, read back with the error
    // flags below (one host wait for the whole colour split)
    exclusive_scan_async(st, H.scan_tmp, H.flag.p, L.n + 1);
    WC_HIP(hipMemcpyAsync(H.nred.p + l, H.flag.p + L.n, sizeof(uint32_t),
                          hipMemcpyDeviceToDevice, st));
    L.crow.ensure((size_t)L.n);
    L.clen.ensure((size_t)L.n);
    L.ccol.ensure((size_t)kSlots * L.n);
    L.cval.ensure((size_t)kSlots * L.n);
    hipLaunchKernelGGL(k_amg_compact, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.A, L.red.p,
                       H.flag.p, L.crow.p, L.ccol.p, L.cval.p, L.clen.p, H.err.p);
    WC_LAUNCH_CHECK();
  }
  int herr = 0;
  uint32_t nred[kAmgMaxLevels] = {};
  WC_HIP(hipMemcpyAsync(&herr, H.err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  WC_HIP(hipMemcpyAsync(nred, H.nred.p, sizeof(uint32_t) * H.nl, hipMemcpyDeviceToHost, st));
  WC_HIP(hipStreamSynchronize(st));
  if (herr & 2)
    throw HipError{hipErrorInvalidValue, "AMG levels: a DOF cell lies outside the grid",
                   __FILE__, __LINE__};
  if (herr & 1)
    throw HipError{hipErrorInvalidValue, "AMG levels: a coarse row has more than 7 entries",
                   __FILE__, __LINE__};
  if (herr & 8)
    throw HipError{hipErrorInvalidValue, "AMG levels: two DOFs share a cell", __FILE__, __LINE__};
  for (int l = 0; l < H.nl; ++l) H.L[l].nred = nred[l];
  H.compact = (herr & 4) == 0;
}

bool amg_spmv(hipStream_t st, AmgHier& H, const double* x, double* y) {
  if (!H.compact || H.nl < 1) return false;
  const AmgLevel& L = H.L[0];
  if (L.n > 0)
    hipLaunchKernelGGL(k_amg_spmv_c, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.n, L.crow.p,
                       L.ccol.p, L.cval.p, L.clen.p, x, y);
  WC_LAUNCH_CHECK();
  return true;
}

// AlgebraicMultigrid.h:179-245
void amg_vcycle(hipStream_t st, AmgHier& H, const double* b, double* z) {
  const int nl = H.nl;
  const double* bl[kAmgMaxLevels];
  double* xl[kAmgMaxLevels];
  for (int l = 0; l < nl; ++l) {
    bl[l] = l == 0 ? b : H.L[l].b.p;
    xl[l] = l == 0 ? z : H.L[l].x.p;
    WC_HIP(hipMemsetAsync(xl[l], 0, sizeof(double) * (size_t)H.L[l].n, st));
  }
  for (int l = 0; l + 1 < nl; ++l) {
    AmgLevel& L = H.L[l];
    smooth(st, H, L, bl[l], xl[l], 4);
    if (H.compact)
      hipLaunchKernelGGL(k_amg_resid_c, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.n, L.crow.p,
                         L.ccol.p, L.cval.p, L.clen.p, bl[l], xl[l], L.r.p);
    else
      hipLaunchKernelGGL(k_amg_resid, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.A, bl[l], xl[l],
                         L.r.p);
    hipLaunchKernelGGL(k_amg_restrict, dim3(nblk(H.L[l + 1].n, 256)), dim3(256), 0, st,
                       H.L[l + 1].n, H.L[l + 1].kids.p, L.r.p, H.L[l + 1].b.p);
    WC_LAUNCH_CHECK();
  }
  AmgLevel& T = H.L[nl - 1];
  if (T.n <= kCoarsestMax) {
    if (T.n > 0 && H.compact)
      hipLaunchKernelGGL(k_amg_coarsest_c, dim3(1), dim3(kCoarsestThreads), 0, st, (int)T.n,
                         (int)T.nred, T.crow.p, T.ccol.p, T.cval.p, T.clen.p, bl[nl - 1],
                         xl[nl - 1], 200);
    else if (T.n > 0)
      hipLaunchKernelGGL(k_amg_coarsest, dim3(1), dim3(kCoarsestThreads), 0, st, T.A, T.red.p,
                         bl[nl - 1], xl[nl - 1], 200);
    WC_LAUNCH_CHECK();
  } else {
    smooth(st, H, T, bl[nl - 1], xl[nl - 1], 200);
  }
  for (int l = nl - 2; l >= 0; --l) {
    AmgLevel& L = H.L[l];
    hipLaunchKernelGGL(k_amg_prolong, dim3(nblk(L.n, 256)), dim3(256), 0, st, L.n, L.parent.p,
                       xl[l + 1], xl[l]);
    WC_LAUNCH_CHECK();
    smooth(st, H, L, bl[l], xl[l], 4);
  }
}

}  // namespace wc
