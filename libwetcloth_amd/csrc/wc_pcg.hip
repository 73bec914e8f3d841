// MIC(0)-preconditioned conjugate gradients on gfx950
// (robertbridson::PCGSolver, PCGSolver/PCGSolver.h:97-330).
//
// The factorization and both triangular solves are level-scheduled: rows are
// grouped by their depth in the lower-triangular dependency graph (for the
// pressure grid that is the i+j+k hyperplane) and each level is one launch in
// which every row only reads rows of earlier levels.  Each row is computed in
// gather ("left-looking") form, applying exactly the updates the reference's
// column-sweeping loops apply to it and in the same order (ascending column),
// so the factor and the preconditioned vector are bitwise the sequential
// ones.  Inner products are two-stage block reductions in FP64 with a fixed
// order (deterministic, but not the reference's SIMD order).
#include <math.h>

#include <algorithm>
#include <chrono>

#include "wc_ctx.h"

namespace wc {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kRedThreads = 256;
constexpr int kMaxUpper = 16;  // upper entries per row the factor supports

// scalar slots in PcgWork::scal
enum { S_RHO = 0, S_ALPHA = 1, S_BETA = 2, S_RES = 3, S_DOT = 4, S_RHO2 = 5, S_NSCAL = 8 };

__device__ inline void row_range(const DevCsr& A, int64_t i, uint32_t& b,
                                 uint32_t& e) {
  if (A.rowptr) {
    b = A.rowptr[i];
    e = A.rowptr[i + 1];
  } else {
    b = A.start ? A.start[i] : (uint32_t)(i * (int64_t)A.stride);
    e = b + A.len[i];
  }
}

// first entry of row i with column > i
__device__ inline uint32_t upper_begin(const DevCsr& A, int64_t i, uint32_t b,
                                       uint32_t e) {
  while (b < e && A.col[b] <= (uint32_t)i) ++b;
  return b;
}

__global__ void k_tpos_adiag(DevCsr A, uint32_t* tpos, double* adiag) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  uint32_t b, e;
  row_range(A, i, b, e);
  double d = 0.0;
  for (uint32_t q = b; q < e; ++q) {
    const uint32_t c = A.col[q];
    uint32_t t = kNone;
    if (c == (uint32_t)i) d = A.val[q];
    if (c < (uint32_t)i) {
      uint32_t kb, ke;
      row_range(A, c, kb, ke);
      for (uint32_t r = kb; r < ke; ++r)
        if (A.col[r] == (uint32_t)i) {
          t = r;
          break;
        }
    }
    tpos[q] = t;
  }
  adiag[i] = d;
}

__global__ void k_level_relax(DevCsr A, int32_t* level, int* changed) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  uint32_t b, e;
  row_range(A, i, b, e);
  int32_t lv = 0;
  for (uint32_t q = b; q < e; ++q) {
    const uint32_t c = A.col[q];
    if (c < (uint32_t)i) lv = max(lv, level[c] + 1);
  }
  if (lv != level[i]) {
    level[i] = lv;
    *changed = 1;
  }
}

__global__ void k_level_max(const int32_t* level, int64_t n, int* out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicMax(out, level[i]);
}

__global__ void k_level_count(const int32_t* level, int64_t n, int32_t base,
                              uint32_t* count) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicAdd(&count[level[i] - base], 1u);
}

__global__ void k_level_scatter(const int32_t* level, int64_t n, int32_t base,
                                uint32_t* cursor, uint32_t* list) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t pos = atomicAdd(&cursor[level[i] - base], 1u);
  list[pos] = (uint32_t)i;
}

// factor_modified_incomplete_cholesky0 (PCGSolver.h:97-205), one level.
// Column j collects, in ascending k, every update the right-looking loop of
// column k applies to it: the diagonal -= L_jk*L_jk, the fill-in matches
// L_rj -= L_jk*L_rk, and the diagonal -= tau*L_jk*missing.
__global__ void k_mic0_level(DevCsr A, const uint32_t* tpos,
                             const double* adiag, const uint32_t* list,
                             uint32_t count, double tau, double mdr,
                             double* invdiag, double* lv, int* overflow) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int64_t j = list[t];
  uint32_t jb, je;
  row_range(A, j, jb, je);
  const uint32_t ju = upper_begin(A, j, jb, je);
  const int nu = (int)(je - ju);
  if (nu > kMaxUpper) {
    *overflow = 1;
    return;
  }
  double lvj[kMaxUpper];
  for (int u = 0; u < nu; ++u) lvj[u] = A.val[ju + u];
  const double aj = adiag[j];
  double d = aj;
  for (uint32_t q = jb; q < ju; ++q) {
    const uint32_t k = A.col[q];
    if (k >= (uint32_t)j) break;  // the diagonal
    const uint32_t pk = tpos[q];
    if (pk == kNone) continue;
    if (adiag[k] == 0) continue;  // null column k is skipped (PCGSolver.h:141)
    uint32_t kb, ke;
    row_range(A, k, kb, ke);
    const uint32_t c0 = upper_begin(A, k, kb, ke);
    const double multiplier = lv[pk];
    double missing = 0;
    uint32_t a = c0;
    uint32_t bb = jb;
    while (a < ke && A.col[a] < (uint32_t)j) {
      while (bb < je) {
        if (A.col[bb] < A.col[a])
          ++bb;
        else if (A.col[bb] == A.col[a])
          break;
        else {
          missing += lv[a];
          break;
        }
      }
      ++a;
    }
    if (a < ke && A.col[a] == (uint32_t)j) d -= multiplier * lv[a];
    ++a;
    int ub = 0;
    while (a < ke && ub < nu) {
      if (A.col[ju + ub] < A.col[a])
        ++ub;
      else if (A.col[ju + ub] == A.col[a]) {
        lvj[ub] -= multiplier * lv[a];
        ++a;
        ++ub;
      } else {
        missing += lv[a];
        ++a;
      }
    }
    while (a < ke) {
      missing += lv[a];
      ++a;
    }
    d -= tau * multiplier * missing;
  }
  if (aj == 0) {
    invdiag[j] = d;
    for (int u = 0; u < nu; ++u) lv[ju + u] = lvj[u];
    return;
  }
  double inv;
  if (d < mdr * aj)
    inv = 1 / sqrt(aj);
  else
    inv = 1 / sqrt(d);
  invdiag[j] = inv;
  for (int u = 0; u < nu; ++u) lv[ju + u] = lvj[u] * inv;
}

// solve_lower (PCGSolver.h:211-223) for one level: z_j = (r_j - sum over
// ascending k of L_jk z_k) * invdiag_j.
__global__ void k_lower_level(DevCsr A, const uint32_t* tpos,
                              const double* invdiag, const double* lv,
                              const uint32_t* list, uint32_t count,
                              const double* r, double* z) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int64_t j = list[t];
  uint32_t jb, je;
  row_range(A, j, jb, je);
  double acc = r[j];
  for (uint32_t q = jb; q < je; ++q) {
    const uint32_t k = A.col[q];
    if (k >= (uint32_t)j) break;
    const uint32_t pk = tpos[q];
    if (pk == kNone) continue;
    acc -= lv[pk] * z[k];
  }
  z[j] = acc * invdiag[j];
}

// solve_lower_transpose_in_place (PCGSolver.h:226-239) for one level.
__global__ void k_upper_level(DevCsr A, const double* invdiag,
                              const double* lv, const uint32_t* list,
                              uint32_t count, double* z) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int64_t j = list[t];
  uint32_t jb, je;
  row_range(A, j, jb, je);
  double acc = z[j];
  for (uint32_t q = upper_begin(A, j, jb, je); q < je; ++q)
    acc -= lv[q] * z[A.col[q]];
  z[j] = acc * invdiag[j];
}

// multiply(FixedSparseMatrix) (SparseMatrix.h:298-322)
__global__ void k_spmv(DevCsr A, const double* x, double* y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  uint32_t b, e;
  row_range(A, i, b, e);
  double acc = 0;
  for (uint32_t q = b; q < e; ++q) acc += A.val[q] * x[A.col[q]];
  y[i] = acc;
}

__device__ inline double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

__device__ inline double block_max(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  return sh[0];
}

// FP64 block partials of a . b (FP64 products and sums for FP32 vectors too)
template <class V>
__global__ __launch_bounds__(kRedThreads) void k_dot_partial(
    const V* a, const V* b, int64_t n, double* partial) {
  __shared__ double sh[kRedThreads];
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads)
    acc += (double)a[i] * (double)b[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// modes of the final reduction
enum { R_SET_RHO = 0, R_ALPHA = 1, R_BETA = 2, R_DOT = 3 };

__global__ __launch_bounds__(kRedThreads) void k_dot_final(
    const double* partial, int nb, double* scal, int mode) {
  __shared__ double sh[kRedThreads];
  double acc = 0;
  for (int i = threadIdx.x; i < nb; i += kRedThreads) acc += partial[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    if (mode == R_SET_RHO) {
      scal[S_RHO] = s;
    } else if (mode == R_ALPHA) {
      scal[S_ALPHA] = scal[S_RHO] / s;  // PCGSolver.h:291
    } else if (mode == R_BETA) {
      scal[S_BETA] = s / scal[S_RHO];  // PCGSolver.h:301
      scal[S_RHO] = s;
    } else {
      scal[S_DOT] = s;
    }
  }
}

__global__ __launch_bounds__(kRedThreads) void k_absmax_partial(
    const double* a, int64_t n, double* partial) {
  __shared__ double sh[kRedThreads];
  double m = 0;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads)
    m = fmax(m, fabs(a[i]));
  const double s = block_max(m, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedThreads) void k_absmax_final(
    const double* partial, int nb, double* scal) {
  __shared__ double sh[kRedThreads];
  double m = 0;
  for (int i = threadIdx.x; i < nb; i += kRedThreads) m = fmax(m, partial[i]);
  const double s = block_max(m, sh);
  if (threadIdx.x == 0) scal[S_RES] = s;
}

// result += s*alpha; r += z*(-alpha) (PCGSolver.h:292-293), plus |r|_inf
// block partials (:294).
__global__ __launch_bounds__(kRedThreads) void k_update_xr(
    double* x, double* r, const double* s, const double* z, int64_t n,
    const double* scal, double* partial) {
  __shared__ double sh[kRedThreads];
  const double alpha = scal[S_ALPHA];
  const double nalpha = -alpha;
  double m = 0;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads) {
    x[i] += s[i] * alpha;
    const double ri = r[i] + z[i] * nalpha;
    r[i] = ri;
    m = fmax(m, fabs(ri));
  }
  const double mm = block_max(m, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = mm;
}

// s = z + s*beta (add_scaled then swap, PCGSolver.h:302-303); FP32 vectors
// are updated in FP64 and rounded
template <class V>
__global__ void k_update_s(V* s, const V* z, int64_t n, const double* scal) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  s[i] = (V)((double)z[i] + (double)s[i] * scal[S_BETA]);
}

// The box loops' update kernels take alpha and beta straight from the block
// partials of the inner products: every block repeats the final reduction
// with the thread mapping and tree of k_dot_final, so each computes the same
// value k_dot_final would (rho_old / s . z, and z . r / rho_old).

// x += s alpha, r -= z alpha (FP64 r and z, or FP32 formed in FP64), |r|_inf
// block partials into rpart (mapped host memory, which the host polls; the
// sweeps' abort flag rides in block 0's entry); and the next forward
// sweep's output set to the all-ones sentinel (reset, the same slots),
// which saves the sweep's own fill launch.
// N values of T, aligned to their size (N defaults to 16 bytes' worth):
// one vector access per lane instead of N scalar ones
template <class T, int NN = 16 / (int)sizeof(T)>
struct alignas(NN * sizeof(T) < 16 ? NN * sizeof(T) : 16) Pack {
  static constexpr int N = NN;
  T v[NN];
};

__device__ inline void set_all_ones(double* p) {
  *reinterpret_cast<unsigned long long*>(p) = ~0ull;
}
__device__ inline void set_all_ones(float* p) { *reinterpret_cast<unsigned int*>(p) = ~0u; }

template <class V>
__global__ __launch_bounds__(kRedThreads) void k_update_xr_fused(
    double* x, V* r, const V* s, const V* z, int64_t n, const double* dpart, int nb,
    const double* scal, int rho_slot, double* rpart, const int* flag, V* reset) {
  __shared__ double sh[kRedThreads];
  double acc = 0;
  for (int i = threadIdx.x; i < nb; i += kRedThreads) acc += dpart[i];
  const double sz = block_sum(acc, sh);
  const double alpha = scal[rho_slot] / sz;  // PCGSolver.h:291
  const double nalpha = -alpha;
  __syncthreads();  // every thread has read the sum before sh is reused
  double m = 0;
  auto one = [&](double& xi, V& ri_, V si, V zi) {
    if constexpr (sizeof(V) == sizeof(double)) {
      xi += si * alpha;
      const double ri = ri_ + zi * nalpha;
      ri_ = ri;
      m = fmax(m, fabs(ri));
    } else {
      xi += (double)si * alpha;
      const V ri = (V)((double)ri_ - (double)zi * alpha);
      ri_ = ri;
      m = fmax(m, (double)fabsf(ri));
    }
  };
  // FP64: two values (16 bytes) of each vector per lane and pass.  FP32:
  // one (two or four FP32 values per lane, with 16 or 32 bytes of x,
  // measured slower: 46.8 and 50.3 against 44.8 us at 256^3)
  constexpr int N = sizeof(V) == sizeof(double) ? 2 : 1;
  using PV = Pack<V, N>;
  const int64_t nv = n / N;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < nv;
       i += (int64_t)gridDim.x * kRedThreads) {
    Pack<double, N> xv = reinterpret_cast<const Pack<double, N>*>(x)[i];
    PV rv = reinterpret_cast<const PV*>(r)[i];
    const PV sv = reinterpret_cast<const PV*>(s)[i];
    const PV zv = reinterpret_cast<const PV*>(z)[i];
#pragma unroll
    for (int k = 0; k < N; ++k) one(xv.v[k], rv.v[k], sv.v[k], zv.v[k]);
    reinterpret_cast<Pack<double, N>*>(x)[i] = xv;
    reinterpret_cast<PV*>(r)[i] = rv;
    PV ones;
#pragma unroll
    for (int k = 0; k < N; ++k) set_all_ones(&ones.v[k]);
    reinterpret_cast<PV*>(reset)[i] = ones;
  }
  for (int64_t i = nv * N + blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads) {
    one(x[i], r[i], s[i], z[i]);
    set_all_ones(&reset[i]);
  }
  const double mm = block_max(m, sh);
  if (threadIdx.x == 0) {
    // the host polls rpart, whose entries it set to -1: each block stores
    // its |r| max (>= 0), block 0 stores -2 instead when the sweeps' abort
    // flag is set (one store, so nothing has to be ordered before it)
    rpart[blockIdx.x] = blockIdx.x == 0 && *flag ? -2.0 : mm;
  }
}

// beta = (z . r) / rho_old (PCGSolver.h:301), s = z + s beta; the new rho is
// written to the other rho slot, which no block reads here.  With f15 / f0,
// also the x-face columns of s that the windowed SpMV reads: s at the slots
// with (t & 15) == 15 to f15[t >> 4], with (t & 15) == 0 to f0[t >> 4].
template <class V>
__global__ __launch_bounds__(kRedThreads) void k_update_s_fused(
    V* s, const V* z, int64_t n, const double* dpart, int nb, double* scal, int rho_old,
    int rho_new, V* f15, V* f0) {
  __shared__ double sh[kRedThreads];
  double acc = 0;
  for (int i = threadIdx.x; i < nb; i += kRedThreads) acc += dpart[i];
  const double zr = block_sum(acc, sh);
  const double beta = zr / scal[rho_old];
  if (blockIdx.x == 0 && threadIdx.x == 0) scal[rho_new] = zr;
  using PV = Pack<V>;
  constexpr int N = PV::N;
  const int64_t nv = n / N;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < nv;
       i += (int64_t)gridDim.x * kRedThreads) {
    const PV zv = reinterpret_cast<const PV*>(z)[i];
    PV sv = reinterpret_cast<const PV*>(s)[i];
#pragma unroll
    for (int k = 0; k < N; ++k) sv.v[k] = (V)((double)zv.v[k] + (double)sv.v[k] * beta);
    reinterpret_cast<PV*>(s)[i] = sv;
    if (f15) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const int64_t t = i * N + k;
        if ((t & 15) == 15) f15[t >> 4] = sv.v[k];
        if ((t & 15) == 0) f0[t >> 4] = sv.v[k];
      }
    }
  }
  for (int64_t i = nv * N + blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads) {
    const V v = (V)((double)z[i] + (double)s[i] * beta);
    s[i] = v;
    if (f15 && (i & 15) == 15) f15[i >> 4] = v;
    if (f15 && (i & 15) == 0) f0[i >> 4] = v;
  }
}

inline unsigned blocks_for(int64_t n, int t) {
  return (unsigned)((n + t - 1) / t);
}

}  // namespace

hipEvent_t KTimer::mark(hipStream_t st) {
  hipEvent_t e;
  if (!spare.empty()) {
    e = spare.back();
    spare.pop_back();
  } else {
    WC_HIP(hipEventCreate(&e));
  }
  WC_HIP(hipEventRecord(e, st));
  return e;
}

void KTimer::harvest() {
  for (const Rec& r : pending) {
    float t = 0.f;
    WC_HIP(hipEventElapsedTime(&t, r.a, r.b));
    ms[r.kind] += t;
    launches[r.kind] += 1;
    units[r.kind] += r.units;
    spare.push_back(r.a);
    spare.push_back(r.b);
  }
  pending.clear();
}

void KTimer::harvest_ready() {
  // the pairs were recorded on one stream, so they complete in order
  size_t k = 0;
  for (; k < pending.size(); ++k) {
    const hipError_t q = hipEventQuery(pending[k].b);
    if (q == hipErrorNotReady) break;
    WC_HIP(q);
  }
  if (k == 0) return;
  std::vector<Rec> rest(pending.begin() + (ptrdiff_t)k, pending.end());
  pending.resize(k);
  harvest();
  pending.swap(rest);
}

void KTimer::reset() {
  for (const Rec& r : pending) {
    spare.push_back(r.a);
    spare.push_back(r.b);
  }
  pending.clear();
  for (int k = 0; k < COUNT; ++k) {
    ms[k] = 0;
    launches[k] = 0;
    units[k] = 0;
  }
}

void KTimer::release() {
  reset();
  for (hipEvent_t e : spare) (void)hipEventDestroy(e);
  spare.clear();
}

void PcgWork::release() {
  timer.release();
  sweep.release();
  invdiag.release();
  adiag.release();
  flags.release();
  lv.release();
  r.release();
  z.release();
  s.release();
  x.release();
  partial.release();
  scal.release();
  tpos.release();
  level.release();
  lvl_count.release();
  lvl_cursor.release();
  lvl_list.release();
  if (host_scal && host_scal != host_scal_own) (void)hipHostFree(host_scal);
  host_scal = nullptr;
  if (host_part) (void)hipHostFree(host_part);
  host_part = nullptr;
  host_part_dev = nullptr;
}

static size_t slot_count(const DevCsr& A) { return (size_t)A.nnz_slots; }

void pcg_prepare(hipStream_t st, PcgWork& w, const DevCsr& A,
                 const int32_t* level, int32_t level_span) {
  const int64_t n = A.n;
  w.tpos.ensure(slot_count(A));
  w.invdiag.ensure(n);
  w.adiag.ensure(n);
  w.flags.ensure(4);
  w.z.ensure(n);
  w.lv.ensure(slot_count(A));
  w.r.ensure(n);
  w.s.ensure(n);
  w.x.ensure(n);
  w.red_blocks = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + kRedThreads - 1) / kRedThreads));
  w.partial.ensure(w.red_blocks);
  w.scal.ensure(S_NSCAL);
  w.level_in = level;
  w.level_span = level_span;
  w.levels_ready = false;
  w.nlevels = 0;
  if (w.no_alloc) {
    // lent buffers only: scalars come back through pageable host memory
    // (the copy is followed by a stream synchronisation in read_scal); the
    // box loops, which need mapped memory, are not used
    static_assert(S_NSCAL <= sizeof(w.host_scal_own) / sizeof(double), "scalar slots");
    w.host_scal = w.host_scal_own;
    return;
  }
  // written by the update kernels of the box loops, read by the host
  if (!w.host_part) {
    WC_HIP(hipHostMalloc((void**)&w.host_part, sizeof(double) * 1024,
                         hipHostMallocMapped | hipHostMallocCoherent));
    WC_HIP(hipHostGetDevicePointer((void**)&w.host_part_dev, w.host_part, 0));
  }
  if (!w.host_scal) WC_HIP(hipHostMalloc((void**)&w.host_scal, sizeof(double) * S_NSCAL));
}

void pcg_prepare_vectors(hipStream_t st, PcgWork& w, int64_t n) {
  w.r.ensure(n);
  w.z.ensure(n);
  w.s.ensure(n);
  w.red_blocks = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + kRedThreads - 1) / kRedThreads));
  w.partial.ensure(w.red_blocks);
  w.scal.ensure(S_NSCAL);
  if (!w.host_scal) WC_HIP(hipHostMalloc((void**)&w.host_scal, sizeof(double) * S_NSCAL));
  w.levels_ready = false;
  w.nlevels = 0;
}

// The level schedule of the level-scheduled kernels (built on first use, so
// systems solved with the pipelined sweeps never pay for it).
static void pcg_levels(hipStream_t st, PcgWork& w, const DevCsr& A) {
  if (w.levels_ready) return;
  w.levels_ready = true;
  const int64_t n = A.n;
  if (n == 0) {
    w.nlevels = 0;
    return;
  }
  hipLaunchKernelGGL(k_tpos_adiag, dim3(blocks_for(n, 256)), dim3(256), 0, st, A,
                     w.tpos.p, w.adiag.p);
  WC_LAUNCH_CHECK();

  const int32_t* lev = w.level_in;
  int32_t span = w.level_span;
  if (!lev) {
    w.level.ensure(n);
    WC_HIP(hipMemsetAsync(w.level.p, 0, sizeof(int32_t) * n, st));
    int* changed = w.flags.p;
    int hchanged = 1;
    int64_t iters = 0;
    while (hchanged) {
      WC_HIP(hipMemsetAsync(changed, 0, sizeof(int), st));
      hipLaunchKernelGGL(k_level_relax, dim3(blocks_for(n, 256)), dim3(256), 0,
                         st, A, w.level.p, changed);
      WC_LAUNCH_CHECK();
      WC_HIP(hipMemcpyAsync(&hchanged, changed, sizeof(int),
                            hipMemcpyDeviceToHost, st));
      WC_HIP(hipStreamSynchronize(st));
      if (++iters > n + 1) break;
    }
    // relaxation in place can converge in fewer passes than the depth
    WC_HIP(hipMemsetAsync(w.flags.p + 2, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_level_max, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                       w.level.p, n, w.flags.p + 2);
    WC_LAUNCH_CHECK();
    int hmax = 0;
    WC_HIP(hipMemcpyAsync(&hmax, w.flags.p + 2, sizeof(int),
                          hipMemcpyDeviceToHost, st));
    WC_HIP(hipStreamSynchronize(st));
    span = hmax + 1;
    lev = w.level.p;
  }
  w.lvl_count.ensure(span);
  w.lvl_cursor.ensure(span);
  w.lvl_list.ensure(n);
  WC_HIP(hipMemsetAsync(w.lvl_count.p, 0, sizeof(uint32_t) * span, st));
  hipLaunchKernelGGL(k_level_count, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                     lev, n, 0, w.lvl_count.p);
  WC_LAUNCH_CHECK();
  std::vector<uint32_t> cnt(span);
  WC_HIP(hipMemcpyAsync(cnt.data(), w.lvl_count.p, sizeof(uint32_t) * span,
                        hipMemcpyDeviceToHost, st));
  WC_HIP(hipStreamSynchronize(st));
  w.lvl_start.clear();
  w.lvl_size.clear();
  std::vector<uint32_t> start(span);
  uint32_t acc = 0;
  for (int32_t l = 0; l < span; ++l) {
    start[l] = acc;
    if (cnt[l]) {
      w.lvl_start.push_back(acc);
      w.lvl_size.push_back(cnt[l]);
    }
    acc += cnt[l];
  }
  w.nlevels = (int)w.lvl_size.size();
  WC_HIP(hipMemcpyAsync(w.lvl_cursor.p, start.data(), sizeof(uint32_t) * span,
                        hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_level_scatter, dim3(blocks_for(n, 256)), dim3(256), 0,
                     st, lev, n, 0, w.lvl_cursor.p, w.lvl_list.p);
  WC_LAUNCH_CHECK();
  // the host vector 'start' must outlive the async copy
  WC_HIP(hipStreamSynchronize(st));
}

void pcg_factor(hipStream_t st, PcgWork& w, const DevCsr& A, double tau,
                double mdr) {
  if (A.n == 0) return;
  pcg_levels(st, w, A);
  int* overflow = w.flags.p + 1;
  WC_HIP(hipMemsetAsync(overflow, 0, sizeof(int), st));
  for (int l = 0; l < w.nlevels; ++l) {
    const uint32_t cnt = w.lvl_size[l];
    hipLaunchKernelGGL(k_mic0_level, dim3(blocks_for(cnt, 64)), dim3(64), 0, st,
                       A, w.tpos.p, w.adiag.p, w.lvl_list.p + w.lvl_start[l], cnt,
                       tau, mdr, w.invdiag.p, w.lv.p, overflow);
    WC_LAUNCH_CHECK();
  }
  int hov = 0;
  WC_HIP(hipMemcpyAsync(&hov, overflow, sizeof(int), hipMemcpyDeviceToHost, st));
  WC_HIP(hipStreamSynchronize(st));
  if (hov) throw HipError{hipErrorInvalidValue, "MIC(0): a row has more than 16 entries above the diagonal", __FILE__, __LINE__};
}

void pcg_apply(hipStream_t st, PcgWork& w, const DevCsr& A, const double* r,
               double* z) {
  pcg_levels(st, w, A);
  for (int l = 0; l < w.nlevels; ++l) {
    const uint32_t cnt = w.lvl_size[l];
    hipLaunchKernelGGL(k_lower_level, dim3(blocks_for(cnt, 64)), dim3(64), 0, st,
                       A, w.tpos.p, w.invdiag.p, w.lv.p,
                       w.lvl_list.p + w.lvl_start[l], cnt, r, z);
    WC_LAUNCH_CHECK();
  }
  for (int l = w.nlevels - 1; l >= 0; --l) {
    const uint32_t cnt = w.lvl_size[l];
    hipLaunchKernelGGL(k_upper_level, dim3(blocks_for(cnt, 64)), dim3(64), 0, st,
                       A, w.invdiag.p, w.lv.p, w.lvl_list.p + w.lvl_start[l],
                       cnt, z);
    WC_LAUNCH_CHECK();
  }
}

void pcg_spmv(hipStream_t st, const DevCsr& A, const double* x, double* y) {
  if (A.n == 0) return;
  hipLaunchKernelGGL(k_spmv, dim3(blocks_for(A.n, 256)), dim3(256), 0, st, A,
                     x, y);
  WC_LAUNCH_CHECK();
}

static double read_scal(hipStream_t st, PcgWork& w, int slot) {
  WC_HIP(hipMemcpyAsync(w.host_scal + slot, w.scal.p + slot, sizeof(double),
                        hipMemcpyDeviceToHost, st));
  WC_HIP(hipStreamSynchronize(st));
  if (!w.timer.pending.empty()) w.timer.harvest();
  return w.host_scal[slot];
}

// The residual of the last update: the max of its block partials, which
// the update kernel wrote into mapped host memory with the sweeps' abort
// flag, once the event recorded after it has completed (work enqueued after
// the event may still be running).
// The update kernel's |r|_inf partials arrive in mapped host memory, each
// block's over the -1 the host left there: a max (>= 0), or -2 from block 0
// when the sweeps' abort flag is set.  Polling them instead of waiting on
// an event keeps the event's marker packet, ~5.6 us of idle GPU per
// iteration, out of the stream.  After 10 ms (an iteration takes well
// under one) the stream is queried as well, which enqueues a marker of its
// own: a stream that failed, or finished without the partials, ends the
// wait with an error.
// Test hook ("debug_abort_sweeps"): the sweeps' abort flag set before the
// loop, as a sweep whose bounded wait ran out would set it, so the solve
// takes its error path (every kernel still runs to its end).
static void force_abort(hipStream_t st, SweepPlan& P) {
  if (P.debug_abort) WC_HIP(hipMemsetAsync(P.progress.p + 2, 1, 1, st));
}

static void clear_residual(PcgWork& w) {
  for (int i = 0; i < w.red_blocks; ++i) w.host_part[i] = -1.0;
}

static double box_residual_max(hipStream_t st, PcgWork& w) {
  volatile const double* hp = w.host_part;
  const int nb = w.red_blocks;
  int i = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (long spin = 1;; ++spin) {
    while (i < nb && hp[i] != -1.0) ++i;
    if (i == nb) break;
    if ((spin & 4095) == 0 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(10)) {
      const hipError_t e = hipStreamQuery(st);
      if (e == hipSuccess) {  // idle: every partial is written by now
        while (i < nb && hp[i] != -1.0) ++i;
        if (i == nb) break;
        throw HipError{hipErrorUnknown, "PCG update: residual partials missing", __FILE__,
                       __LINE__};
      }
      if (e != hipErrorNotReady) WC_HIP(e);
    }
    __builtin_ia32_pause();
  }
  __atomic_thread_fence(__ATOMIC_ACQUIRE);
  if (hp[0] == -2.0)
    throw HipError{hipErrorLaunchTimeOut,
                   "pipelined MIC(0) sweep: a tile wait exceeded its bound", __FILE__,
                   __LINE__};
  double m = 0;
  for (int k = 0; k < nb; ++k) m = fmax(m, hp[k]);
  clear_residual(w);  // for the next update, which the host enqueues after this
  return m;
}

// Harvest the kernel timer's completed launches without waiting (the rest
// are harvested later, at the latest when the statistics are read), so the
// solve ends without draining the stream.
static void settle_timer(hipStream_t, PcgWork& w) {
  if (w.timer.on && !w.timer.pending.empty()) w.timer.harvest_ready();
}

// The PCG loop of PCGSolver::solve (PCGSolver.h:280-304) on box vectors: the
// pipelined sweeps apply MIC(0) in place and the SpMV is the box stencil.
// b is gathered into the box once and x scattered back once.  Each iteration
// is the SpMV with the s . z partials, one update kernel (alpha, x, r and the
// |r|_inf partials), then -- enqueued before the host has read the residual,
// since they touch only s and the sweep outputs, which are dead once the
// residual meets the tolerance -- the MIC(0) apply (whose backward sweep
// also leaves the z . r partials, one per tile) and the update of s with
// beta.
static bool pcg_loop_box(hipStream_t st, PcgWork& w, int64_t ndof, const double* b,
                         double* x_dof, double tolr, int maxit, double* residual_out,
                         int* iters_out) {
  SweepPlan& P = w.sweep;
  const int64_t nc = P.ncell;
  const int nb = w.red_blocks;
  double* r = P.rb.p;
  double* s = P.sb.p;
  double* x = P.xb.p;
  double* z = P.zb.p;  // the preconditioned vector, then A s
  KTimer* tm = w.timer.on ? &w.timer : nullptr;
  force_abort(st, P);
  box_gather(st, P, b, r);
  WC_HIP(hipMemsetAsync(x, 0, sizeof(double) * nc, st));
  sweep_apply(st, P, r, tm ? tm->sample() : nullptr, ndof);  // z and the z . r tile partials
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, st, P.tdot.p, P.ndot,
                     w.scal.p, (int)R_SET_RHO);
  WC_LAUNCH_CHECK();
  const double rho = read_scal(st, w, S_RHO);
  if (rho == 0 || rho != rho) {
    *iters_out = 0;
    return false;
  }
  WC_HIP(hipMemcpyAsync(s, z, sizeof(double) * nc, hipMemcpyDeviceToDevice, st));
  clear_residual(w);
  int rho_slot = S_RHO;
  int it;
  bool ok = false;
  for (it = 0; it < maxit; ++it) {
    KTimer* ti = tm ? tm->sample() : nullptr;
    hipEvent_t e0 = ti ? ti->mark(st) : nullptr;
    // z = A s and the partials of s . z
    // from the second iteration on, the s update has written s's face columns
    const DotParts dp = box_spmv_dot(st, P, s, z, w.partial.p, nb, it > 0);
    if (e0) ti->add(KTimer::SPMV, ndof, e0, ti->mark(st));
    hipLaunchKernelGGL(k_update_xr_fused<double>, dim3(nb), dim3(kRedThreads), 0, st, x, r, s,
                       z, nc, dp.p, dp.n, w.scal.p, rho_slot, w.host_part_dev,
                       P.progress.p + 2, P.zf.p);
    WC_LAUNCH_CHECK();
    if (it + 1 < maxit) {
      sweep_apply(st, P, r, ti, ndof, true);  // the update has set zf to the sentinel
      const int rho_next = rho_slot == S_RHO ? S_RHO2 : S_RHO;
      hipLaunchKernelGGL(k_update_s_fused<double>, dim3(nb), dim3(kRedThreads), 0, st, s, z,
                         nc, P.tdot.p, P.ndot, w.scal.p, rho_slot, rho_next,
                         P.spmv_win ? P.xf15.p : nullptr, P.xf0.p);
      WC_LAUNCH_CHECK();
      rho_slot = rho_next;
    }
    const double res = box_residual_max(st, w);
    *residual_out = res;
    if (res <= tolr) {
      ++it;
      ok = true;
      break;
    }
  }
  *iters_out = it;
  box_scatter(st, P, x, x_dof);
  settle_timer(st, w);
  return ok;
}

// The mixed-precision form of the solve (no reference counterpart): PCG
// with FP32 vectors, FP32 factor records and matrix entries, FP64 inner
// products and alpha/beta, inside FP64 iterative refinement.  The factor is
// computed in FP64 (pivots and the min-diagonal guard included) and rounded.
// Each pass solves A d = r from d = 0, adding d into the FP64 solution as it
// goes, until its FP32 residual has dropped by 10^-digits (or to the target);
// the FP64 residual r = b - A x is then recomputed, and the solve stops when
// it meets |r|_inf <= tol |b|_inf, after 1 + refinements passes, or when the
// iterations of all passes reach maxit.  iterations_out counts every inner
// iteration.
static bool pcg_loop_box_mixed(hipStream_t st, PcgWork& w, int64_t ndof, const double* b,
                               double* x_dof, double tolr, int maxit,
                               double* residual_out, int* iters_out) {
  SweepPlan& P = w.sweep;
  const int64_t nc = P.ncell;
  const int nb = w.red_blocks;
  sweep_narrow(st, P);
  double* x = P.xb.p;
  float* r = P.rb32.p;
  float* s = P.sb32.p;
  float* z = P.zb32.p;  // the preconditioned vector, then A s
  KTimer* tm = w.timer.on ? &w.timer : nullptr;
  force_abort(st, P);
  box_gather(st, P, b, P.bb.p);
  WC_HIP(hipMemsetAsync(x, 0, sizeof(double) * nc, st));
  double eta = 1.0;
  for (int d = 0; d < w.mixed_digits; ++d) eta *= 0.1;
  int total = 0;
  w.passes = 0;
  for (;;) {
    // FP64 residual of the current solution
    box_residual(st, P, P.bb.p, x, r, w.partial.p, nb);
    hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(kRedThreads), 0, st, w.partial.p, nb,
                       w.scal.p);
    WC_LAUNCH_CHECK();
    // the solution so far, in case this pass is the last (enqueued before
    // the read-back, so the device is busy while the host wakes up)
    box_scatter(st, P, x, x_dof);
    const double res = read_scal(st, w, S_RES);
    *residual_out = res;
    *iters_out = total;
    if (res <= tolr) {
      settle_timer(st, w);
      return true;
    }
    if (total >= maxit || w.passes > w.mixed_refinements || res != res) {
      settle_timer(st, w);
      return false;
    }
    ++w.passes;
    const double inner_tol = std::max(tolr, eta * res);
    sweep_apply32(st, P, r, tm ? tm->sample() : nullptr, ndof);  // z and the z . r partials
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, st, P.tdot.p, P.ndot,
                       w.scal.p, (int)R_SET_RHO);
    WC_LAUNCH_CHECK();
    const double rho = read_scal(st, w, S_RHO);
    if (rho == 0 || rho != rho) {
      box_scatter(st, P, x, x_dof);
      settle_timer(st, w);
      return false;
    }
    WC_HIP(hipMemcpyAsync(s, z, sizeof(float) * nc, hipMemcpyDeviceToDevice, st));
    // inner PCG iterations, with the next apply enqueued before the host reads
    // the residual (as in pcg_loop_box: it touches only s and the sweep
    // outputs, which the next refinement pass recomputes)
    clear_residual(w);
    int rho_slot = S_RHO;
    bool faces = false;  // s's face columns, written by the s update
    while (total < maxit) {
      KTimer* ti = tm ? tm->sample() : nullptr;
      hipEvent_t e0 = ti ? ti->mark(st) : nullptr;
      // z = A s and the partials of s . z
      const DotParts dp = box_spmv32_dot(st, P, s, z, w.partial.p, nb, faces);
      faces = true;
      if (e0) ti->add(KTimer::SPMV, ndof, e0, ti->mark(st));
      hipLaunchKernelGGL(k_update_xr_fused<float>, dim3(nb), dim3(kRedThreads), 0, st, x, r,
                         s, z, nc, dp.p, dp.n, w.scal.p, rho_slot, w.host_part_dev,
                         P.progress.p + 2, P.zf32.p);
      WC_LAUNCH_CHECK();
      ++total;
      if (total < maxit) {
        sweep_apply32(st, P, r, ti, ndof, true);  // the update has set zf32 to the sentinel
        const int rho_next = rho_slot == S_RHO ? S_RHO2 : S_RHO;
        hipLaunchKernelGGL(k_update_s_fused<float>, dim3(nb), dim3(kRedThreads), 0, st, s, z,
                           nc, P.tdot.p, P.ndot, w.scal.p, rho_slot, rho_next,
                           P.spmv_win ? P.xf15_32.p : nullptr, P.xf0_32.p);
        WC_LAUNCH_CHECK();
        rho_slot = rho_next;
      }
      const double r32 = box_residual_max(st, w);
      if (r32 <= inner_tol || r32 != r32) break;
    }
  }
}

bool pcg_solve(hipStream_t st, PcgWork& w, const DevCsr& A, const double* b,
               double* x, double tol, int maxit, double tau, double mdr,
               double* residual_out, int* iters_out) {
  const int64_t n = A.n;
  w.sweep.valid = false;  // the plan belongs to the system factored below
  w.factored = false;
  if (tol < 1e-30) tol = 1e-30;  // PCGSolver.h:253
  if (n == 0) {
    *residual_out = 0;
    *iters_out = 0;
    return true;
  }
  const int nb = w.red_blocks;
  double* r = w.r.p;
  double* z = w.z.p;
  double* s = w.s.p;
  WC_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, st));
  WC_HIP(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_absmax_partial, dim3(nb), dim3(kRedThreads), 0, st, r, n,
                     w.partial.p);
  hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(kRedThreads), 0, st,
                     w.partial.p, nb, w.scal.p);
  WC_LAUNCH_CHECK();
  double res = read_scal(st, w, S_RES);
  *residual_out = res;
  if (res < 1e-30) {
    *iters_out = 0;
    return true;
  }
  const double tolr = tol * res;
  if (w.use_sweeps && w.grid.g)
    sweep_prepare(st, w.sweep, *w.grid.g, w.grid.act, w.grid.isdof, w.grid.dofscan,
                  w.grid.rowlen, w.grid.col, A.val, n);
  if (w.sweep.valid) {
    sweep_factor(st, w.sweep, tau, mdr);
    if (sweep_aborted(st, w.sweep))
      throw HipError{hipErrorLaunchTimeOut,
                     "pipelined MIC(0) factor: a tile wait exceeded its bound", __FILE__,
                     __LINE__};
    w.factored = true;
    if (w.mixed)
      return pcg_loop_box_mixed(st, w, n, b, x, tolr, maxit, residual_out, iters_out);
    return pcg_loop_box(st, w, n, b, x, tolr, maxit, residual_out, iters_out);
  }
  pcg_factor(st, w, A, tau, mdr);
  w.factored = true;
  pcg_apply(st, w, A, r, z);
  hipLaunchKernelGGL(k_dot_partial<double>, dim3(nb), dim3(kRedThreads), 0, st, z, r, n,
                     w.partial.p);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, st, w.partial.p,
                     nb, w.scal.p, (int)R_SET_RHO);
  WC_LAUNCH_CHECK();
  const double rho = read_scal(st, w, S_RHO);
  if (rho == 0 || rho != rho) {
    *iters_out = 0;
    return false;
  }
  WC_HIP(hipMemcpyAsync(s, z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  const unsigned gb = blocks_for(n, 256);
  int it;
  for (it = 0; it < maxit; ++it) {
    hipEvent_t e0 = w.timer.on ? w.timer.mark(st) : nullptr;
    hipLaunchKernelGGL(k_spmv, dim3(gb), dim3(256), 0, st, A, s, z);
    if (e0) w.timer.add(KTimer::SPMV, n, e0, w.timer.mark(st));
    hipLaunchKernelGGL(k_dot_partial<double>, dim3(nb), dim3(kRedThreads), 0, st, s, z,
                       n, w.partial.p);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, st,
                       w.partial.p, nb, w.scal.p, (int)R_ALPHA);
    hipLaunchKernelGGL(k_update_xr, dim3(nb), dim3(kRedThreads), 0, st, x, r, s,
                       z, n, w.scal.p, w.partial.p);
    hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(kRedThreads), 0, st,
                       w.partial.p, nb, w.scal.p);
    WC_LAUNCH_CHECK();
    res = read_scal(st, w, S_RES);
    *residual_out = res;
    if (res <= tolr) {
      *iters_out = it + 1;
      return true;
    }
    pcg_apply(st, w, A, r, z);
    hipLaunchKernelGGL(k_dot_partial<double>, dim3(nb), dim3(kRedThreads), 0, st, z, r,
                       n, w.partial.p);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, st,
                       w.partial.p, nb, w.scal.p, (int)R_BETA);
    hipLaunchKernelGGL(k_update_s<double>, dim3(gb), dim3(256), 0, st, s, z, n,
                       w.scal.p);
    WC_LAUNCH_CHECK();
  }
  *iters_out = it;
  return false;
}

// AMGPCGSolveSparse (AlgebraicMultigrid.h:247-359): x = 0, r = b; the
// tolerance is tol |b|_inf (no lower clamp, unlike PCGSolver); z = one
// V-cycle of r; false with 0 iterations if rho = z . r is 0 or NaN; then the
// PCG loop with the V-cycle as the preconditioner.
bool amgpcg_solve(hipStream_t st, PcgWork& w, AmgHier& H, const DevCsr& A, const double* b,
                  double* x, double tol, int maxit, double* residual_out, int* iters_out) {
  const int64_t n = A.n;
  if (n == 0) {
    *residual_out = 0;
    *iters_out = 0;
    return true;
  }
  const int nb = w.red_blocks;
  double* r = w.r.p;
  double* z = w.z.p;
  double* s = w.s.p;
  WC_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, st));
  WC_HIP(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_absmax_partial, dim3(nb), dim3(kRedThreads), 0, st, r, n, w.partial.p);
  hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(kRedThreads), 0, st, w.partial.p, nb,
                     w.scal.p);
  WC_LAUNCH_CHECK();
  double res = read_scal(st, w, S_RES);
  *residual_out = res;
  if (res == 0) {
    *iters_out = 0;
    return true;
  }
  const double tolr = tol * res;
  amg_vcycle(st, H, r, z);
  hipLaunchKernelGGL(k_dot_partial<double>, dim3(nb), dim3(kRedThreads), 0, st, z, r, n,
                     w.partial.p);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, st, w.partial.p, nb, w.scal.p,
                     (int)R_SET_RHO);
  WC_LAUNCH_CHECK();
  const double rho = read_scal(st, w, S_RHO);
  if (rho == 0 || rho != rho) {
    *iters_out = 0;
    return false;
  }
  WC_HIP(hipMemcpyAsync(s, z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  const unsigned gb = blocks_for(n, 256);
  int it;
  for (it = 0; it < maxit; ++it) {
    if (!amg_spmv(st, H, s, z))
      hipLaunchKernelGGL(k_spmv, dim3(gb), dim3(256), 0, st, A, s, z);
    hipLaunchKernelGGL(k_dot_partial<double>, dim3(nb), dim3(kRedThreads), 0, st, s, z, n,
                       w.partial.p);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, st, w.partial.p, nb,
                       w.scal.p, (int)R_ALPHA);
    hipLaunchKernelGGL(k_update_xr, dim3(nb), dim3(kRedThreads), 0, st, x, r, s, z, n,
                       w.scal.p, w.partial.p);
    hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(kRedThreads), 0, st, w.partial.p, nb,
                       w.scal.p);
    WC_LAUNCH_CHECK();
    res = read_scal(st, w, S_RES);
    *residual_out = res;
    if (res <= tolr) {
      *iters_out = it + 1;
      return true;
    }
    amg_vcycle(st, H, r, z);
    hipLaunchKernelGGL(k_dot_partial<double>, dim3(nb), dim3(kRedThreads), 0, st, z, r, n,
                       w.partial.p);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, st, w.partial.p, nb,
                       w.scal.p, (int)R_BETA);
    hipLaunchKernelGGL(k_update_s<double>, dim3(gb), dim3(256), 0, st, s, z, n, w.scal.p);
    WC_LAUNCH_CHECK();
  }
  *iters_out = it;
  return false;
}

}  // namespace wc
