/*
 * FP64 CPU restatement of the reference's AMG-preconditioned PCG
 * (TEST INFRASTRUCTURE, part of liboracle.so; never linked into the product).
 *
 * Follows, operation for operation in the order the reference's loops run:
 *   - level generation: GeometricLevelGen.h:125-171 (generateRPSparse:
 *     coarse cell (i/2, j/2, k/2), coarse DOFs numbered in order of first
 *     appearance over the fine DOFs, P = 1, R = 0.125 P^T),
 *     :173-184 (generatePatternSparse: red where i+j+k is odd) and
 *     :186-249 (generateLevelsGalerkinCoarseningSparse: coarsen while more
 *     than 4096 unknowns, dims ceil(n/2), A_c = 0.5 R (A P));
 *   - the products: SparseMatrix.h:327-348 (multiplyMat: c(i, B.col) +=
 *     scale * A_ik * B_kc over A's row i, then B's row k, in order);
 *   - smoothing: AlgebraicMultigrid.h:117-177 (RBGS_with_pattern: red rows,
 *     then black rows, each x_i = (b_i - sum_{j != i} a_ij x_j) / a_ii in row
 *     order, 0 where a_ii == 0);
 *   - restriction / prolongation: :96-116 (b_c = R (b - A x) with the
 *     residual subtracted entry by entry in row order; x_f += P x_c);
 *   - the V-cycle: :179-232 (4 sweeps down and up, 200 on the coarsest
 *     level, x = 0 at every level);
 *   - the solver: :247-359 (AMGPCGSolveSparse).
 * The BLAS dot is a sequential sum (the reference's Eigen dot blocks its sum),
 * so iterates agree with the reference to rounding, not bit for bit.
 *
 * Parity status: the coarse operator is pinned against a dense 0.5 R A P
 * (numpy), the smoother against the exact solution of small SPD systems, the
 * solver against scipy's direct solve (tests/test_oracle_amg.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "wc_oracle.h"

typedef struct {
  int64_t n;
  uint32_t* rs;  /* n + 1 */
  uint32_t* col;
  double* val;
  int32_t* ijk;  /* 3 per DOF */
  int ni, nj, nk;
  char* red;     /* pattern: 1 where i+j+k is odd */
  uint32_t* parent;  /* coarse DOF of each DOF (levels above the coarsest) */
} amg_level;

static void* amg_alloc(size_t n, size_t sz) {
  void* p = calloc(n ? n : 1, sz);
  if (!p) abort();
  return p;
}

/* one sparse row being accumulated: sorted columns (add_to_element) */
typedef struct {
  uint32_t* c;
  double* v;
  int len, cap;
} accrow;

static void acc_add(accrow* r, uint32_t c, double v) {
  int lo = 0, hi = r->len;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (r->c[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < r->len && r->c[lo] == c) {
    r->v[lo] += v;
    return;
  }
  if (r->len == r->cap) {
    r->cap = r->cap ? 2 * r->cap : 8;
    r->c = (uint32_t*)realloc(r->c, sizeof(uint32_t) * r->cap);
    r->v = (double*)realloc(r->v, sizeof(double) * r->cap);
    if (!r->c || !r->v) abort();
  }
  memmove(r->c + lo + 1, r->c + lo, sizeof(uint32_t) * (r->len - lo));
  memmove(r->v + lo + 1, r->v + lo, sizeof(double) * (r->len - lo));
  r->c[lo] = c;
  r->v[lo] = v;
  r->len++;
}

static void pattern(amg_level* L) {
  L->red = (char*)amg_alloc((size_t)L->n, 1);
  for (int64_t i = 0; i < L->n; ++i)
    L->red[i] = ((L->ijk[3 * i] + L->ijk[3 * i + 1] + L->ijk[3 * i + 2]) % 2) == 1;
}

/* coarsen L into C (GeometricLevelGen.h:125-171, 228-233) */
static void coarsen(amg_level* L, amg_level* C) {
  const int nni = (int)ceilf((float)L->ni / 2.0f);
  const int nnj = (int)ceilf((float)L->nj / 2.0f);
  const int nnk = (int)ceilf((float)L->nk / 2.0f);
  const int64_t ncell = (int64_t)nni * nnj * nnk;
  int64_t* map = (int64_t*)amg_alloc((size_t)ncell, sizeof(int64_t));
  for (int64_t c = 0; c < ncell; ++c) map[c] = -1;
  L->parent = (uint32_t*)amg_alloc((size_t)L->n, sizeof(uint32_t));
  int32_t* cijk = (int32_t*)amg_alloc((size_t)L->n * 3, sizeof(int32_t));
  int64_t nc = 0;
  for (int64_t f = 0; f < L->n; ++f) {
    const int ci = L->ijk[3 * f] / 2, cj = L->ijk[3 * f + 1] / 2, ck = L->ijk[3 * f + 2] / 2;
    const int64_t key = ci + (int64_t)cj * nni + (int64_t)ck * nni * nnj;
    if (map[key] < 0) {
      map[key] = nc;
      cijk[3 * nc] = ci;
      cijk[3 * nc + 1] = cj;
      cijk[3 * nc + 2] = ck;
      ++nc;
    }
    L->parent[f] = (uint32_t)map[key];
  }
  free(map);
  /* temp = A P (scale 1): temp(i, parent(k)) += 1 * A_ik * 1 over row i */
  accrow* tmp = (accrow*)amg_alloc((size_t)L->n, sizeof(accrow));
  for (int64_t i = 0; i < L->n; ++i)
    for (uint32_t e = L->rs[i]; e < L->rs[i + 1]; ++e)
      acc_add(&tmp[i], L->parent[L->col[e]], 1.0 * L->val[e] * 1.0);
  /* A_c = R temp (scale 0.5): R's row I holds its children in ascending fine
   * order with 0.125; c(I, J) += 0.5 * 0.125 * temp(k, J) */
  accrow* ac = (accrow*)amg_alloc((size_t)nc, sizeof(accrow));
  for (int64_t k = 0; k < L->n; ++k) {
    const uint32_t I = L->parent[k];
    for (int e = 0; e < tmp[k].len; ++e) acc_add(&ac[I], tmp[k].c[e], 0.5 * 0.125 * tmp[k].v[e]);
    free(tmp[k].c);
    free(tmp[k].v);
  }
  free(tmp);
  C->n = nc;
  C->ni = nni;
  C->nj = nnj;
  C->nk = nnk;
  C->ijk = (int32_t*)amg_alloc((size_t)nc * 3, sizeof(int32_t));
  memcpy(C->ijk, cijk, sizeof(int32_t) * 3 * (size_t)nc);
  free(cijk);
  C->rs = (uint32_t*)amg_alloc((size_t)nc + 1, sizeof(uint32_t));
  int64_t nnz = 0;
  for (int64_t I = 0; I < nc; ++I) nnz += ac[I].len;
  C->col = (uint32_t*)amg_alloc((size_t)nnz, sizeof(uint32_t));
  C->val = (double*)amg_alloc((size_t)nnz, sizeof(double));
  int64_t o = 0;
  for (int64_t I = 0; I < nc; ++I) {
    C->rs[I] = (uint32_t)o;
    for (int e = 0; e < ac[I].len; ++e) {
      C->col[o] = ac[I].c[e];
      C->val[o] = ac[I].v[e];
      ++o;
    }
    free(ac[I].c);
    free(ac[I].v);
  }
  C->rs[nc] = (uint32_t)o;
  free(ac);
  C->red = NULL;
  C->parent = NULL;
}

typedef struct {
  amg_level L[32];
  int nl;
} amg_hier;

static void hier_build(amg_hier* H, int64_t n, const uint32_t* rs, const uint32_t* col,
                       const double* val, const int32_t* ijk, int ni, int nj, int nk) {
  memset(H, 0, sizeof(*H));
  amg_level* L0 = &H->L[0];
  L0->n = n;
  L0->rs = (uint32_t*)amg_alloc((size_t)n + 1, sizeof(uint32_t));
  memcpy(L0->rs, rs, sizeof(uint32_t) * ((size_t)n + 1));
  L0->col = (uint32_t*)amg_alloc(rs[n], sizeof(uint32_t));
  memcpy(L0->col, col, sizeof(uint32_t) * rs[n]);
  L0->val = (double*)amg_alloc(rs[n], sizeof(double));
  memcpy(L0->val, val, sizeof(double) * rs[n]);
  L0->ijk = (int32_t*)amg_alloc((size_t)n * 3, sizeof(int32_t));
  memcpy(L0->ijk, ijk, sizeof(int32_t) * 3 * (size_t)n);
  L0->ni = ni;
  L0->nj = nj;
  L0->nk = nk;
  H->nl = 1;
  while (H->L[H->nl - 1].n > 4096 && H->nl < 32) {
    amg_level* F = &H->L[H->nl - 1];
    pattern(F);
    coarsen(F, &H->L[H->nl]);
    H->nl++;
  }
  pattern(&H->L[H->nl - 1]);
}

static void hier_free(amg_hier* H) {
  for (int l = 0; l < H->nl; ++l) {
    free(H->L[l].rs);
    free(H->L[l].col);
    free(H->L[l].val);
    free(H->L[l].ijk);
    free(H->L[l].red);
    free(H->L[l].parent);
  }
}

/* AlgebraicMultigrid.h:117-177 */
static void rbgs(const amg_level* L, const double* b, double* x, int iters) {
  for (int it = 0; it < iters; ++it)
    for (int color = 1; color >= 0; --color)
      for (int64_t i = 0; i < L->n; ++i) {
        if (L->red[i] != color) continue;
        double sum = 0, diag = 0;
        for (uint32_t e = L->rs[i]; e < L->rs[i + 1]; ++e) {
          if (L->col[e] != (uint32_t)i && L->col[e] < (uint32_t)L->n)
            sum += L->val[e] * x[L->col[e]];
          else
            diag = L->val[e];
        }
        x[i] = diag != 0 ? (b[i] - sum) / diag : 0;
      }
}

/* AlgebraicMultigrid.h:179-245: z = one V-cycle from z = 0 */
static void vcycle(const amg_hier* H, const double* b, double* z) {
  double* xs[32];
  double* bs[32];
  for (int l = 0; l < H->nl; ++l) {
    xs[l] = (double*)amg_alloc((size_t)H->L[l].n, sizeof(double));
    bs[l] = (double*)amg_alloc((size_t)H->L[l].n, sizeof(double));
  }
  memcpy(bs[0], b, sizeof(double) * (size_t)H->L[0].n);
  for (int l = 0; l < H->nl - 1; ++l) {
    const amg_level* L = &H->L[l];
    rbgs(L, bs[l], xs[l], 4);
    /* restriction: r = b - A x (row order), b_next = R r */
    double* r = (double*)amg_alloc((size_t)L->n, sizeof(double));
    memcpy(r, bs[l], sizeof(double) * (size_t)L->n);
    for (int64_t i = 0; i < L->n; ++i)
      for (uint32_t e = L->rs[i]; e < L->rs[i + 1]; ++e) r[i] -= L->val[e] * xs[l][L->col[e]];
    /* multiply(R, r): row I = children in ascending fine order */
    for (int64_t I = 0; I < H->L[l + 1].n; ++I) bs[l + 1][I] = 0;
    for (int64_t f = 0; f < L->n; ++f) bs[l + 1][L->parent[f]] += 0.125 * r[f];
    free(r);
  }
  rbgs(&H->L[H->nl - 1], bs[H->nl - 1], xs[H->nl - 1], 200);
  for (int l = H->nl - 2; l >= 0; --l) {
    const amg_level* L = &H->L[l];
    for (int64_t f = 0; f < L->n; ++f) xs[l][f] += 1.0 * (0.0 + 1.0 * xs[l + 1][L->parent[f]]);
    rbgs(L, bs[l], xs[l], 4);
  }
  memcpy(z, xs[0], sizeof(double) * (size_t)H->L[0].n);
  for (int l = 0; l < H->nl; ++l) {
    free(xs[l]);
    free(bs[l]);
  }
}

int wco_amg_levels(int64_t n, const uint32_t* rs, const uint32_t* col, const double* val,
                   const int32_t* ijk, int ni, int nj, int nk, int64_t* sizes, int max_levels) {
  amg_hier H;
  hier_build(&H, n, rs, col, val, ijk, ni, nj, nk);
  for (int l = 0; l < H.nl && l < max_levels; ++l) sizes[l] = H.L[l].n;
  const int nl = H.nl;
  hier_free(&H);
  return nl;
}

int64_t wco_amg_coarse(int64_t n, const uint32_t* rs, const uint32_t* col, const double* val,
                       const int32_t* ijk, int ni, int nj, int nk, uint32_t* parent,
                       uint32_t* crs, uint32_t* ccol, double* cval, int64_t cap) {
  amg_level F;
  memset(&F, 0, sizeof(F));
  F.n = n;
  F.rs = (uint32_t*)rs;
  F.col = (uint32_t*)col;
  F.val = (double*)val;
  F.ijk = (int32_t*)ijk;
  F.ni = ni;
  F.nj = nj;
  F.nk = nk;
  amg_level C;
  memset(&C, 0, sizeof(C));
  coarsen(&F, &C);
  const int64_t nc = C.n;
  if (parent) memcpy(parent, F.parent, sizeof(uint32_t) * (size_t)n);
  if (C.rs[nc] <= (uint64_t)cap) {
    memcpy(crs, C.rs, sizeof(uint32_t) * ((size_t)nc + 1));
    memcpy(ccol, C.col, sizeof(uint32_t) * C.rs[nc]);
    memcpy(cval, C.val, sizeof(double) * C.rs[nc]);
  }
  free(F.parent);
  free(C.rs);
  free(C.col);
  free(C.val);
  free(C.ijk);
  return nc;
}

void wco_amg_rbgs(int64_t n, const uint32_t* rs, const uint32_t* col, const double* val,
                  const int32_t* ijk, const double* b, double* x, int iters) {
  amg_level L;
  memset(&L, 0, sizeof(L));
  L.n = n;
  L.rs = (uint32_t*)rs;
  L.col = (uint32_t*)col;
  L.val = (double*)val;
  L.ijk = (int32_t*)ijk;
  pattern(&L);
  rbgs(&L, b, x, iters);
  free(L.red);
}

void wco_amg_vcycle(int64_t n, const uint32_t* rs, const uint32_t* col, const double* val,
                    const int32_t* ijk, int ni, int nj, int nk, const double* b, double* z) {
  amg_hier H;
  hier_build(&H, n, rs, col, val, ijk, ni, nj, nk);
  vcycle(&H, b, z);
  hier_free(&H);
}

/* AlgebraicMultigrid.h:247-359 */
int wco_amgpcg_solve(int64_t n, const uint32_t* rs, const uint32_t* col, const double* val,
                     const int32_t* ijk, int ni, int nj, int nk, const double* rhs,
                     double* result, double tol_factor, int maxit, double* residual_out,
                     int* iters_out) {
  amg_hier H;
  hier_build(&H, n, rs, col, val, ijk, ni, nj, nk);
  double* r = (double*)amg_alloc((size_t)n, sizeof(double));
  double* z = (double*)amg_alloc((size_t)n, sizeof(double));
  double* s = (double*)amg_alloc((size_t)n, sizeof(double));
  int ok = 0;
  for (int64_t i = 0; i < n; ++i) {
    result[i] = 0;
    r[i] = rhs[i];
  }
  *residual_out = wco_abs_max(n, r);
  if (*residual_out == 0) {
    *iters_out = 0;
    ok = 1;
    goto done;
  }
  {
    const double tol = tol_factor * *residual_out;
    vcycle(&H, r, z);
    double rho = wco_dot(n, z, r);
    if (rho == 0 || rho != rho) {
      *iters_out = 0;
      goto done;
    }
    memcpy(s, z, sizeof(double) * (size_t)n);
    int it;
    for (it = 0; it < maxit; ++it) {
      wco_spmv(n, rs, col, val, s, z);
      const double alpha = rho / wco_dot(n, s, z);
      for (int64_t i = 0; i < n; ++i) result[i] += alpha * s[i];
      for (int64_t i = 0; i < n; ++i) r[i] += -alpha * z[i];
      *residual_out = wco_abs_max(n, r);
      if (*residual_out <= tol) {
        *iters_out = it + 1;
        ok = 1;
        goto done;
      }
      vcycle(&H, r, z);
      const double rho_new = wco_dot(n, z, r);
      const double beta = rho_new / rho;
      for (int64_t i = 0; i < n; ++i) z[i] += beta * s[i];
      double* t = s;
      s = z;
      z = t;
      rho = rho_new;
    }
    *iters_out = it;
  }
done:
  free(r);
  free(z);
  free(s);
  hier_free(&H);
  return ok;
}
