/*
 * bilinear.c -- alpha x beta bilinear structure of the oracle (test
 * infrastructure only).
 *
 * Restates src/Determinants/spindeterminants.irp.f:
 *   spin_det_search_key            :9-24
 *   psi_det_{alpha,beta}_unique    :56-136 (sort by key, remove duplicates)
 *   psi_bilinear_matrix_{rows,columns,order}        :387-446 (key N_a*(col-1)+row)
 *   psi_bilinear_matrix_order_reverse               :448-463
 *   psi_bilinear_matrix_columns_loc                 :465-494
 *   psi_bilinear_matrix_transp_{rows,columns,order} :496-555 (key N_b*(row-1)+col)
 *   psi_bilinear_matrix_transp_rows_loc             :557-578
 *   psi_bilinear_matrix_order_transp_reverse        :580-597
 * Ties among equal search keys (possible for N_int > 1) are broken by the
 * words themselves, so the unique order is deterministic.
 */
#include <stdlib.h>
#include <string.h>
#include "qp_oracle.h"

static int g_nint;

static uint64_t search_key(const bk_t *s, int nint) {
  uint64_t k = s[0];
  for (int i = 1; i < nint; ++i) k ^= s[i];
  return k; /* compared unsigned: same order as the reference's shifted signed key */
}

static int cmp_str(const bk_t *a, const bk_t *b, int nint) {
  uint64_t ka = search_key(a, nint), kb = search_key(b, nint);
  if (ka != kb) return ka < kb ? -1 : 1;
  for (int i = 0; i < nint; ++i)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}

typedef struct { const bk_t *s; int64_t idx; } sref;
static int cmp_sref(const void *x, const void *y) {
  const sref *a = (const sref *)x, *b = (const sref *)y;
  int c = cmp_str(a->s, b->s, g_nint);
  if (c) return c;
  return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

static int build_unique(const bk_t *psi_det, int64_t n_det, int nint, int spin, bk_t **out) {
  sref *r = (sref *)malloc(sizeof(sref) * n_det);
  for (int64_t k = 0; k < n_det; ++k) { r[k].s = psi_det + (2 * k + spin) * nint; r[k].idx = k; }
  g_nint = nint;
  qsort(r, n_det, sizeof(sref), cmp_sref);
  bk_t *u = (bk_t *)malloc(sizeof(bk_t) * nint * (n_det > 0 ? n_det : 1));
  int n = 0;
  for (int64_t k = 0; k < n_det; ++k) {
    if (n > 0 && cmp_str(u + (int64_t)(n - 1) * nint, r[k].s, nint) == 0) continue;
    memcpy(u + (int64_t)n * nint, r[k].s, sizeof(bk_t) * nint);
    ++n;
  }
  free(r);
  *out = u;
  return n;
}

static int find_unique(const bk_t *u, int n, const bk_t *s, int nint) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int c = cmp_str(u + (int64_t)mid * nint, s, nint);
    if (c == 0) return mid + 1;
    if (c < 0) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

typedef struct { int64_t key; int idx; } kpair;
static int cmp_kpair(const void *x, const void *y) {
  const kpair *a = (const kpair *)x, *b = (const kpair *)y;
  if (a->key != b->key) return a->key < b->key ? -1 : 1;
  return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

oq_bilinear *oq_bilinear_create(const bk_t *psi_det, int64_t n_det, int nint) {
  oq_bilinear *B = (oq_bilinear *)calloc(1, sizeof(oq_bilinear));
  B->n_det = n_det;
  B->nint = nint;
  B->n_alpha = build_unique(psi_det, n_det, nint, 0, &B->alpha_unique);
  B->n_beta = build_unique(psi_det, n_det, nint, 1, &B->beta_unique);
  int64_t N = n_det;
  B->rows = (int *)malloc(sizeof(int) * N);
  B->columns = (int *)malloc(sizeof(int) * N);
  B->order = (int *)malloc(sizeof(int) * N);
  B->order_reverse = (int *)malloc(sizeof(int) * N);
  B->columns_loc = (int *)malloc(sizeof(int) * (B->n_beta + 1));
  B->transp_rows = (int *)malloc(sizeof(int) * N);
  B->transp_columns = (int *)malloc(sizeof(int) * N);
  B->transp_order = (int *)malloc(sizeof(int) * N);
  B->transp_rows_loc = (int *)malloc(sizeof(int) * (B->n_alpha + 1));
  B->order_transp_reverse = (int *)malloc(sizeof(int) * N);
  int *ri = (int *)malloc(sizeof(int) * N), *ci = (int *)malloc(sizeof(int) * N);
  kpair *kp = (kpair *)malloc(sizeof(kpair) * N);
  for (int64_t k = 0; k < N; ++k) {
    ri[k] = find_unique(B->alpha_unique, B->n_alpha, psi_det + 2 * k * nint, nint);
    ci[k] = find_unique(B->beta_unique, B->n_beta, psi_det + (2 * k + 1) * nint, nint);
    kp[k].key = (int64_t)B->n_alpha * (ci[k] - 1) + ri[k];
    kp[k].idx = (int)(k + 1);
  }
  qsort(kp, N, sizeof(kpair), cmp_kpair);
  for (int64_t k = 0; k < N; ++k) {
    int o = kp[k].idx;
    B->order[k] = o;
    B->rows[k] = ri[o - 1];
    B->columns[k] = ci[o - 1];
  }
  for (int64_t k = 0; k < N; ++k) B->order_reverse[B->order[k] - 1] = (int)(k + 1);
  for (int j = 0; j <= B->n_beta; ++j) B->columns_loc[j] = (int)N + 1;
  for (int64_t k = N - 1; k >= 0; --k) B->columns_loc[B->columns[k] - 1] = (int)(k + 1);
  B->columns_loc[B->n_beta] = (int)N + 1;
  /* transposed */
  for (int64_t k = 0; k < N; ++k) {
    kp[k].key = (int64_t)B->n_beta * (B->rows[k] - 1) + B->columns[k];
    kp[k].idx = (int)(k + 1);
  }
  qsort(kp, N, sizeof(kpair), cmp_kpair);
  for (int64_t k = 0; k < N; ++k) {
    int o = kp[k].idx;
    B->transp_order[k] = o;
    B->transp_rows[k] = B->rows[o - 1];
    B->transp_columns[k] = B->columns[o - 1];
  }
  for (int i = 0; i <= B->n_alpha; ++i) B->transp_rows_loc[i] = (int)N + 1;
  for (int64_t k = N - 1; k >= 0; --k) B->transp_rows_loc[B->transp_rows[k] - 1] = (int)(k + 1);
  B->transp_rows_loc[B->n_alpha] = (int)N + 1;
  for (int64_t k = 0; k < N; ++k) B->order_transp_reverse[B->transp_order[k] - 1] = (int)(k + 1);
  free(ri); free(ci); free(kp);
  return B;
}

void oq_bilinear_destroy(oq_bilinear *B) {
  if (!B) return;
  free(B->alpha_unique); free(B->beta_unique);
  free(B->rows); free(B->columns); free(B->order); free(B->order_reverse); free(B->columns_loc);
  free(B->transp_rows); free(B->transp_columns); free(B->transp_order); free(B->transp_rows_loc);
  free(B->order_transp_reverse);
  free(B);
}
