/*
 * sigma.c -- sigma build H.u and S^2.u of the oracle (test infrastructure
 * only).
 *
 * Restates src/Davidson/u0Hu0.irp.f:
 *   H_S2_u_0_nstates_openmp           :15-66  (reorder to bilinear order,
 *                                              transpose to u_t(N_st,N_det))
 *   H_S2_u_0_nstates_openmp_work_$N_int :97-458
 *     loop A (alpha/beta doubles)     :177-246
 *     loop B (same-spin + diagonal)   :248-444
 * with the spin-string filters get_all_spin_singles(_and_doubles)_$N_int of
 * src/Determinants/spindeterminants.irp.f:920-1191, and OpenMP
 * schedule(dynamic,64) over rows as the reference.
 */
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "qp_oracle.h"

/* popcount(xor) over one spin string; the reference stops adding once the
 * count exceeds 4 (spindeterminants.irp.f:1052-1060); only 2 and 4 matter */
static inline int spin_xor_degree(const bk_t *a, const bk_t *b, int nint) {
  int d = 0;
  for (int k = 0; k < nint; ++k) {
    if (d > 4) break;
    d += __builtin_popcountll(a[k] ^ b[k]);
  }
  return d;
}

void oq_H_S2_u_0_work(const oq_ham *H, const oq_bilinear *B, double *v_t, double *s_t,
                      const double *u_t, int N_st, int64_t istart, int64_t iend, int64_t ishift,
                      int64_t istep, int nthreads) {
  const int nint = B->nint;
  const int64_t N_det = B->n_det;
  const int maxab = (B->n_alpha > B->n_beta ? B->n_alpha : B->n_beta) + 1;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
  {
    bk_t *buffer = (bk_t *)malloc(sizeof(bk_t) * nint * maxab);
    int *singles_a = (int *)malloc(sizeof(int) * maxab);
    int *singles_b = (int *)malloc(sizeof(int) * maxab);
    int *doubles = (int *)malloc(sizeof(int) * maxab);
    int *idx = (int *)malloc(sizeof(int) * maxab);
    bk_t tmp_det[2 * OQ_NINT_MAX], tmp_det2[2 * OQ_NINT_MAX];
    int n_singles_b = 0;
    int kcol_prev = -1;

    /* loop A: alpha/beta doubles */
#pragma omp for schedule(dynamic, 64)
    for (int64_t k_a = istart + ishift; k_a <= iend; k_a += istep) {
      int krow = B->rows[k_a - 1], kcol = B->columns[k_a - 1];
      memcpy(tmp_det, B->alpha_unique + (int64_t)(krow - 1) * nint, sizeof(bk_t) * nint);
      memcpy(tmp_det + nint, B->beta_unique + (int64_t)(kcol - 1) * nint, sizeof(bk_t) * nint);
      if (kcol != kcol_prev) {
        n_singles_b = 0;
        for (int j = 1; j <= B->n_beta; ++j)
          if (spin_xor_degree(B->beta_unique + (int64_t)(j - 1) * nint, tmp_det + nint, nint) == 2)
            singles_b[n_singles_b++] = j;
      }
      kcol_prev = kcol;
      for (int i = 0; i < n_singles_b; ++i) {
        int lcol = singles_b[i];
        memcpy(tmp_det2 + nint, B->beta_unique + (int64_t)(lcol - 1) * nint, sizeof(bk_t) * nint);
        int l_a = B->columns_loc[lcol - 1];
        int j = 0;
        for (; l_a < B->columns_loc[lcol]; ++l_a, ++j) {
          int lrow = B->rows[l_a - 1];
          memcpy(buffer + (int64_t)j * nint, B->alpha_unique + (int64_t)(lrow - 1) * nint, sizeof(bk_t) * nint);
          idx[j] = l_a;
        }
        int n_singles_a = 0;
        for (int t = 0; t < j; ++t)
          if (spin_xor_degree(tmp_det, buffer + (int64_t)t * nint, nint) == 2) singles_a[n_singles_a++] = idx[t];
        for (int k = 0; k < n_singles_a; ++k) {
          int la = singles_a[k];
          int lrow = B->rows[la - 1];
          memcpy(tmp_det2, B->alpha_unique + (int64_t)(lrow - 1) * nint, sizeof(bk_t) * nint);
          double hij = oq_i_H_j_double_alpha_beta(H, tmp_det, tmp_det2);
          double sij = oq_get_s2(tmp_det, tmp_det2, nint);
          for (int l = 0; l < N_st; ++l) {
            v_t[l + (int64_t)N_st * (k_a - 1)] += hij * u_t[l + (int64_t)N_st * (la - 1)];
            s_t[l + (int64_t)N_st * (k_a - 1)] += sij * u_t[l + (int64_t)N_st * (la - 1)];
          }
        }
      }
    }

    /* loop B: same-spin singles/doubles and the diagonal */
#pragma omp for schedule(dynamic, 64)
    for (int64_t k_a = istart + ishift; k_a <= iend; k_a += istep) {
      int krow = B->rows[k_a - 1], kcol = B->columns[k_a - 1];
      memcpy(tmp_det, B->alpha_unique + (int64_t)(krow - 1) * nint, sizeof(bk_t) * nint);
      memcpy(tmp_det + nint, B->beta_unique + (int64_t)(kcol - 1) * nint, sizeof(bk_t) * nint);
      /* alpha excitations inside the beta column */
      int n = 0;
      for (int l_a = B->columns_loc[kcol - 1]; l_a < B->columns_loc[kcol]; ++l_a, ++n) {
        int lrow = B->rows[l_a - 1];
        memcpy(buffer + (int64_t)n * nint, B->alpha_unique + (int64_t)(lrow - 1) * nint, sizeof(bk_t) * nint);
        idx[n] = l_a;
      }
      int ns = 0, nd = 0;
      for (int t = 0; t < n; ++t) {
        int d = spin_xor_degree(tmp_det, buffer + (int64_t)t * nint, nint);
        if (d == 4) doubles[nd++] = idx[t];
        else if (d == 2) singles_a[ns++] = idx[t];
      }
      memcpy(tmp_det2 + nint, B->beta_unique + (int64_t)(kcol - 1) * nint, sizeof(bk_t) * nint);
      for (int i = 0; i < ns; ++i) {
        int la = singles_a[i];
        int lrow = B->rows[la - 1];
        memcpy(tmp_det2, B->alpha_unique + (int64_t)(lrow - 1) * nint, sizeof(bk_t) * nint);
        double hij = oq_i_H_j_mono_spin(H, tmp_det, tmp_det2, 1);
        for (int l = 0; l < N_st; ++l)
          v_t[l + (int64_t)N_st * (k_a - 1)] += hij * u_t[l + (int64_t)N_st * (la - 1)];
      }
      for (int i = 0; i < nd; ++i) {
        int la = doubles[i];
        int lrow = B->rows[la - 1];
        double hij = oq_i_H_j_double_spin(H, tmp_det, B->alpha_unique + (int64_t)(lrow - 1) * nint);
        for (int l = 0; l < N_st; ++l)
          v_t[l + (int64_t)N_st * (k_a - 1)] += hij * u_t[l + (int64_t)N_st * (la - 1)];
      }
      /* beta excitations inside the alpha row (transposed CSR) */
      int k_b = B->order_transp_reverse[k_a - 1];
      int lrow0 = B->transp_rows[k_b - 1];
      n = 0;
      for (int l_b = B->transp_rows_loc[lrow0 - 1]; l_b < B->transp_rows_loc[lrow0]; ++l_b, ++n) {
        int lcol = B->transp_columns[l_b - 1];
        memcpy(buffer + (int64_t)n * nint, B->beta_unique + (int64_t)(lcol - 1) * nint, sizeof(bk_t) * nint);
        idx[n] = l_b;
      }
      ns = 0; nd = 0;
      for (int t = 0; t < n; ++t) {
        int d = spin_xor_degree(tmp_det + nint, buffer + (int64_t)t * nint, nint);
        if (d == 4) doubles[nd++] = idx[t];
        else if (d == 2) singles_b[ns++] = idx[t];
      }
      memcpy(tmp_det2, B->alpha_unique + (int64_t)(krow - 1) * nint, sizeof(bk_t) * nint);
      for (int i = 0; i < ns; ++i) {
        int lb = singles_b[i];
        int lcol = B->transp_columns[lb - 1];
        memcpy(tmp_det2 + nint, B->beta_unique + (int64_t)(lcol - 1) * nint, sizeof(bk_t) * nint);
        double hij = oq_i_H_j_mono_spin(H, tmp_det, tmp_det2, 2);
        int la = B->transp_order[lb - 1];
        for (int l = 0; l < N_st; ++l)
          v_t[l + (int64_t)N_st * (k_a - 1)] += hij * u_t[l + (int64_t)N_st * (la - 1)];
      }
      for (int i = 0; i < nd; ++i) {
        int lb = doubles[i];
        int lcol = B->transp_columns[lb - 1];
        double hij = oq_i_H_j_double_spin(H, tmp_det + nint, B->beta_unique + (int64_t)(lcol - 1) * nint);
        int la = B->transp_order[lb - 1];
        for (int l = 0; l < N_st; ++l)
          v_t[l + (int64_t)N_st * (k_a - 1)] += hij * u_t[l + (int64_t)N_st * (la - 1)];
      }
      /* diagonal */
      double hij = oq_diag_H_mat_elem(H, tmp_det);
      double sij = oq_diag_S_mat_elem(tmp_det, nint);
      for (int l = 0; l < N_st; ++l) {
        v_t[l + (int64_t)N_st * (k_a - 1)] += hij * u_t[l + (int64_t)N_st * (k_a - 1)];
        s_t[l + (int64_t)N_st * (k_a - 1)] += sij * u_t[l + (int64_t)N_st * (k_a - 1)];
      }
    }
    free(buffer); free(singles_a); free(singles_b); free(doubles); free(idx);
  }
  (void)N_det;
}

void oq_H_S2_u_0_nstates(const oq_ham *H, const oq_bilinear *B, double *v_0, double *s_0,
                         const double *u_0, int N_st, int64_t sze, int nthreads) {
  int64_t N = B->n_det;
  double *u_t = (double *)malloc(sizeof(double) * N_st * N);
  double *v_t = (double *)calloc((size_t)N_st * N, sizeof(double));
  double *s_t = (double *)calloc((size_t)N_st * N, sizeof(double));
  /* dset_order(u_0, psi_bilinear_matrix_order) + dtranspose */
  for (int64_t k = 0; k < N; ++k)
    for (int l = 0; l < N_st; ++l) u_t[l + (int64_t)N_st * k] = u_0[(B->order[k] - 1) + sze * l];
  oq_H_S2_u_0_work(H, B, v_t, s_t, u_t, N_st, 1, N, 0, 1, nthreads);
  /* dtranspose + dset_order(., psi_bilinear_matrix_order_reverse) */
  for (int64_t k = 0; k < N; ++k)
    for (int l = 0; l < N_st; ++l) {
      v_0[(B->order[k] - 1) + sze * l] = v_t[l + (int64_t)N_st * k];
      s_0[(B->order[k] - 1) + sze * l] = s_t[l + (int64_t)N_st * k];
    }
  free(u_t); free(v_t); free(s_t);
}

int64_t oq_sigma_count_pairs(const oq_ham *H, const oq_bilinear *B, int nthreads) {
  /* count the off-diagonal (row, column) pairs the traversal visits */
  const int nint = B->nint;
  int64_t total = 0;
  (void)H;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) num_threads(nthreads)
  for (int64_t k_a = 1; k_a <= B->n_det; ++k_a) {
    int krow = B->rows[k_a - 1], kcol = B->columns[k_a - 1];
    const bk_t *a = B->alpha_unique + (int64_t)(krow - 1) * nint;
    const bk_t *b = B->beta_unique + (int64_t)(kcol - 1) * nint;
    int64_t cnt = 0;
    for (int j = 1; j <= B->n_beta; ++j) {
      if (spin_xor_degree(B->beta_unique + (int64_t)(j - 1) * nint, b, nint) != 2) continue;
      for (int l = B->columns_loc[j - 1]; l < B->columns_loc[j]; ++l)
        if (spin_xor_degree(a, B->alpha_unique + (int64_t)(B->rows[l - 1] - 1) * nint, nint) == 2) ++cnt;
    }
    for (int l = B->columns_loc[kcol - 1]; l < B->columns_loc[kcol]; ++l) {
      int d = spin_xor_degree(a, B->alpha_unique + (int64_t)(B->rows[l - 1] - 1) * nint, nint);
      if (d == 2 || d == 4) ++cnt;
    }
    int k_b = B->order_transp_reverse[k_a - 1];
    int r = B->transp_rows[k_b - 1];
    for (int l = B->transp_rows_loc[r - 1]; l < B->transp_rows_loc[r]; ++l) {
      int d = spin_xor_degree(b, B->beta_unique + (int64_t)(B->transp_columns[l - 1] - 1) * nint, nint);
      if (d == 2 || d == 4) ++cnt;
    }
    total += cnt;
  }
  return total;
}
