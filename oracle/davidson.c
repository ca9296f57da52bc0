/*
 * davidson.c -- Davidson diagonaliser of the oracle (test infrastructure
 * only).
 *
 * Restates src/Davidson/diagonalization_hs2.irp.f:
 *   davidson_diag_hs2              :1-49   (H_jj = diag_H_mat_elem)
 *   davidson_diag_hjj_sjj          :52-444
 * src/Davidson/parameters.irp.f:9-49 (davidson_converged, criterion
 * 'residual': max squared residual norm < threshold, never before iter 2),
 * src/Utils/LinearAlgebra.irp.f:123-153 (ortho_qr: Householder QR with the
 * LAPACK dgeqrf/dorgqr conventions) and :446-512 (lapack_diag: here a cyclic
 * Jacobi eigensolver, eigenvalues ascending; LAPACK itself is not present).
 * The random extra guess vectors (:188-197, Fortran random_number) use a
 * SplitMix64 stream + Box-Muller so that they are reproducible.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>
#include "qp_oracle.h"

/* ---- cyclic Jacobi eigensolver for a small symmetric matrix ---- */
void oq_jacobi_eigh(int n, const double *Ain, double *w, double *V) {
  double *A = (double *)malloc(sizeof(double) * n * n);
  /* the upper triangle defines the matrix (dsyev 'U') */
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) A[i + n * j] = (i <= j) ? Ain[i + n * j] : Ain[j + n * i];
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) V[i + n * j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        if (i != j) off += A[i + n * j] * A[i + n * j];
        else diag += A[i + n * j] * A[i + n * j];
      }
    if (off <= 1e-30 * (diag > 1e-300 ? diag : 1e-300)) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p + n * q];
        if (fabs(apq) < 1e-300) continue;
        double app = A[p + n * p], aqq = A[q + n * q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[k + n * p], akq = A[k + n * q];
          A[k + n * p] = c * akp - s * akq;
          A[k + n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[p + n * k], aqk = A[q + n * k];
          A[p + n * k] = c * apk - s * aqk;
          A[q + n * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k + n * p], vkq = V[k + n * q];
          V[k + n * p] = c * vkp - s * vkq;
          V[k + n * q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = A[i + n * i];
  /* sort ascending (selection sort, stable) */
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j)
      if (w[j] < w[m]) m = j;
    if (m != i) {
      double t = w[i]; w[i] = w[m]; w[m] = t;
      for (int k = 0; k < n; ++k) { t = V[k + n * i]; V[k + n * i] = V[k + n * m]; V[k + n * m] = t; }
    }
  }
  free(A);
}

/* ---- Householder QR, dgeqrf + dorgqr conventions ---- */
static void ortho_qr(double *A, int64_t lda, int64_t m, int n, int nthreads) {
  double *tau = (double *)calloc(n, sizeof(double));
  for (int j = 0; j < n; ++j) {
    double *a = A + lda * j;
    double alpha = a[j], xn2 = 0.0;
#pragma omp parallel for reduction(+ : xn2) num_threads(nthreads) schedule(static)
    for (int64_t i = j + 1; i < m; ++i) xn2 += a[i] * a[i];
    if (xn2 == 0.0) { tau[j] = 0.0; continue; }
    double beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    tau[j] = (beta - alpha) / beta;
    double sc = 1.0 / (alpha - beta);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = j + 1; i < m; ++i) a[i] *= sc;
    a[j] = beta;
    /* apply H_j = I - tau v v^T to columns j+1..n-1 (v_j = 1) */
    for (int c = j + 1; c < n; ++c) {
      double *b = A + lda * c;
      double dot = b[j];
#pragma omp parallel for reduction(+ : dot) num_threads(nthreads) schedule(static)
      for (int64_t i = j + 1; i < m; ++i) dot += a[i] * b[i];
      dot *= tau[j];
      b[j] -= dot;
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (int64_t i = j + 1; i < m; ++i) b[i] -= dot * a[i];
    }
  }
  /* dorgqr: Q = H_0 ... H_{n-1} I(:,0:n-1), accumulated backwards */
  for (int j = n - 1; j >= 0; --j) {
    double *a = A + lda * j;
    for (int c = j + 1; c < n; ++c) {
      double *b = A + lda * c;
      double dot = 0.0; /* b[j] is 0 in the partially formed Q */
#pragma omp parallel for reduction(+ : dot) num_threads(nthreads) schedule(static)
      for (int64_t i = j + 1; i < m; ++i) dot += a[i] * b[i];
      dot *= tau[j];
      b[j] = -dot;
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (int64_t i = j + 1; i < m; ++i) b[i] -= dot * a[i];
    }
    double t = tau[j];
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = j + 1; i < m; ++i) a[i] *= -t;
    a[j] = 1.0 - t;
    for (int64_t i = 0; i < j; ++i) a[i] = 0.0;
  }
  free(tau);
}

static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double u01(uint64_t *s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* C(m x n) = A(k x m)^T B(k x n), with leading dims */
static void gemm_tn(int m, int n, int64_t k, const double *A, int64_t lda, const double *B, int64_t ldb,
                    double *C, int ldc, int nthreads) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      const double *a = A + lda * i, *b = B + ldb * j;
#pragma omp parallel for reduction(+ : s) num_threads(nthreads) schedule(static)
      for (int64_t t = 0; t < k; ++t) s += a[t] * b[t];
      C[i + (int64_t)ldc * j] = s;
    }
}

/* C(k x n) = A(k x m) Y(m x n) */
static void gemm_nn(int64_t k, int n, int m, const double *A, int64_t lda, const double *Y, int ldy,
                    double *C, int64_t ldc, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t t = 0; t < k; ++t)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int i = 0; i < m; ++i) s += A[t + lda * i] * Y[i + (int64_t)ldy * j];
      C[t + ldc * j] = s;
    }
}

int oq_davidson_diag_hs2(const oq_ham *H, const oq_bilinear *B, const bk_t *dets, double *u_in,
                         int64_t dim_in, int64_t sze, int N_st, int N_st_diag,
                         const oq_davidson_params *P, double *energies, double *s2_out) {
  int nthreads = P->nthreads > 0 ? P->nthreads : omp_get_max_threads();
  int nint = H->nint;
  if (N_st_diag * 3 > sze) return -1;
  int itermax = (int)(sze / N_st_diag);
  if (itermax > P->sze_max) itermax = P->sze_max;
  if (itermax < 3) itermax = 3;
  int64_t ncol = (int64_t)N_st_diag * itermax;
  double *H_jj = (double *)malloc(sizeof(double) * sze);
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < sze; ++i) H_jj[i] = oq_diag_H_mat_elem(H, dets + 2 * i * nint);
  double *W = (double *)calloc((size_t)sze * ncol, sizeof(double));
  double *U = (double *)calloc((size_t)sze * ncol, sizeof(double));
  double *S = (double *)calloc((size_t)sze * ncol, sizeof(double));
  int nc = (int)ncol;
  double *h = (double *)calloc((size_t)nc * nc, sizeof(double));
  double *y = (double *)calloc((size_t)nc * nc, sizeof(double));
  double *s_ = (double *)calloc((size_t)nc * nc, sizeof(double));
  double *s_tmp = (double *)calloc((size_t)nc * nc, sizeof(double));
  double *lambda = (double *)calloc(nc, sizeof(double));
  double *s2 = (double *)calloc(nc, sizeof(double));
  double *resid = (double *)calloc(N_st_diag, sizeof(double));
  int *state_ok = (int *)calloc(nc, sizeof(int));
  int n_sigma = 0, total_iter = 0, failed = 0;

  /* guess: the extra vectors are random (reproducible stream) */
  uint64_t rs = 0x5EEDULL;
  for (int k = N_st; k < N_st_diag; ++k)
    for (int64_t i = 0; i < sze; ++i) {
      double r1 = 1.0 - u01(&rs), r2 = u01(&rs);
      u_in[i + dim_in * k] = sqrt(-2.0 * log(r1)) * cos(6.283185307179586 * r2);
    }
  for (int k = 0; k < N_st_diag; ++k) {
    double nrm = 0.0;
    for (int64_t i = 0; i < sze; ++i) nrm += u_in[i + dim_in * k] * u_in[i + dim_in * k];
    nrm = 1.0 / sqrt(nrm);
    for (int64_t i = 0; i < sze; ++i) u_in[i + dim_in * k] *= nrm;
  }
  int converged = 0;
  int shift2 = 0;
  while (!converged && !failed) {
    for (int k = 0; k < N_st_diag; ++k)
      memcpy(U + sze * k, u_in + dim_in * k, sizeof(double) * sze);
    for (int iter = 1; iter <= itermax - 1; ++iter) {
      int shift = N_st_diag * (iter - 1);
      shift2 = N_st_diag * iter;
      ortho_qr(U, sze, sze, shift2, nthreads);
      {
        /* H_S2_u_0_nstates_openmp on the new block */
        double *Ub = (double *)malloc(sizeof(double) * sze * N_st_diag);
        memcpy(Ub, U + sze * shift, sizeof(double) * sze * N_st_diag);
        oq_H_S2_u_0_nstates(H, B, W + sze * shift, S + sze * shift, Ub, N_st_diag, sze, nthreads);
        free(Ub);
        ++n_sigma;
      }
      gemm_tn(shift2, shift2, sze, U, sze, W, sze, h, nc, nthreads);
      gemm_tn(shift2, shift2, sze, U, sze, S, sze, s_, nc, nthreads);
      {
        double *hh = (double *)malloc(sizeof(double) * shift2 * shift2);
        double *yy = (double *)malloc(sizeof(double) * shift2 * shift2);
        for (int j = 0; j < shift2; ++j)
          for (int i = 0; i < shift2; ++i) hh[i + shift2 * j] = h[i + nc * j];
        oq_jacobi_eigh(shift2, hh, lambda, yy);
        for (int j = 0; j < shift2; ++j)
          for (int i = 0; i < shift2; ++i) y[i + nc * j] = yy[i + shift2 * j];
        free(hh); free(yy);
      }
      /* S^2 of the Ritz vectors */
      for (int j = 0; j < shift2; ++j)
        for (int i = 0; i < shift2; ++i) {
          double t = 0.0;
          for (int k = 0; k < shift2; ++k) t += s_[i + nc * k] * y[k + nc * j];
          s_tmp[i + nc * j] = t;
        }
      for (int j = 0; j < shift2; ++j)
        for (int i = 0; i < shift2; ++i) {
          double t = 0.0;
          for (int k = 0; k < shift2; ++k) t += y[k + nc * i] * s_tmp[k + nc * j];
          s_[i + nc * j] = t;
        }
      for (int k = 0; k < shift2; ++k) s2[k] = s_[k + nc * k] + P->s_z2_sz;
      for (int k = 0; k < nc; ++k) state_ok[k] = 1;
      if (P->s2_eig)
        for (int k = 0; k < shift2; ++k) state_ok[k] = fabs(s2[k] - P->expected_s2) < 0.6;
      for (int k = 0; k < shift2; ++k) {
        if (state_ok[k]) continue;
        for (int l = k + 1; l < shift2; ++l) {
          if (!state_ok[l]) continue;
          for (int i = 0; i < shift2; ++i) {
            double t = y[i + nc * k]; y[i + nc * k] = y[i + nc * l]; y[i + nc * l] = t;
          }
          double t = s2[k]; s2[k] = s2[l]; s2[l] = t;
          t = lambda[k]; lambda[k] = lambda[l]; lambda[l] = t;
          state_ok[k] = 1; state_ok[l] = 0;
          break;
        }
      }
      if (P->state_following) {
        /* diagonalization_hs2.irp.f:328-370 */
        double *ov = (double *)malloc(sizeof(double) * nc * nc);
        int *order = (int *)malloc(sizeof(int) * N_st_diag);
        for (int i = 0; i < nc * nc; ++i) ov[i] = -1.0;
        for (int k = 0; k < shift2; ++k)
          for (int i = 0; i < shift2; ++i) ov[k + nc * i] = fabs(y[k + nc * i]);
        for (int k = 0; k < N_st; ++k) {
          double cmax = -1.0;
          for (int i = 0; i < N_st; ++i)
            if (ov[i + nc * k] > cmax) { cmax = ov[i + nc * k]; order[k] = i; }
          for (int i = 0; i < N_st_diag; ++i) ov[order[k] + nc * i] = -1.0;
        }
        for (int i = 0; i < nc * nc; ++i) ov[i] = y[i];
        for (int k = 0; k < N_st; ++k) {
          int l = order[k];
          if (k != l)
            for (int i = 0; i < shift2; ++i) y[i + nc * k] = ov[i + nc * l];
        }
        for (int k = 0; k < N_st; ++k) { ov[k] = lambda[k]; ov[k + nc] = s2[k]; }
        for (int k = 0; k < N_st; ++k) {
          int l = order[k];
          if (k != l) { lambda[k] = ov[l]; s2[k] = ov[l + nc]; }
        }
        free(ov); free(order);
      }
      gemm_nn(sze, N_st_diag, shift2, U, sze, y, nc, U + sze * shift2, sze, nthreads);
      gemm_nn(sze, N_st_diag, shift2, W, sze, y, nc, W + sze * shift2, sze, nthreads);
      gemm_nn(sze, N_st_diag, shift2, S, sze, y, nc, S + sze * shift2, sze, nthreads);
      for (int k = 0; k < N_st_diag; ++k) {
        double *u = U + sze * (shift2 + k), *wv = W + sze * (shift2 + k), *sv = S + sze * (shift2 + k);
        double lk = lambda[k], s2k = s2[k], rn = 0.0;
#pragma omp parallel for reduction(+ : rn) num_threads(nthreads) schedule(static)
        for (int64_t i = 0; i < sze; ++i) {
          double den = H_jj[i] - lk;
          if (den < 1e-2) den = 1e-2;
          u[i] = (lk * u[i] - wv[i]) * (1.0 + s2k * u[i] - sv[i] - P->s_z2_sz) / den;
          rn += u[i] * u[i];
        }
        if (k < N_st) resid[k] = rn;
      }
      ++total_iter;
      if (P->verbose) {
        printf("%4d", iter);
        for (int k = 0; k < N_st; ++k) printf("  %18.12f %10.6f %10.3e", lambda[k], s2[k], resid[k]);
        printf("\n");
        fflush(stdout);
      }
      double rmax = 0.0;
      for (int k = 0; k < N_st; ++k) rmax = resid[k] > rmax ? resid[k] : rmax;
      converged = (iter >= 2) && (fabs(rmax) < P->threshold);
      for (int k = 0; k < N_st; ++k)
        if (resid[k] > 1e8) failed = 1;
      if (P->max_iter > 0 && total_iter >= P->max_iter) { if (!converged) failed = 2; }
      if (converged || failed) break;
    }
    /* re-contract to u_in */
    gemm_nn(sze, N_st_diag, shift2, U, sze, y, nc, u_in, dim_in, nthreads);
  }
  for (int k = 0; k < N_st_diag; ++k) { energies[k] = lambda[k]; s2_out[k] = s2[k]; }
  free(H_jj); free(W); free(U); free(S); free(h); free(y); free(s_); free(s_tmp);
  free(lambda); free(s2); free(resid); free(state_ok);
  if (failed == 1) return -2;
  if (failed == 2) return -3;
  return n_sigma;
}
