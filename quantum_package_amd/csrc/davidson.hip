// davidson.hip -- block Davidson with S^2 tracking on the GPU.
//
// Reference: davidson_diag_hs2 / davidson_diag_hjj_sjj
// (src/Davidson/diagonalization_hs2.irp.f:1-444), davidson_converged
// (src/Davidson/parameters.irp.f:9-49), ortho_qr and lapack_diag
// (src/Utils/LinearAlgebra.irp.f:123-153,446-512).
//
// The tall vectors U, W, S (n_det x N_st_diag*itermax, FP64, column-major,
// bilinear order) stay resident in HBM.  Per iteration:
//   1. the new block is orthonormalised against the previous columns
//      (two passes of block Gram-Schmidt) and within itself (two passes of
//      modified Gram-Schmidt): the same subspace as the reference's
//      Householder QR of U(:,1:shift2), whose leading columns are already
//      orthonormal;
//   2. W, S of the new block from the sigma kernels (FP64, or FP32 for the
//      correction vectors in mixed mode);
//   3. h = U^T W and s = U^T S (tall-skinny reductions, two-stage and
//      deterministic), then the small eigenproblem, S^2 filter and state
//      following on the host (cyclic Jacobi: no LAPACK on this image);
//   4. Ritz vectors U y, W y, S y and the reference's preconditioned
//      residual (lambda U - W)(1 + s2 U - S - S_z2_Sz)/max(H_jj - lambda,
//      1e-2) in one fused pass; convergence on max squared residual.
// Mixed mode: FP32 sigma until the squared residual falls below the switch
// threshold, then a restart on the current Ritz vectors with FP64 sigma,
// so the reported energies come from an all-FP64 subspace.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#include "qpx_internal.h"

using namespace qpx;

namespace {

constexpr int RB = 256;        // threads per reduction block
constexpr int RBLOCKS = 1024;  // reduction blocks

__global__ void permute_in_kernel(const double *__restrict__ src, int64_t ld_src, const int64_t *__restrict__ perm,
                                  int64_t n, int ncol, double *__restrict__ dst, int64_t ld_dst) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n * ncol; x += (int64_t)gridDim.x * blockDim.x) {
    int64_t j = x / n, k = x - j * n;
    dst[k + ld_dst * j] = src[perm[k] + ld_src * j];
  }
}

__global__ void permute_out_kernel(const double *__restrict__ src, int64_t ld_src, const int64_t *__restrict__ perm,
                                   int64_t n, int ncol, double *__restrict__ dst, int64_t ld_dst) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n * ncol; x += (int64_t)gridDim.x * blockDim.x) {
    int64_t j = x / n, k = x - j * n;
    dst[perm[k] + ld_dst * j] = src[k + ld_src * j];
  }
}

// column block (n x nb, ld) -> state-fastest (nb, n) of type T
template <typename T>
__global__ void to_t_kernel(const double *__restrict__ X, int64_t ld, int64_t n, int nb, T *__restrict__ xt) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n * nb; x += (int64_t)gridDim.x * blockDim.x) {
    int64_t k = x / nb;
    int j = (int)(x - k * nb);
    xt[x] = (T)X[k + ld * j];
  }
}

template <typename T>
__global__ void from_t_kernel(const T *__restrict__ xt, int64_t n, int nb, double *__restrict__ X, int64_t ld) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n * nb; x += (int64_t)gridDim.x * blockDim.x) {
    int64_t k = x / nb;
    int j = (int)(x - k * nb);
    X[k + ld * j] = (double)xt[x];
  }
}

// partial[blk][i*nb+j] = sum over the block's rows of A(k,i) B(k,j)
// (na, nb <= 32), rows staged through LDS in tiles of 64
// B2 (optional): nb more columns, so that one pass over A serves A^T [B B2]
__global__ __launch_bounds__(RB) void gram_partial_kernel(const double *__restrict__ A, int64_t lda, int na,
                                                          const double *__restrict__ B,
                                                          const double *__restrict__ B2, int64_t ldb, int nb1,
                                                          int64_t n, double *__restrict__ partial) {
  __shared__ double ta[64 * 33], tb[64 * 33];
  const int nb = B2 ? 2 * nb1 : nb1;
  const int nout = na * nb;
  double acc[4] = {0, 0, 0, 0};
  int64_t rows_per = (n + gridDim.x - 1) / gridDim.x;
  int64_t r0 = blockIdx.x * rows_per, r1 = min(n, r0 + rows_per);
  for (int64_t t0 = r0; t0 < r1; t0 += 64) {
    int m = (int)min<int64_t>(64, r1 - t0);
    __syncthreads();
    for (int x = threadIdx.x; x < 64 * na; x += RB) {
      int i = x / 64, r = x - i * 64;
      ta[r * 33 + i] = (r < m) ? A[t0 + r + lda * i] : 0.0;
    }
    for (int x = threadIdx.x; x < 64 * nb; x += RB) {
      int j = x / 64, r = x - j * 64;
      const double *src = (j < nb1) ? B + ldb * j : B2 + ldb * (j - nb1);
      tb[r * 33 + j] = (r < m) ? src[t0 + r] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int o = threadIdx.x + q * RB;
      if (o < nout) {
        int i = o / nb, j = o - i * nb;
        double s = 0.0;
        for (int r = 0; r < m; ++r) s += ta[r * 33 + i] * tb[r * 33 + j];
        acc[q] += s;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int o = threadIdx.x + q * RB;
    if (o < nout) partial[(int64_t)blockIdx.x * nout + o] = acc[q];
  }
}

// out[o] = sum over the nblk partial rows; one workgroup per output, a
// strided sum per thread and a fixed tree, so the result is deterministic
__global__ __launch_bounds__(RB) void reduce_partial_kernel(const double *__restrict__ partial, int nblk, int nout,
                                                            double *__restrict__ out) {
  __shared__ double red[RB];
  const int o = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += RB) s += partial[(int64_t)b * nout + o];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = RB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[o] = red[0];
}

// X(:, 0:nb) -= A(:, 0:na) C(na x nb)   (C column-major na x nb, in LDS)
__global__ void sub_gemm_kernel(double *__restrict__ X, int64_t ldx, int nb, const double *__restrict__ A,
                                int64_t lda, int na, const double *__restrict__ Cm, int64_t n) {
  __shared__ double cs[32 * 32];
  for (int x = threadIdx.x; x < na * nb; x += blockDim.x) cs[x] = Cm[x];
  __syncthreads();
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    // the row of A stays in registers: the loops run to the compile-time
    // bound (na <= 32 is checked at entry) with the unused slots skipped
    double a[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) a[i] = (i < na) ? A[k + lda * i] : 0.0;
    for (int j = 0; j < nb; ++j) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < 32; ++i)
        if (i < na) s += a[i] * cs[i + na * j];
      X[k + ldx * j] -= s;
    }
  }
}

__global__ void axpy_col_kernel(double *__restrict__ x, const double *__restrict__ y, double alpha, int64_t n) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    x[k] += alpha * y[k];
}

__global__ void scale_col_kernel(double *__restrict__ x, double alpha, int64_t n) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    x[k] *= alpha;
}

// Ritz vectors and preconditioned residual, fused:
//   Un = U y, Wn = W y, Sn = S y (first nsd columns of y, rows 0..shift2-1)
//   U(:, shift2+j) = residual_j, W(:, shift2+j) = Wn_j, S(:, shift2+j) = Sn_j
//   partial[blk][j] = sum residual_j^2
struct RitzArgs {
  double *U, *W, *S;
  int64_t ld, n;
  int shift2, nsd;
  const double *hjj;
  double sz2;
};

__global__ __launch_bounds__(RB) void ritz_residual_kernel(RitzArgs R, const double *__restrict__ y,
                                                           const double *__restrict__ lam,
                                                           const double *__restrict__ s2v,
                                                           double *__restrict__ partial) {
  __shared__ double ys[32 * 8];
  __shared__ double red[RB / 64][8];
  const int ncol = R.shift2;
  for (int x = threadIdx.x; x < ncol * R.nsd; x += blockDim.x) ys[x] = y[x];
  __syncthreads();
  double rs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < R.n; k += (int64_t)gridDim.x * blockDim.x) {
    double un[8] = {0}, wn[8] = {0}, sn[8] = {0};
    for (int i = 0; i < ncol; ++i) {
      double uu = R.U[k + R.ld * i], ww = R.W[k + R.ld * i], ss = R.S[k + R.ld * i];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (j < R.nsd) {
          double c = ys[i + ncol * j];
          un[j] += uu * c;
          wn[j] += ww * c;
          sn[j] += ss * c;
        }
    }
    const double hk = R.hjj[k];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j < R.nsd) {
        double den = hk - lam[j];
        if (den < 1e-2) den = 1e-2;
        double r = (lam[j] * un[j] - wn[j]) * (1.0 + s2v[j] * un[j] - sn[j] - R.sz2) / den;
        R.U[k + R.ld * (R.shift2 + j)] = r;
        R.W[k + R.ld * (R.shift2 + j)] = wn[j];
        R.S[k + R.ld * (R.shift2 + j)] = sn[j];
        rs[j] += r * r;
      }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    double x = rs[j];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) red[wv][j] = x;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    double x = 0.0;
    for (int w = 0; w < RB / 64; ++w) x += red[w][threadIdx.x];
    partial[(int64_t)blockIdx.x * 8 + threadIdx.x] = x;
  }
}

// out(:, j) = A(:, 0:ncol) y(:, j), j < nsd
__global__ void gemm_y_kernel(const double *__restrict__ A, int64_t lda, int ncol, const double *__restrict__ y,
                              int nsd, int64_t n, double *__restrict__ out, int64_t ldo) {
  __shared__ double ys[32 * 8];
  for (int x = threadIdx.x; x < ncol * nsd; x += blockDim.x) ys[x] = y[x];
  __syncthreads();
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    double o[8] = {0};
    for (int i = 0; i < ncol; ++i) {
      double a = A[k + lda * i];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (j < nsd) o[j] += a * ys[i + ncol * j];
    }
    for (int j = 0; j < nsd; ++j) out[k + ldo * j] = o[j];
  }
}

// X(:, 0:nb) <- X(:, 0:nb) R (R upper triangular nb x nb, column-major),
// row by row in place
__global__ void tri_mul_kernel(double *__restrict__ X, int64_t ld, int nb, const double *__restrict__ R, int64_t n) {
  __shared__ double rs[8 * 8];
  for (int x = threadIdx.x; x < nb * nb; x += blockDim.x) rs[x] = R[x];
  __syncthreads();
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    double a[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = (i < nb) ? X[k + ld * i] : 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (i <= j && j < nb) t += a[i] * rs[i + nb * j];
      o[j] = t;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j < nb) X[k + ld * j] = o[j];
  }
}

// ---- small dense helpers on the host ----
void jacobi_eigh(int n, const std::vector<double> &Ain, int lda, std::vector<double> &w, std::vector<double> &V) {
  std::vector<double> A((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) A[i + (size_t)n * j] = (i <= j) ? Ain[i + (size_t)lda * j] : Ain[j + (size_t)lda * i];
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i + (size_t)n * i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) (i == j ? dg : off) += A[i + (size_t)n * j] * A[i + (size_t)n * j];
    if (off <= 1e-30 * std::max(dg, 1e-300)) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p + (size_t)n * q];
        if (std::fabs(apq) < 1e-300) continue;
        double theta = (A[q + (size_t)n * q] - A[p + (size_t)n * p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {
          double akp = A[k + (size_t)n * p], akq = A[k + (size_t)n * q];
          A[k + (size_t)n * p] = cs * akp - sn * akq;
          A[k + (size_t)n * q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[p + (size_t)n * k], aqk = A[q + (size_t)n * k];
          A[p + (size_t)n * k] = cs * apk - sn * aqk;
          A[q + (size_t)n * k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k + (size_t)n * p], vkq = V[k + (size_t)n * q];
          V[k + (size_t)n * p] = cs * vkp - sn * vkq;
          V[k + (size_t)n * q] = sn * vkp + cs * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A[i + (size_t)n * i];
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j)
      if (w[j] < w[m]) m = j;
    if (m != i) {
      std::swap(w[i], w[m]);
      for (int k = 0; k < n; ++k) std::swap(V[k + (size_t)n * i], V[k + (size_t)n * m]);
    }
  }
}

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct Dav {
  qpx_ctx *c;
  hipStream_t st;
  int64_t n;
  DBuf<double> part;
  DBuf<double> small;  // up to 32*32 results
  int grid_rows;

  explicit Dav(qpx_ctx *cc) : c(cc), st(cc->stream), n(cc->W.n_det) {
    part.alloc((size_t)RBLOCKS * 32 * 32);
    small.alloc(32 * 32 + 64);
    grid_rows = (int)std::min<int64_t>((n + 255) / 256, 8192);
  }

  // C (na x nb, column-major, host) = A^T B; with B2, C is na x 2nb = A^T [B B2]
  void gram(const double *A, int64_t lda, int na, const double *B, int64_t ldb, int nb, std::vector<double> &C,
            const double *B2 = nullptr) {
    int nblk = (int)std::min<int64_t>(RBLOCKS, std::max<int64_t>(1, n / 256));
    hipLaunchKernelGGL(gram_partial_kernel, dim3(nblk), dim3(RB), 0, st, A, lda, na, B, B2, ldb, nb, n, part.p);
    if (B2) nb *= 2;
    int nout = na * nb;
    hipLaunchKernelGGL(reduce_partial_kernel, dim3(nout), dim3(RB), 0, st, part.p, nblk, nout, small.p);
    QPX_HIP(hipGetLastError());
    std::vector<double> tmp(nout);
    QPX_HIP(hipMemcpyAsync(tmp.data(), small.p, sizeof(double) * nout, hipMemcpyDeviceToHost, st));
    QPX_HIP(hipStreamSynchronize(st));
    // tmp[i*nb+j] = A_i . B_j  -> column-major C(i,j)
    C.assign(nout, 0.0);
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < nb; ++j) C[i + (size_t)na * j] = tmp[(size_t)i * nb + j];
  }

  double dot(const double *x, const double *y) {
    std::vector<double> C;
    gram(x, n, 1, y, n, 1, C);
    return C[0];
  }

  // X (n x nb) <- X R^-1 with X^T X = R^T R; false (X untouched) when the
  // Cholesky factorisation breaks down
  bool chol_qr(double *X, int nb) {
    std::vector<double> G;
    gram(X, n, nb, X, n, nb, G);
    std::vector<double> R((size_t)nb * nb, 0.0);
    for (int j = 0; j < nb; ++j) {
      for (int i = 0; i <= j; ++i) {
        double t = G[i + (size_t)nb * j];
        for (int k = 0; k < i; ++k) t -= R[k + (size_t)nb * i] * R[k + (size_t)nb * j];
        if (i == j) {
          if (!(t > 1e-20 * std::max(1.0, G[j + (size_t)nb * j]))) return false;
          R[j + (size_t)nb * j] = std::sqrt(t);
        } else {
          R[i + (size_t)nb * j] = t / R[i + (size_t)nb * i];
        }
      }
    }
    // R^-1 (upper triangular)
    std::vector<double> Ri((size_t)nb * nb, 0.0);
    for (int j = 0; j < nb; ++j) {
      Ri[j + (size_t)nb * j] = 1.0 / R[j + (size_t)nb * j];
      for (int i = j - 1; i >= 0; --i) {
        double t = 0.0;
        for (int k = i + 1; k <= j; ++k) t += R[i + (size_t)nb * k] * Ri[k + (size_t)nb * j];
        Ri[i + (size_t)nb * j] = -t / R[i + (size_t)nb * i];
      }
    }
    for (double v : Ri)
      if (!std::isfinite(v)) return false;
    QPX_HIP(hipMemcpyAsync(small.p, Ri.data(), sizeof(double) * Ri.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(tri_mul_kernel, dim3(grid_rows), dim3(256), 0, st, X, n, nb, small.p, n);
    QPX_HIP(hipGetLastError());
    return true;
  }

  // orthonormalise columns [c0, c0+nb) of U against [0, c0) and each other
  void orthonormalize(double *U, int c0, int nb) {
    std::vector<double> C;
    if (c0 > 0) {
      for (int pass = 0; pass < 2; ++pass) {
        gram(U, n, c0, U + n * c0, n, nb, C);
        QPX_HIP(hipMemcpyAsync(small.p, C.data(), sizeof(double) * C.size(), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(sub_gemm_kernel, dim3(grid_rows), dim3(256), 0, st, U + n * c0, n, nb, U, n, c0, small.p,
                           n);
        QPX_HIP(hipGetLastError());
      }
    }
    // within the block: Cholesky QR, twice (same orthonormal basis as
    // Gram-Schmidt in column order); Gram-Schmidt when the block is too
    // close to rank deficient for the Cholesky factor
    if (chol_qr(U + n * c0, nb) && chol_qr(U + n * c0, nb)) return;
    for (int j = 0; j < nb; ++j) {
      double *x = U + n * (c0 + j);
      for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i < j; ++i) {
          double *q = U + n * (c0 + i);
          double d = dot(q, x);
          hipLaunchKernelGGL(axpy_col_kernel, dim3(grid_rows), dim3(256), 0, st, x, q, -d, n);
        }
        if (c0 > 0 && pass == 1) {
          // one more projection against the old block keeps double precision
          gram(U, n, c0, x, n, 1, C);
          QPX_HIP(hipMemcpyAsync(small.p, C.data(), sizeof(double) * C.size(), hipMemcpyHostToDevice, st));
          hipLaunchKernelGGL(sub_gemm_kernel, dim3(grid_rows), dim3(256), 0, st, x, n, 1, U, n, c0, small.p, n);
        }
        double nr = std::sqrt(dot(x, x));
        QPX_REQUIRE(nr > 1e-300 && std::isfinite(nr), "Davidson: correction vector vanished or is not finite");
        hipLaunchKernelGGL(scale_col_kernel, dim3(grid_rows), dim3(256), 0, st, x, 1.0 / nr, n);
      }
      QPX_HIP(hipGetLastError());
    }
  }
};

}  // namespace

extern "C" int qpx_davidson(qpx_ctx *c, int N_st, int N_st_diag, double threshold, int sze_max, int s2_eig,
                            int state_following, int mixed, double switch_threshold, uint64_t seed, double *u_in,
                            int64_t dim_in, double *energies, double *s2_out, int *n_sigma_out) {
  return qpx_guard(c, [&] {
    QPX_HIP(hipSetDevice(c->device));
    Wfn &Wf = c->W;
    QPX_REQUIRE(Wf.valid, "no wavefunction set");
    const int64_t sze = Wf.n_det;
    QPX_REQUIRE(N_st >= 1 && N_st_diag >= N_st && N_st_diag <= 8, "need 1 <= N_st <= N_st_diag <= 8");
    QPX_REQUIRE(dim_in >= sze && u_in && energies && s2_out, "bad Davidson arguments");
    // diagonalization_hs2.irp.f:110-114
    QPX_REQUIRE((int64_t)N_st_diag * 3 <= sze, "error in Davidson: increase n_det_max_jacobi (N_st_diag*3 > sze)");
    int itermax = (int)std::max<int64_t>(3, std::min<int64_t>(sze_max, sze / N_st_diag));
    // The subspace holds at most 32 vectors (the projection and Ritz kernels
    // keep a row of it in registers / LDS): with more vectors per block the
    // restart comes earlier than the reference's davidson_sze_max would put
    // it.  32 / N_st_diag >= 4 for N_st_diag <= 8, so every restart still
    // keeps at least three blocks of history.
    QPX_REQUIRE(sze_max >= 1, "davidson_sze_max must be >= 1");
    itermax = std::min(itermax, 32 / N_st_diag);
    const int nc = N_st_diag * itermax;
    const double sz = 0.5 * (c->H.na - c->H.nb);
    const double sz2 = sz * (sz - 1.0);
    const double expected_s2 = sz * (sz + 1.0);
    hipStream_t st = c->stream;
    Dav D(c);
    const int64_t n = sze;
    // random extra guess vectors (reference: random_number + Box-Muller)
    uint64_t rs = seed;
    for (int k = N_st; k < N_st_diag; ++k)
      for (int64_t i = 0; i < sze; ++i) {
        double r1 = 1.0 - (double)(splitmix64(rs) >> 11) * (1.0 / 9007199254740992.0);
        double r2 = (double)(splitmix64(rs) >> 11) * (1.0 / 9007199254740992.0);
        u_in[i + dim_in * k] = std::sqrt(-2.0 * std::log(r1)) * std::cos(6.283185307179586 * r2);
      }
    for (int k = 0; k < N_st_diag; ++k) {
      double nr = 0.0;
      for (int64_t i = 0; i < sze; ++i) nr += u_in[i + dim_in * k] * u_in[i + dim_in * k];
      QPX_REQUIRE(nr > 0.0, "Davidson guess vector is zero");
      nr = 1.0 / std::sqrt(nr);
      for (int64_t i = 0; i < sze; ++i) u_in[i + dim_in * k] *= nr;
    }
    DBuf<double> U, W, S, uin_d, host_io;
    U.alloc((size_t)n * nc);
    W.alloc((size_t)n * nc);
    S.alloc((size_t)n * nc);
    uin_d.alloc((size_t)n * N_st_diag);
    host_io.alloc((size_t)dim_in * N_st_diag);
    QPX_HIP(hipMemcpyAsync(host_io.p, u_in, sizeof(double) * dim_in * N_st_diag, hipMemcpyHostToDevice, st));
    int gridp = (int)std::min<int64_t>((n * N_st_diag + 255) / 256, 65536);
    hipLaunchKernelGGL(permute_in_kernel, dim3(gridp), dim3(256), 0, st, host_io.p, dim_in, Wf.perm_d.p, n,
                       N_st_diag, uin_d.p, n);
    QPX_HIP(hipGetLastError());
    // sigma staging (state-fastest)
    DBuf<char> tbuf;
    tbuf.alloc(sizeof(double) * (size_t)n * N_st_diag * 3);
    DBuf<double> partial;
    partial.alloc((size_t)RBLOCKS * 8);
    DBuf<double> ydev, lamdev, s2dev;
    ydev.alloc(32 * 8);
    lamdev.alloc(8);
    s2dev.alloc(8);

    std::vector<double> hm((size_t)nc * nc, 0.0), sm((size_t)nc * nc, 0.0), y((size_t)nc * nc, 0.0),
        lambda(nc, 0.0), s2(nc, 0.0), resid(N_st_diag, 0.0);
    bool use_fp32 = mixed != 0;
    int n64 = 0, n32 = 0;
    bool converged = false;
    int shift2 = 0;
    int iter_total = 0;
    constexpr int kMaxIterations = 1000;
    while (!converged) {
      QPX_HIP(hipMemcpyAsync(U.p, uin_d.p, sizeof(double) * n * N_st_diag, hipMemcpyDeviceToDevice, st));
      bool switched = false;
      for (int iter = 1; iter <= itermax - 1; ++iter) {
        const int shift = N_st_diag * (iter - 1);
        shift2 = N_st_diag * iter;
        D.orthonormalize(U.p, shift, N_st_diag);
        // sigma on the new block
        {
          int g = (int)std::min<int64_t>((n * N_st_diag + 255) / 256, 65536);
          if (use_fp32) {
            float *ut = (float *)tbuf.p, *vt = ut + n * N_st_diag, *stt = vt + n * N_st_diag;
            hipLaunchKernelGGL(to_t_kernel<float>, dim3(g), dim3(256), 0, st, U.p + n * shift, n, n, N_st_diag, ut);
            sigma_device(c, QPX_FP32, N_st_diag, ut, vt, stt, st);
            hipLaunchKernelGGL(from_t_kernel<float>, dim3(g), dim3(256), 0, st, vt, n, N_st_diag, W.p + n * shift, n);
            hipLaunchKernelGGL(from_t_kernel<float>, dim3(g), dim3(256), 0, st, stt, n, N_st_diag, S.p + n * shift, n);
            ++n32;
          } else {
            double *ut = (double *)tbuf.p, *vt = ut + n * N_st_diag, *stt = vt + n * N_st_diag;
            hipLaunchKernelGGL(to_t_kernel<double>, dim3(g), dim3(256), 0, st, U.p + n * shift, n, n, N_st_diag, ut);
            sigma_device(c, QPX_FP64, N_st_diag, ut, vt, stt, st);
            hipLaunchKernelGGL(from_t_kernel<double>, dim3(g), dim3(256), 0, st, vt, n, N_st_diag, W.p + n * shift, n);
            hipLaunchKernelGGL(from_t_kernel<double>, dim3(g), dim3(256), 0, st, stt, n, N_st_diag, S.p + n * shift, n);
            ++n64;
          }
          QPX_HIP(hipGetLastError());
        }
        // h = U^T W and s = U^T S: the columns before the new block are
        // unchanged since the last iteration, so only U^T [W S](:, new
        // block) is formed; h and s are kept symmetric
        {
          std::vector<double> G;
          D.gram(U.p, n, shift2, W.p + n * shift, n, N_st_diag, G, S.p + n * shift);
          for (int j = 0; j < N_st_diag; ++j) {
            const int cj = shift + j;
            for (int i = 0; i < shift2; ++i) {
              const double hv = G[i + (size_t)shift2 * j], sv = G[i + (size_t)shift2 * (N_st_diag + j)];
              hm[i + (size_t)nc * cj] = hv;
              sm[i + (size_t)nc * cj] = sv;
              if (i < shift) {
                hm[cj + (size_t)nc * i] = hv;
                sm[cj + (size_t)nc * i] = sv;
              }
            }
          }
        }
        std::vector<double> w, yy;
        jacobi_eigh(shift2, hm, nc, w, yy);
        for (int j = 0; j < shift2; ++j) {
          lambda[j] = w[j];
          for (int i = 0; i < shift2; ++i) y[i + (size_t)nc * j] = yy[i + (size_t)shift2 * j];
        }
        // S^2 of the Ritz vectors: y^T s_ y
        for (int k = 0; k < shift2; ++k) {
          double acc = 0.0;
          for (int i = 0; i < shift2; ++i) {
            double t = 0.0;
            for (int l = 0; l < shift2; ++l) t += sm[i + (size_t)nc * l] * y[l + (size_t)nc * k];
            acc += y[i + (size_t)nc * k] * t;
          }
          s2[k] = acc + sz2;
        }
        std::vector<int> ok(nc, 1);
        if (s2_eig)
          for (int k = 0; k < shift2; ++k) ok[k] = std::fabs(s2[k] - expected_s2) < 0.6;
        for (int k = 0; k < shift2; ++k) {
          if (ok[k]) continue;
          for (int l = k + 1; l < shift2; ++l) {
            if (!ok[l]) continue;
            for (int i = 0; i < shift2; ++i) std::swap(y[i + (size_t)nc * k], y[i + (size_t)nc * l]);
            std::swap(s2[k], s2[l]);
            std::swap(lambda[k], lambda[l]);
            ok[k] = 1;
            ok[l] = 0;
            break;
          }
        }
        if (state_following) {
          std::vector<double> ov((size_t)nc * nc, -1.0);
          std::vector<int> order(N_st_diag, 0);
          for (int k = 0; k < shift2; ++k)
            for (int i = 0; i < shift2; ++i) ov[k + (size_t)nc * i] = std::fabs(y[k + (size_t)nc * i]);
          for (int k = 0; k < N_st; ++k) {
            double cmax = -1.0;
            for (int i = 0; i < N_st; ++i)
              if (ov[i + (size_t)nc * k] > cmax) { cmax = ov[i + (size_t)nc * k]; order[k] = i; }
            for (int i = 0; i < N_st_diag; ++i) ov[order[k] + (size_t)nc * i] = -1.0;
          }
          std::vector<double> yc = y;
          std::vector<double> lc(lambda), sc(s2);
          for (int k = 0; k < N_st; ++k) {
            int l = order[k];
            if (k != l) {
              for (int i = 0; i < shift2; ++i) y[i + (size_t)nc * k] = yc[i + (size_t)nc * l];
              lambda[k] = lc[l];
              s2[k] = sc[l];
            }
          }
        }
        // y(:, 0:N_st_diag) packed shift2 x N_st_diag
        std::vector<double> yp((size_t)shift2 * N_st_diag);
        for (int j = 0; j < N_st_diag; ++j)
          for (int i = 0; i < shift2; ++i) yp[i + (size_t)shift2 * j] = y[i + (size_t)nc * j];
        QPX_HIP(hipMemcpyAsync(ydev.p, yp.data(), sizeof(double) * yp.size(), hipMemcpyHostToDevice, st));
        QPX_HIP(hipMemcpyAsync(lamdev.p, lambda.data(), sizeof(double) * N_st_diag, hipMemcpyHostToDevice, st));
        QPX_HIP(hipMemcpyAsync(s2dev.p, s2.data(), sizeof(double) * N_st_diag, hipMemcpyHostToDevice, st));
        RitzArgs R{U.p, W.p, S.p, n, n, shift2, N_st_diag, Wf.hdiag.p, sz2};
        int nblk = (int)std::min<int64_t>(RBLOCKS, std::max<int64_t>(1, (n + 255) / 256));
        hipLaunchKernelGGL(ritz_residual_kernel, dim3(nblk), dim3(RB), 0, st, R, ydev.p, lamdev.p, s2dev.p, partial.p);
        QPX_HIP(hipGetLastError());
        std::vector<double> ph((size_t)nblk * 8);
        QPX_HIP(hipMemcpyAsync(ph.data(), partial.p, sizeof(double) * ph.size(), hipMemcpyDeviceToHost, st));
        QPX_HIP(hipStreamSynchronize(st));
        double rmax = 0.0;
        for (int k = 0; k < N_st_diag; ++k) {
          double r = 0.0;
          for (int b = 0; b < nblk; ++b) r += ph[(size_t)b * 8 + k];
          resid[k] = r;
          if (k < N_st) rmax = std::max(rmax, r);
        }
        ++iter_total;
        if (c->verbosity > 0) {
          std::fprintf(stderr, "davidson %4d %s", iter_total, use_fp32 ? "fp32" : "fp64");
          for (int k = 0; k < N_st; ++k) std::fprintf(stderr, "  %18.10f %9.6f %10.3e", lambda[k], s2[k], resid[k]);
          std::fprintf(stderr, "\n");
          std::fflush(stderr);
        }
        for (int k = 0; k < N_st; ++k)
          QPX_REQUIRE(resid[k] <= 1e8 && std::isfinite(resid[k]), "Davidson failed (residual > 1e8)");
        QPX_REQUIRE(iter_total < kMaxIterations, "Davidson did not converge in 1000 iterations");
        if (use_fp32 && rmax < switch_threshold) {
          // restart on the Ritz vectors with FP64 sigma
          use_fp32 = false;
          switched = true;
          break;
        }
        converged = (iter >= 2) && (std::fabs(rmax) < threshold);
        if (converged) break;
      }
      // re-contract to u_in (Ritz vectors)
      std::vector<double> yp((size_t)shift2 * N_st_diag);
      for (int j = 0; j < N_st_diag; ++j)
        for (int i = 0; i < shift2; ++i) yp[i + (size_t)shift2 * j] = y[i + (size_t)nc * j];
      QPX_HIP(hipMemcpyAsync(ydev.p, yp.data(), sizeof(double) * yp.size(), hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(gemm_y_kernel, dim3(D.grid_rows), dim3(256), 0, st, U.p, n, shift2, ydev.p, N_st_diag, n,
                         uin_d.p, n);
      QPX_HIP(hipGetLastError());
      (void)switched;
    }
    hipLaunchKernelGGL(permute_out_kernel, dim3(gridp), dim3(256), 0, st, uin_d.p, n, Wf.perm_d.p, n, N_st_diag,
                       host_io.p, dim_in);
    QPX_HIP(hipGetLastError());
    QPX_HIP(hipMemcpyAsync(u_in, host_io.p, sizeof(double) * dim_in * N_st_diag, hipMemcpyDeviceToHost, st));
    QPX_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < N_st_diag; ++k) {
      energies[k] = lambda[k];
      s2_out[k] = s2[k];
    }
    if (n_sigma_out) {
      n_sigma_out[0] = n64;
      n_sigma_out[1] = n32;
    }
  });
}
