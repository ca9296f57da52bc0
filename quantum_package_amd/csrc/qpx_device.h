// qpx_device.h -- device helpers for 64-bit determinant strings (0-based
// orbitals inside kernels).
//
// Phase rule (src/Determinants/slater_rules.irp.f:171-313, the published
// Scemama-Giner bit-string method): a same-spin hole/particle pair (h,p)
// contributes the number of occupied orbitals of the initial string strictly
// between h and p; a same-spin double adds one more when the two pairs cross
// (a<c<b<d with a,b the sorted first pair and c,d the sorted second pair,
// holes and particles each taken in ascending order).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace qpx {

__host__ __device__ inline int64_t pair_index(int p, int q) {
  int lo = p < q ? p : q, hi = p < q ? q : p;
  return (int64_t)hi * (hi + 1) / 2 + lo;
}

// bits of word w inside [x, y) (global bit positions)
__device__ inline uint64_t range_mask(int w, int x, int y) {
  int s = x - 64 * w, e = y - 64 * w;
  if (s < 0) s = 0;
  if (e > 64) e = 64;
  if (s >= e) return 0;
  uint64_t hi = (e == 64) ? ~0ULL : ((1ULL << e) - 1);
  uint64_t lo = (1ULL << s) - 1;
  return hi & ~lo;
}

template <int NI>
__device__ inline int count_between(const uint64_t *a, int h, int p) {
  int lo = h < p ? h : p, hi = h < p ? p : h;
  int c = 0;
#pragma unroll
  for (int w = 0; w < NI; ++w) c += __popcll(a[w] & range_mask(w, lo + 1, hi));
  return c;
}

template <int NI>
__device__ inline int between_parity(const uint64_t *a, int h, int p) {
  return count_between<NI>(a, h, p) & 1;
}

template <int NI>
__device__ inline int double_parity(const uint64_t *a, int h1, int h2, int p1, int p2) {
  int n = count_between<NI>(a, h1, p1) + count_between<NI>(a, h2, p2);
  int A = h1 < p1 ? h1 : p1, B = h1 < p1 ? p1 : h1;
  int Cc = h2 < p2 ? h2 : p2, D = h2 < p2 ? p2 : h2;
  if (A < Cc && Cc < B && B < D) n += 1;
  return n & 1;
}

// single excitation a -> b: hole h (in a), particle p (in b)
template <int NI>
__device__ inline void single_hp(const uint64_t *a, const uint64_t *b, int *h, int *p) {
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    uint64_t x = a[w] ^ b[w];
    uint64_t hh = x & a[w], pp = x & b[w];
    if (hh) *h = 64 * w + __ffsll((unsigned long long)hh) - 1;
    if (pp) *p = 64 * w + __ffsll((unsigned long long)pp) - 1;
  }
}

// double excitation a -> b (same spin): holes h1<h2, particles p1<p2
template <int NI>
__device__ inline void double_hp(const uint64_t *a, const uint64_t *b, int *h1, int *h2, int *p1, int *p2) {
  int nh = 0, np = 0;
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    uint64_t x = a[w] ^ b[w];
    uint64_t hh = x & a[w], pp = x & b[w];
    while (hh) {
      int o = 64 * w + __ffsll((unsigned long long)hh) - 1;
      hh &= hh - 1;
      if (nh == 0) *h1 = o; else *h2 = o;
      ++nh;
    }
    while (pp) {
      int o = 64 * w + __ffsll((unsigned long long)pp) - 1;
      pp &= pp - 1;
      if (np == 0) *p1 = o; else *p2 = o;
      ++np;
    }
  }
}

// <I|H|I> from orbital occupations: one-electron terms, same-spin pairs
// J-K, opposite-spin pairs J.
template <int NI>
__device__ inline double diag_energy(const uint64_t *a, const uint64_t *b, int M, const double *hcore,
                                     const double *J, const double *K) {
  double e = 0.0;
  for (int s = 0; s < 2; ++s) {
    const uint64_t *x = s ? b : a;
#pragma unroll
    for (int w = 0; w < NI; ++w) {
      uint64_t y = x[w];
      while (y) {
        int i = 64 * w + __ffsll((unsigned long long)y) - 1;
        y &= y - 1;
        e += hcore[(size_t)i * M + i];
        // same spin j > i
        for (int w2 = w; w2 < NI; ++w2) {
          uint64_t z = x[w2] & range_mask(w2, i + 1, 64 * NI);
          while (z) {
            int j = 64 * w2 + __ffsll((unsigned long long)z) - 1;
            z &= z - 1;
            e += J[(size_t)i * M + j] - K[(size_t)i * M + j];
          }
        }
        if (s == 0) {
          for (int w2 = 0; w2 < NI; ++w2) {
            uint64_t z = b[w2];
            while (z) {
              int j = 64 * w2 + __ffsll((unsigned long long)z) - 1;
              z &= z - 1;
              e += J[(size_t)i * M + j];
            }
          }
        }
      }
    }
  }
  return e;
}

}  // namespace qpx
