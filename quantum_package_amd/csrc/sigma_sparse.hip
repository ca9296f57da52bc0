// sigma_sparse.hip -- v = H u, s = S^2 u for selected wavefunctions whose
// unique spin-string sets are too large for the dense-map path.
//
// Reference: H_S2_u_0_nstates_openmp_work (src/Davidson/u0Hu0.irp.f:97-458)
// finds the partners of a determinant by scanning whole columns and rows of
// the alpha x beta matrix and testing excitation degrees.  Selected
// wavefunctions have a few very long rows and columns (the alpha and beta
// strings close to HF), so that scan costs orders of magnitude more than
// the partners it finds.  Here the partners are found exactly instead,
// through shared sub-strings: two spin strings are single excitations of
// each other iff they share an (n-1)-electron sub-string, double
// excitations iff they share an (n-2)-electron one, and that shared
// sub-string is unique.  So every connected pair meets in exactly one
// bucket of one of five classes:
//
//   ab  (alpha minus one, beta minus one)  alpha/beta doubles
//   a1  (alpha minus one, beta)            alpha singles
//   a2  (alpha minus two, beta)            alpha doubles
//   b1  (alpha, beta minus one)            beta singles
//   b2  (alpha, beta minus two)            beta doubles
//
// Set-up (once per wavefunction, on first use): the sub-string ids of every
// unique string (host sort), then per class one (key, slot) entry per
// determinant and removed orbital(s), sorted by key on the device (hipCUB
// radix sort); a bucket is a run of equal keys.  Each entry keeps its
// determinant and the orbital(s) it removed, and each (determinant, slot)
// knows its bucket's range.
//
// Sigma: one wavefront per determinant walks the buckets of all its slots
// (64 slots at a time, their entries concatenated and spread over the lanes
// by a prefix-sum search); an entry is a partner when the orbitals it
// removed differ from this slot's, and then hole and particle orbitals are
// known without touching the partner's strings.  Matrix elements and
// phases follow the same rules as the other paths (qpx_device.h).
#include <algorithm>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>
#include <hipcub/hipcub.hpp>
#include "qpx_internal.h"
#include "qpx_device.h"

using namespace qpx;

namespace {

constexpr int SP_MAXE = 32;    // electrons per spin handled by this path
constexpr int SP_MAXST = 8;
// ids of the (ne - drop)-electron sub-strings of n strings: ids[s * nslot +
// slot], slot = index of the removed orbital (drop 1) or of the removed pair
// i < j in lexicographic order over the occupied list (drop 2)
static int substring_ids(const std::vector<uint64_t> &str, int n, int nint, int ne, int drop,
                         std::vector<int> &ids) {
  const int nslot = drop == 1 ? ne : ne * (ne - 1) / 2;
  struct E {
    uint64_t w[4];
    int64_t at;
  };
  std::vector<E> v((size_t)n * nslot);
  std::vector<int> occ(ne);
  for (int s = 0; s < n; ++s) {
    const uint64_t *x = &str[(size_t)s * nint];
    int k = 0;
    for (int w = 0; w < nint; ++w)
      for (uint64_t y = x[w]; y; y &= y - 1) occ[k++] = 64 * w + __builtin_ctzll(y);
    int sl = 0;
    for (int i = 0; i < ne; ++i) {
      if (drop == 1) {
        E &e = v[(size_t)s * nslot + sl];
        for (int w = 0; w < 4; ++w) e.w[w] = w < nint ? x[w] : 0;
        e.w[occ[i] >> 6] &= ~(1ULL << (occ[i] & 63));
        e.at = (int64_t)s * nslot + sl;
        ++sl;
        continue;
      }
      for (int j = i + 1; j < ne; ++j) {
        E &e = v[(size_t)s * nslot + sl];
        for (int w = 0; w < 4; ++w) e.w[w] = w < nint ? x[w] : 0;
        e.w[occ[i] >> 6] &= ~(1ULL << (occ[i] & 63));
        e.w[occ[j] >> 6] &= ~(1ULL << (occ[j] & 63));
        e.at = (int64_t)s * nslot + sl;
        ++sl;
      }
    }
  }
  std::sort(v.begin(), v.end(), [](const E &a, const E &b) {
    for (int w = 3; w >= 0; --w)
      if (a.w[w] != b.w[w]) return a.w[w] < b.w[w];
    return false;
  });
  ids.assign(v.size(), 0);
  int id = -1;
  for (size_t i = 0; i < v.size(); ++i) {
    if (i == 0 || std::memcmp(v[i].w, v[i - 1].w, sizeof(v[i].w)) != 0) ++id;
    ids[v[i].at] = id;
  }
  return id + 1;
}

// the k-th set bit (0-based) of a string
template <int NI>
__device__ inline int kth_bit(const uint64_t *x, int k) {
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    const int c = __popcll(x[w]);
    if (k < c) {
      uint64_t y = x[w];
      for (int q = 0; q < k; ++q) y &= y - 1;
      return 64 * w + __ffsll((unsigned long long)y) - 1;
    }
    k -= c;
  }
  return -1;
}

__device__ inline void pair_of(int sl, int ne, int *i, int *j) {   // lexicographic (i < j) of slot sl
  int a = 0;
  while (sl >= ne - 1 - a) {
    sl -= ne - 1 - a;
    ++a;
  }
  *i = a;
  *j = a + 1 + sl;
}

struct ClassArgs {
  int cls;        // 0 ab, 1 a1, 2 a2, 3 b1, 4 b2
  int nslot;
  const int *sa;  // alpha sub-string ids (n_a * nslot_a) or null
  const int *sb;
  int nsa, nsb;   // slots per string of sa / sb
  int64_t mult;   // key = hi * mult + lo
};

// one entry per (determinant, slot): key and original index
__global__ void bucket_keys_kernel(ClassArgs C, const int *__restrict__ csc_a, const int *__restrict__ csc_b,
                                   int64_t n_det, uint64_t *__restrict__ key, int *__restrict__ orig) {
  const int64_t n = n_det * C.nslot;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = x / C.nslot;
    const int sl = (int)(x - k * C.nslot);
    const int a = csc_a[k], b = csc_b[k];
    uint64_t hi, lo;
    switch (C.cls) {
      case 0: hi = (uint64_t)C.sa[(int64_t)a * C.nsa + sl / C.nsb]; lo = (uint64_t)C.sb[(int64_t)b * C.nsb + sl % C.nsb]; break;
      case 1: case 2: hi = (uint64_t)C.sa[(int64_t)a * C.nsa + sl]; lo = (uint64_t)b; break;
      default: hi = (uint64_t)a; lo = (uint64_t)C.sb[(int64_t)b * C.nsb + sl]; break;
    }
    key[x] = hi * (uint64_t)C.mult + lo;
    orig[x] = (int)x;
  }
}

__global__ void bucket_heads_kernel(const uint64_t *__restrict__ key, int64_t n, int *__restrict__ head) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
    head[p] = (p == 0 || key[p] != key[p - 1]) ? 1 : 0;
}

__global__ void bucket_starts_kernel(const int *__restrict__ head, const int *__restrict__ sid, int64_t n,
                                     int *__restrict__ start) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
    if (head[p]) start[sid[p] - 1] = (int)p;
}

// per-block partial sums of len^2 over the buckets: the number of entries
// the walk visits for this class (every member visits its whole bucket)
__global__ void bucket_visits_kernel(const int *__restrict__ start, int nseg, int64_t n,
                                     long long *__restrict__ partial) {
  __shared__ long long red[4];
  long long acc = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nseg; q += (int64_t)gridDim.x * blockDim.x) {
    const long long len = (q + 1 < nseg ? (int64_t)start[q + 1] : n) - start[q];
    acc += len * len;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// payload of sorted position p: determinant | removed orbital(s) << 32 /
// << 40; range of the (determinant, slot) that sits at p
template <int NI>
__global__ void bucket_finalize_kernel(ClassArgs C, const int *__restrict__ orig, const int *__restrict__ sid,
                                       const int *__restrict__ start, int64_t n, int nseg,
                                       const int *__restrict__ csc_a, const int *__restrict__ csc_b,
                                       const uint64_t *__restrict__ astr, const uint64_t *__restrict__ bstr,
                                       int ne_a, int ne_b, uint64_t *__restrict__ pay, int2 *__restrict__ range) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const int o = orig[p];
    const int64_t k = o / C.nslot;
    const int sl = o - (int)(k * C.nslot);
    const int s = sid[p] - 1;
    range[o] = make_int2(start[s], s + 1 < nseg ? start[s + 1] : (int)n);
    const uint64_t *aw = astr + (int64_t)csc_a[k] * NI, *bw = bstr + (int64_t)csc_b[k] * NI;
    int o1 = 0, o2 = 0;
    switch (C.cls) {
      case 0: o1 = kth_bit<NI>(aw, sl / C.nsb); o2 = kth_bit<NI>(bw, sl % C.nsb); break;
      case 1: o1 = kth_bit<NI>(aw, sl); break;
      case 2: { int i, j; pair_of(sl, ne_a, &i, &j); o1 = kth_bit<NI>(aw, i); o2 = kth_bit<NI>(aw, j); } break;
      case 3: o1 = kth_bit<NI>(bw, sl); break;
      default: { int i, j; pair_of(sl, ne_b, &i, &j); o1 = kth_bit<NI>(bw, i); o2 = kth_bit<NI>(bw, j); } break;
    }
    pay[p] = (uint64_t)(uint32_t)k | ((uint64_t)(o1 & 0xff) << 32) | ((uint64_t)(o2 & 0xff) << 40);
  }
}

// single excitation h -> p of the string `same` (other spin `other`):
// h_hp + sum_{k in same} GS[hp][k] + sum_{k in other} GO[hp][k]
template <typename T, int NI>
__device__ inline T single_elem_sp(const uint64_t *same, const uint64_t *other, int h, int p, const T *GS,
                                   const T *GO, const T *hcore, int M) {
  const int64_t row = ((int64_t)h * M + p) * M;
  T e = hcore[h * M + p];
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    uint64_t y = same[w];
    while (y) {
      const int k = 64 * w + __ffsll((unsigned long long)y) - 1;
      y &= y - 1;
      e += GS[row + k];
    }
    y = other[w];
    while (y) {
      const int k = 64 * w + __ffsll((unsigned long long)y) - 1;
      y &= y - 1;
      e += GO[row + k];
    }
  }
  return e;
}

struct SparseArgs {
  const int *csc_a, *csc_b;
  const uint64_t *astr, *bstr;
  const double *hdiag, *sdiag;
  const int2 *range[5];
  const uint64_t *pay[5];
  int nslot[5];
  int64_t n_det;
  int M, npair, ne_a, ne_b;
};

template <typename T, int NI, int NST>
__global__ __launch_bounds__(256) void sigma_sparse_kernel(SparseArgs A, const T *__restrict__ V,
                                                           const T *__restrict__ G4, const T *__restrict__ GS,
                                                           const T *__restrict__ GO, const T *__restrict__ hcore,
                                                           const T *__restrict__ u, T *__restrict__ v,
                                                           T *__restrict__ s) {
  __shared__ uint8_t occ_s[4][2][SP_MAXE];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs; give
  // each XCD a contiguous run of determinants, whose partners overlap, so
  // that they share that XCD's L2
  const int nblk = (int)gridDim.x, base = nblk >> 3, rem = nblk & 7;
  const int xcd = (int)(blockIdx.x & 7);
  const int logical = xcd * base + min(xcd, rem) + (int)(blockIdx.x >> 3);
  const int64_t k = (int64_t)logical * (blockDim.x >> 6) + wv;
  if (k >= A.n_det) return;   // whole wavefront
  constexpr int nst = NST;
  const int M = A.M, npair = A.npair;
  const int a = A.csc_a[k], b = A.csc_b[k];
  uint64_t aw[NI], bw[NI];
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    aw[w] = A.astr[(int64_t)a * NI + w];
    bw[w] = A.bstr[(int64_t)b * NI + w];
  }
  if (lane < A.ne_a) occ_s[wv][0][lane] = (uint8_t)kth_bit<NI>(aw, lane);
  if (lane < A.ne_b) occ_s[wv][1][lane] = (uint8_t)kth_bit<NI>(bw, lane);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  T acc[NST], accs[NST];
#pragma unroll
  for (int st = 0; st < NST; ++st) acc[st] = accs[st] = T(0);

  int cum[6];
  cum[0] = 0;
#pragma unroll
  for (int c = 0; c < 5; ++c) cum[c + 1] = cum[c] + A.nslot[c];
  const int S = cum[5];
  for (int s0 = 0; s0 < S; s0 += 64) {
    const int sx = s0 + lane;
    int cls = 0, sl = 0, len = 0, beg = 0;
    if (sx < S) {
      cls = (sx >= cum[1]) + (sx >= cum[2]) + (sx >= cum[3]) + (sx >= cum[4]);
      sl = sx - cum[cls];
      const int2 r = A.range[cls][k * A.nslot[cls] + sl];
      beg = r.x;
      len = r.y - r.x;
    }
    int pre = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(pre, off, 64);
      if (lane >= off) pre += y;
    }
    const int total = __shfl(pre, 63, 64);
    // flat index f of slot p maps to entry f + off_p of its bucket
    const int off = beg - (pre - len);
    const int csl = (cls << 24) | sl;
    for (int f0 = 0; f0 < total; f0 += 64) {
      const int f = f0 + lane;
      int lo = 0, hi = 63;
#pragma unroll
      for (int it = 0; it < 6; ++it) {
        const int mid = (lo + hi) >> 1;
        const int pm = __shfl(pre, mid, 64);
        if (lo < hi) {
          if (pm > f) hi = mid;
          else lo = mid + 1;
        }
      }
      const int off_p = __shfl(off, lo, 64), csl_p = __shfl(csl, lo, 64);
      const int cls_p = csl_p >> 24, sl_p = csl_p & 0xffffff;
      if (f >= total) continue;
      const int e = f + off_p;
      const uint64_t py = A.pay[cls_p][e];
      const int64_t d2 = (int64_t)(uint32_t)py;
      const int x = (int)((py >> 32) & 0xff), y = (int)((py >> 40) & 0xff);
      T hij;
      bool flip = false, neg = false;
      if (cls_p == 0) {
        const int i = occ_s[wv][0][sl_p / A.ne_b], j = occ_s[wv][1][sl_p % A.ne_b];
        if (x == i || y == j) continue;
        // alpha i -> x, beta j -> y
        neg = between_parity<NI>(aw, i, x) ^ between_parity<NI>(bw, j, y);
        hij = G4 ? G4[((i * M + j) * M + y) * M + x] : V[pair_index(i, x) * npair + pair_index(j, y)];
        flip = (i == y) && (x == j);
      } else if (cls_p == 1) {
        // the spin strings are passed through explicit branches: choosing a
        // pointer between the two private arrays gave wrong beta-double
        // phases in the FP32 instantiation
        const int i = occ_s[wv][0][sl_p];
        if (x == i) continue;
        hij = single_elem_sp<T, NI>(aw, bw, i, x, GS, GO, hcore, M);
        neg = between_parity<NI>(aw, i, x);
      } else if (cls_p == 3) {
        const int i = occ_s[wv][1][sl_p];
        if (x == i) continue;
        hij = single_elem_sp<T, NI>(bw, aw, i, x, GS, GO, hcore, M);
        neg = between_parity<NI>(bw, i, x);
      } else {
        int ii, jj, h1, h2;
        if (cls_p == 2) {
          pair_of(sl_p, A.ne_a, &ii, &jj);
          h1 = occ_s[wv][0][ii];
          h2 = occ_s[wv][0][jj];
        } else {
          pair_of(sl_p, A.ne_b, &ii, &jj);
          h1 = occ_s[wv][1][ii];
          h2 = occ_s[wv][1][jj];
        }
        if (x == h1 || x == h2 || y == h1 || y == h2) continue;
        // (h1 x|h2 y) - (h1 y|h2 x), the second as (h2 x|h1 y) so that both
        // reads run along x
        hij = G4 ? G4[((h1 * M + h2) * M + y) * M + x] - G4[((h2 * M + h1) * M + y) * M + x]
                 : V[pair_index(h1, x) * npair + pair_index(h2, y)] - V[pair_index(h1, y) * npair + pair_index(h2, x)];
        neg = (cls_p == 2) ? double_parity<NI>(aw, h1, h2, x, y) : double_parity<NI>(bw, h1, h2, x, y);
      }
      if (neg) hij = -hij;
      T uu[NST];
#pragma unroll
      for (int st = 0; st < NST; ++st) uu[st] = u[d2 * NST + st];
#pragma unroll
      for (int st = 0; st < NST; ++st) {
        acc[st] += hij * uu[st];
        if (flip) accs[st] += neg ? uu[st] : -uu[st];
      }
    }
  }
#pragma unroll
  for (int st = 0; st < NST; ++st) {
    T xx = acc[st], yy = accs[st];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      xx += __shfl_xor(xx, off, 64);
      yy += __shfl_xor(yy, off, 64);
    }
    acc[st] = xx;
    accs[st] = yy;
  }
  if (lane == 0) {
    const T hd = (T)A.hdiag[k], sd = (T)A.sdiag[k];
#pragma unroll
    for (int st = 0; st < NST; ++st) {
      const T uu = u[k * NST + st];
      v[k * NST + st] = acc[st] + hd * uu;
      s[k * NST + st] = accs[st] + sd * uu;
    }
  }
}

template <int NI>
static void build_class(qpx_ctx *c, Buckets &Bk, ClassArgs C) {
  Wfn &W = c->W;
  hipStream_t st = c->stream;
  const int64_t n = W.n_det * C.nslot;
  QPX_REQUIRE(n < ((int64_t)1 << 31), "sparse sigma: too many (determinant, slot) entries");
  Bk.nslot = C.nslot;
  Bk.visited = 0;
  if (n == 0) return;   // e.g. no same-spin doubles with one electron of a spin
  DBuf<uint64_t> key, key2;
  DBuf<int> orig, orig2, head, sid, start;
  key.alloc(n);
  key2.alloc(n);
  orig.alloc(n);
  orig2.alloc(n);
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(bucket_keys_kernel, dim3(grid), dim3(256), 0, st, C, W.csc_a.p, W.csc_b.p, W.n_det, key.p,
                     orig.p);
  QPX_HIP(hipGetLastError());
  size_t tmp = 0;
  QPX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key.p, key2.p, orig.p, orig2.p, (int)n, 0, 64, st));
  DBuf<char> temp;
  temp.alloc(tmp);
  QPX_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, tmp, key.p, key2.p, orig.p, orig2.p, (int)n, 0, 64, st));
  key.release();
  orig.release();
  head.alloc(n);
  sid.alloc(n);
  hipLaunchKernelGGL(bucket_heads_kernel, dim3(grid), dim3(256), 0, st, key2.p, n, head.p);
  QPX_HIP(hipGetLastError());
  size_t tmp2 = 0;
  QPX_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp2, head.p, sid.p, (int)n, st));
  temp.alloc(std::max(tmp, tmp2));
  QPX_HIP(hipcub::DeviceScan::InclusiveSum(temp.p, tmp2, head.p, sid.p, (int)n, st));
  int nseg = 0;
  QPX_HIP(hipMemcpyAsync(&nseg, sid.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st));
  QPX_HIP(hipStreamSynchronize(st));
  start.alloc(nseg);
  hipLaunchKernelGGL(bucket_starts_kernel, dim3(grid), dim3(256), 0, st, head.p, sid.p, n, start.p);
  QPX_HIP(hipGetLastError());
  {
    const int vg = (int)std::min<int64_t>((nseg + 255) / 256, 1024);
    DBuf<long long> part;
    part.alloc(vg);
    hipLaunchKernelGGL(bucket_visits_kernel, dim3(vg), dim3(256), 0, st, start.p, nseg, n, part.p);
    QPX_HIP(hipGetLastError());
    std::vector<long long> h(vg);
    QPX_HIP(hipMemcpyAsync(h.data(), part.p, vg * sizeof(long long), hipMemcpyDeviceToHost, st));
    QPX_HIP(hipStreamSynchronize(st));
    Bk.visited = 0;
    for (long long x : h) Bk.visited += x;
  }
  key2.release();
  Bk.pay.alloc(n);
  Bk.range.alloc(n);
  hipLaunchKernelGGL(bucket_finalize_kernel<NI>, dim3(grid), dim3(256), 0, st, C, orig2.p, sid.p, start.p, n, nseg,
                     W.csc_a.p, W.csc_b.p, W.astr.p, W.bstr.p, c->H.na, c->H.nb, Bk.pay.p, Bk.range.p);
  QPX_HIP(hipGetLastError());
  QPX_HIP(hipStreamSynchronize(st));
}

template <int NI>
static void build_buckets(qpx_ctx *c) {
  Wfn &W = c->W;
  const int na = c->H.na, nb = c->H.nb;
  std::vector<int> a1, a2, b1, b2;
  const int na1 = substring_ids(W.astr_h, W.n_a, NI, na, 1, a1);
  const int na2 = na >= 2 ? substring_ids(W.astr_h, W.n_a, NI, na, 2, a2) : 0;
  const int nb1 = substring_ids(W.bstr_h, W.n_b, NI, nb, 1, b1);
  const int nb2 = nb >= 2 ? substring_ids(W.bstr_h, W.n_b, NI, nb, 2, b2) : 0;
  (void)na1; (void)na2;
  SparseIdx &X = W.sp;
  hipStream_t st = c->stream;
  X.a1.upload(a1, st);
  X.a2.upload(a2, st);
  X.b1.upload(b1, st);
  X.b2.upload(b2, st);
  const int na2s = na * (na - 1) / 2, nb2s = nb * (nb - 1) / 2;
  build_class<NI>(c, X.cls[0], ClassArgs{0, na * nb, X.a1.p, X.b1.p, na, nb, (int64_t)std::max(nb1, 1)});
  build_class<NI>(c, X.cls[1], ClassArgs{1, na, X.a1.p, nullptr, na, 0, (int64_t)W.n_b});
  build_class<NI>(c, X.cls[2], ClassArgs{2, na2s, X.a2.p, nullptr, na2s, 0, (int64_t)W.n_b});
  build_class<NI>(c, X.cls[3], ClassArgs{3, nb, nullptr, X.b1.p, 0, nb, (int64_t)std::max(nb1, 1)});
  build_class<NI>(c, X.cls[4], ClassArgs{4, nb2s, nullptr, X.b2.p, 0, nb2s, (int64_t)std::max(nb2, 1)});
  X.valid = true;
}

template <typename T, int NI>
static void launch_sparse(qpx_ctx *c, int nst, const T *u, T *v, T *s, hipStream_t st) {
  Wfn &W = c->W;
  SparseArgs A;
  A.csc_a = W.csc_a.p; A.csc_b = W.csc_b.p; A.astr = W.astr.p; A.bstr = W.bstr.p;
  A.hdiag = W.hdiag.p; A.sdiag = W.sdiag.p;
  for (int q = 0; q < 5; ++q) {
    A.range[q] = W.sp.cls[q].range.p;
    A.pay[q] = W.sp.cls[q].pay.p;
    A.nslot[q] = W.sp.cls[q].nslot;
  }
  A.n_det = W.n_det; A.M = c->H.M; A.npair = c->H.npair; A.ne_a = c->H.na; A.ne_b = c->H.nb;
  const int64_t blocks = (W.n_det + 3) / 4;
  QPX_REQUIRE(blocks < (1LL << 31), "too many determinants for one launch");
  const bool f64 = std::is_same<T, double>::value;
  const T *V = f64 ? (const T *)c->H.V.p : (const T *)c->H.Vf.p;
  const T *hc = f64 ? (const T *)c->H.hcore.p : (const T *)c->H.hcoref.p;
  const T *GS = f64 ? (const T *)c->H.GS.p : (const T *)c->H.GSf.p;
  const T *GO = f64 ? (const T *)c->H.GO.p : (const T *)c->H.GOf.p;
  // G4[((i*M + j)*M + y)*M + x] = (ix|jy); a bucket lists its partners in
  // determinant order (beta string, then alpha string), so consecutive
  // lanes share (i, j, y) and read neighbouring x.  Null above kG4MaxOrb.
  const T *G4 = f64 ? (const T *)c->H.G4.p : (const T *)c->H.G4f.p;
  if (c->ktiming) QPX_HIP(hipEventRecord(c->kev[0], st));
  const dim3 g((unsigned)blocks), bl(256);
  switch (nst) {
    case 1: hipLaunchKernelGGL((sigma_sparse_kernel<T, NI, 1>), g, bl, 0, st, A, V, G4, GS, GO, hc, u, v, s); break;
    case 2: hipLaunchKernelGGL((sigma_sparse_kernel<T, NI, 2>), g, bl, 0, st, A, V, G4, GS, GO, hc, u, v, s); break;
    case 3: hipLaunchKernelGGL((sigma_sparse_kernel<T, NI, 3>), g, bl, 0, st, A, V, G4, GS, GO, hc, u, v, s); break;
    case 4: hipLaunchKernelGGL((sigma_sparse_kernel<T, NI, 4>), g, bl, 0, st, A, V, G4, GS, GO, hc, u, v, s); break;
    case 5: hipLaunchKernelGGL((sigma_sparse_kernel<T, NI, 5>), g, bl, 0, st, A, V, G4, GS, GO, hc, u, v, s); break;
    case 6: hipLaunchKernelGGL((sigma_sparse_kernel<T, NI, 6>), g, bl, 0, st, A, V, G4, GS, GO, hc, u, v, s); break;
    case 7: hipLaunchKernelGGL((sigma_sparse_kernel<T, NI, 7>), g, bl, 0, st, A, V, G4, GS, GO, hc, u, v, s); break;
    default: hipLaunchKernelGGL((sigma_sparse_kernel<T, NI, 8>), g, bl, 0, st, A, V, G4, GS, GO, hc, u, v, s); break;
  }
  QPX_HIP(hipGetLastError());
  if (c->ktiming) {
    QPX_HIP(hipEventRecord(c->kev[1], st));
    QPX_HIP(hipEventRecord(c->kev[2], st));
  }
}

}  // namespace

namespace qpx {

bool sparse_ok(qpx_ctx *c, int nst) {
  const Wfn &W = c->W;
  return nst <= SP_MAXST && c->H.na <= SP_MAXE && c->H.nb <= SP_MAXE && c->H.na >= 1 && c->H.nb >= 1 &&
         c->H.M <= 256 && W.nint <= 4;
}

void sigma_sparse(qpx_ctx *c, int prec, int nst, const void *u, void *v, void *s, hipStream_t st) {
  Wfn &W = c->W;
  ensure_g4(c);   // partners in a bucket then read neighbouring integrals
  if (!W.sp.valid) {
    switch (W.nint) {
      case 1: build_buckets<1>(c); break;
      case 2: build_buckets<2>(c); break;
      case 3: build_buckets<3>(c); break;
      default: build_buckets<4>(c); break;
    }
  }
  if (prec == QPX_FP64) {
    switch (W.nint) {
      case 1: launch_sparse<double, 1>(c, nst, (const double *)u, (double *)v, (double *)s, st); break;
      case 2: launch_sparse<double, 2>(c, nst, (const double *)u, (double *)v, (double *)s, st); break;
      case 3: launch_sparse<double, 3>(c, nst, (const double *)u, (double *)v, (double *)s, st); break;
      default: launch_sparse<double, 4>(c, nst, (const double *)u, (double *)v, (double *)s, st); break;
    }
  } else {
    switch (W.nint) {
      case 1: launch_sparse<float, 1>(c, nst, (const float *)u, (float *)v, (float *)s, st); break;
      case 2: launch_sparse<float, 2>(c, nst, (const float *)u, (float *)v, (float *)s, st); break;
      case 3: launch_sparse<float, 3>(c, nst, (const float *)u, (float *)v, (float *)s, st); break;
      default: launch_sparse<float, 4>(c, nst, (const float *)u, (float *)v, (float *)s, st); break;
    }
  }
}

}  // namespace qpx

// benchmark hook: entries the sparse walk visits per class (ab, a1, a2, b1,
// b2) for the current wavefunction; builds the buckets if needed
extern "C" int qpx_debug_sparse_stats(qpx_ctx *c, int64_t *visited) {
  return qpx_guard(c, [&] {
    QPX_REQUIRE(visited, "null output");
    QPX_REQUIRE(c->ham_set && c->W.valid, "Hamiltonian and wavefunction must be set");
    QPX_HIP(hipSetDevice(c->device));
    QPX_REQUIRE(qpx::sparse_ok(c, 1), "the sparse path does not cover this system");
    Wfn &W = c->W;
    if (!W.sp.valid) {
      switch (W.nint) {
        case 1: build_buckets<1>(c); break;
        case 2: build_buckets<2>(c); break;
        case 3: build_buckets<3>(c); break;
        default: build_buckets<4>(c); break;
      }
    }
    for (int q = 0; q < 5; ++q) visited[q] = W.sp.cls[q].visited;
  });
}
