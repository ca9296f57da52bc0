// sigma.hip -- v = H u, s = S^2 u over the determinant space (MI355X).
//
// Reference: H_S2_u_0_nstates_openmp{,_work} (src/Davidson/u0Hu0.irp.f:
// 15-96,97-458).  The reference walks, for every determinant (a,b), (A) the
// beta-singles b' of b and inside each column b' the alpha-singles of a,
// and (B) the alpha singles/doubles inside column b and the beta
// singles/doubles inside row a, plus the diagonal; rows are owned by one
// thread (no write conflicts).  The same ownership is kept here, with two
// device strategies:
//
//  * dense-map path (unique alpha/beta sets small enough that one column of
//    u fits in LDS): one workgroup owns a beta column b.  For every
//    beta-single b' of b it stages column b' of u into LDS as a dense
//    alpha-indexed map (absent determinants read as zero) together with the
//    integral slice (h_a p_a | h_b p_b) for the beta excitation, and every
//    thread gathers its alpha-singles from LDS.  The same-spin alpha part
//    reads column b itself; the same-spin beta part runs as a second kernel
//    over alpha rows with a beta-indexed map.  u is read from HBM once per
//    staged column instead of once per connected pair.
//  * generic path (any size): one wavefront owns a determinant and scans,
//    with coalesced lane-strided loads, the column b and row a for
//    same-spin partners and, for the alpha/beta doubles, either the rows of
//    the alpha-singles of a or the columns of the beta-singles of b,
//    whichever holds fewer entries (the reference scans the columns).
//    Hits are reduced across the wavefront at the end.
//
// Matrix elements: alpha/beta double = phase_a phase_b (h_a p_a|h_b p_b);
// same-spin double = phase ((h1 p1|h2 p2) - (h1 p2|h2 p1)); single h->p =
// phase (h_hp + sum_{k in same-spin string} ((hp|kk)-(hk|kp)) + sum_{k in
// other-spin string} (hp|kk)); S^2 off-diagonal = -phase for the alpha/beta
// spin exchange (src/Determinants/s2.irp.f:32-63).
#include <algorithm>
#include "qpx_internal.h"
#include "qpx_device.h"

using namespace qpx;

template <typename T>
struct Tab;
template <>
struct Tab<double> {
  static const double *V(const Ham &H) { return H.V.p; }
  static const double *sval(const SpinConn &c) { return c.s_val.p; }
  static const double *dval(const SpinConn &c) { return c.d_val.p; }
  static const double *hcore(const Ham &H) { return H.hcore.p; }
};
template <>
struct Tab<float> {
  static const float *V(const Ham &H) { return H.Vf.p; }
  static const float *sval(const SpinConn &c) { return c.s_valf.p; }
  static const float *dval(const SpinConn &c) { return c.d_valf.p; }
  static const float *hcore(const Ham &H) { return H.hcoref.p; }
};

// ---------------------------------------------------------------------------
// dense-map path
// ---------------------------------------------------------------------------
struct DenseArgs {
  const uint64_t *astr, *bstr;
  const int *col_ptr, *csc_a, *csc_b, *row_ptr, *csr_b, *csr2csc;
  const int *as_ptr, *as_idx; const uint32_t *as_info;  // alpha singles
  const int *ad_ptr, *ad_idx;                           // alpha doubles
  const int *bs_ptr, *bs_idx; const uint32_t *bs_info;  // beta singles
  const int *bd_ptr, *bd_idx;
  const double *hdiag, *sdiag;
  int n_a, n_b, M, npair;
};

// Same-spin partners of the 64 determinants of a wavefront, 16 lanes per
// determinant: in round r, lane group q = lane/16 takes the determinant of
// lane 4r+q and walks its singles and doubles lists with stride 16, so the
// list loads are coalesced instead of one list per lane.  The partial sums
// are reduced over the 16 lanes and added to acc of the owning lane.  Every
// lane of the wavefront must call it (lanes without a determinant pass
// empty ranges).
template <typename T, int NST>
__device__ inline void same_spin_x16(const int *__restrict__ s_idx, const uint32_t *__restrict__ s_info,
                                     const T *__restrict__ s_val, int s0, int s1, const int *__restrict__ d_idx,
                                     const T *__restrict__ d_val, int d0, int d1, const T *map, const T *g, int M,
                                     T *acc) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, sub = lane & 15;
#pragma unroll 1
  for (int r = 0; r < 16; ++r) {
    const int src = 4 * r + q;
    const int rs0 = __shfl(s0, src), rs1 = __shfl(s1, src), rd0 = __shfl(d0, src), rd1 = __shfl(d1, src);
    T p[NST];
#pragma unroll
    for (int st = 0; st < NST; ++st) p[st] = T(0);
    for (int es = rs0 + sub; es < rs1; es += 16) {
      const int ap = s_idx[es];
      const uint32_t inf = s_info[es];
      const T gg = g[(inf & 0xff) * M + ((inf >> 8) & 0xff)];
      const T val = s_val[es] + (((inf >> 16) & 1) ? -gg : gg);
#pragma unroll
      for (int st = 0; st < NST; ++st) p[st] += val * map[ap * NST + st];
    }
    int ed = rd0 + sub;
    for (; ed + 16 < rd1; ed += 32) {
      const int ap0 = d_idx[ed], ap1 = d_idx[ed + 16];
      const T v0 = d_val[ed], v1 = d_val[ed + 16];
#pragma unroll
      for (int st = 0; st < NST; ++st) p[st] += v0 * map[ap0 * NST + st] + v1 * map[ap1 * NST + st];
    }
    if (ed < rd1) {
      const int ap0 = d_idx[ed];
      const T v0 = d_val[ed];
#pragma unroll
      for (int st = 0; st < NST; ++st) p[st] += v0 * map[ap0 * NST + st];
    }
#pragma unroll
    for (int st = 0; st < NST; ++st) {
#pragma unroll
      for (int off = 8; off > 0; off >>= 1) p[st] += __shfl_xor(p[st], off);
      const T t = __shfl(p[st], (lane & 3) * 16);
      if ((lane >> 2) == r) acc[st] += t;
    }
  }
}

// g(h,p) = sum_{k in string} (hp|kk) for all M*M (h,p), into LDS
template <typename T, int NI>
__device__ inline void coulomb_slice(const uint64_t *sw, const T *__restrict__ V, int npair, int M, T *out) {
  for (int x = threadIdx.x; x < M * M; x += blockDim.x) {
    int h = x / M, p = x - h * M;
    const int64_t P = pair_index(h, p);
    T g = T(0);
#pragma unroll
    for (int w = 0; w < NI; ++w) {
      uint64_t y = sw[w];
      while (y) {
        int k = 64 * w + __ffsll((unsigned long long)y) - 1;
        y &= y - 1;
        g += V[P * npair + pair_index(k, k)];
      }
    }
    out[x] = g;
  }
}

// Column kernel: workgroup = (beta column b, chunk of BS of its
// determinants), one determinant per thread.  The thread keeps the
// alpha-singles of its alpha string in registers, packed
// a' | (h*M + p) << 16 | phase << 30, for the whole beta-single loop;
// entries past its count point at a map slot that stays zero, so the inner
// loop runs branch-free to a wavefront-uniform trip count.
// LDS layout: map[(n_a + 1) * NST] (slot n_a = 0) | slice[M*M] |
// sacc[BS * NST] (each thread's S^2 accumulators, off the register file)
template <typename T, int NI, int NST, int MAXS, int BS>
__global__ __launch_bounds__(BS) void sigma_dense_col_kernel(DenseArgs A, int nchunk, const T *__restrict__ V,
                                                               const T *__restrict__ as_val,
                                                               const T *__restrict__ ad_val,
                                                               const T *__restrict__ u, T *__restrict__ v,
                                                               T *__restrict__ s) {
  static_assert(MAXS % 4 == 0, "alpha-singles are consumed in groups of 4");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *map = reinterpret_cast<T *>(lds_raw);
  T *slice = map + (size_t)(A.n_a + 1) * NST;
  T *sacc = slice + (size_t)A.M * A.M + (size_t)threadIdx.x * NST;
  // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so
  // logical block = (id mod 8) * (grid/8) + id/8 keeps the chunks of one
  // column, and neighbouring columns, on one XCD and its L2
  const int per_xcd = gridDim.x >> 3;
  const int logical = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (logical >= A.n_b * nchunk) return;
  const int b = logical / nchunk, chunk = logical - b * nchunk;
  const int c0 = A.col_ptr[b], c1 = A.col_ptr[b + 1];
  const int n_col = c1 - c0;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int M = A.M;
  const int e = c0 + chunk * nth + tid;
  const bool mine = e < c1;
  const int a = mine ? A.csc_a[e] : 0;
  int ns = 0, sbeg = 0;
  if (mine) {
    sbeg = A.as_ptr[a];
    ns = A.as_ptr[a + 1] - sbeg;
  }
  const uint32_t zero_slot = (uint32_t)A.n_a;
  if (tid < NST) map[(size_t)A.n_a * NST + tid] = T(0);
  uint32_t L[MAXS];
#pragma unroll
  for (int j = 0; j < MAXS; ++j) {
    uint32_t x = zero_slot;
    if (j < ns) {
      uint32_t inf = A.as_info[sbeg + j];
      x = (uint32_t)A.as_idx[sbeg + j] | (((inf & 0xff) * M + ((inf >> 8) & 0xff)) << 16) | (((inf >> 16) & 1) << 30);
    }
    L[j] = x;
  }
  int nsw = ns;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) nsw = max(nsw, __shfl_xor(nsw, off));
  nsw = __builtin_amdgcn_readfirstlane(nsw);
  T acc[NST];
#pragma unroll
  for (int st = 0; st < NST; ++st) acc[st] = sacc[st] = T(0);
  uint64_t bw[NI];
#pragma unroll
  for (int w = 0; w < NI; ++w) bw[w] = A.bstr[(int64_t)b * NI + w];

  // ---- (A) alpha/beta doubles: beta-singles b' of b ----
  for (int eb = A.bs_ptr[b]; eb < A.bs_ptr[b + 1]; ++eb) {
    const int bp = A.bs_idx[eb];
    const uint32_t binfo = A.bs_info[eb];
    const int hb = binfo & 0xff, pb = (binfo >> 8) & 0xff, phb = (binfo >> 16) & 1;
    const int l0 = A.col_ptr[bp], l1 = A.col_ptr[bp + 1];
    __syncthreads();
    if (l1 - l0 != A.n_a) {
      for (int x = tid; x < A.n_a * NST; x += nth) map[x] = T(0);
      __syncthreads();
    }
    if (l1 - l0 == A.n_a) {
      // full column: contiguous copy
      for (int x = tid; x < A.n_a * NST; x += nth) map[x] = u[(int64_t)l0 * NST + x];
    } else {
      for (int ee = l0 + tid; ee < l1; ee += nth) {
        int ai = A.csc_a[ee];
#pragma unroll
        for (int st = 0; st < NST; ++st) map[ai * NST + st] = u[(int64_t)ee * NST + st];
      }
    }
    const int64_t Pb = pair_index(hb, pb);
    for (int x = tid; x < M * M; x += nth) {
      int h = x / M, p = x - h * M;
      slice[x] = V[pair_index(h, p) * A.npair + Pb];
    }
    __syncthreads();
    // binfo has no bits above 16, so keep is zero; xor-ing it in keeps the
    // decode of L[j] inside this loop (hoisted, it would be spilled)
    const uint32_t keep = (binfo >> 24) * 0x01010101u;
    // the alpha single (pb -> hb) is the spin exchange of this beta single
    // (S^2 off-diagonal); at most one j matches, remembered in fo
    const uint32_t fidx = (uint32_t)(pb * M + hb);
    uint32_t fo = zero_slot;
#pragma unroll
    for (int j0 = 0; j0 < MAXS; j0 += 4) {
      if (j0 < nsw) {
        T uu[4][NST], sv[4];
        uint32_t xs[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t x = L[j0 + q] ^ keep;
          xs[q] = x;
          const T *mp = map + (x & 0xffff) * NST;
#pragma unroll
          for (int st = 0; st < NST; ++st) uu[q][st] = mp[st];
          sv[q] = slice[(x >> 16) & 0x3fff];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t negb = ((xs[q] >> 30) & 1) ^ (uint32_t)phb;
          const T val = negb ? -sv[q] : sv[q];
#pragma unroll
          for (int st = 0; st < NST; ++st) acc[st] += val * uu[q][st];
          fo = (((xs[q] >> 16) & 0x3fff) == fidx) ? ((xs[q] & 0xffff) | (negb << 31)) : fo;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    {
      const T *mp = map + (fo & 0xffff) * NST;
      const bool fneg = (fo >> 31) != 0;
#pragma unroll
      for (int st = 0; st < NST; ++st) {
        const T uu = mp[st];
        sacc[st] += fneg ? uu : -uu;
      }
    }
  }

  // ---- (B) same-spin alpha inside column b, and the diagonal ----
  __syncthreads();
  if (n_col != A.n_a) {
    for (int x = tid; x < A.n_a * NST; x += nth) map[x] = T(0);
    __syncthreads();
  }
  for (int ee = c0 + tid; ee < c1; ee += nth) {
    int ai = A.csc_a[ee];
#pragma unroll
    for (int st = 0; st < NST; ++st) map[ai * NST + st] = u[(int64_t)ee * NST + st];
  }
  coulomb_slice<T, NI>(bw, V, A.npair, M, slice);
  __syncthreads();
  same_spin_x16<T, NST>(A.as_idx, A.as_info, as_val, sbeg, sbeg + ns, A.ad_idx, ad_val, mine ? A.ad_ptr[a] : 0,
                        mine ? A.ad_ptr[a + 1] : 0, map, slice, M, acc);
  if (mine) {
    const T hd = (T)A.hdiag[e], sd = (T)A.sdiag[e];
#pragma unroll
    for (int st = 0; st < NST; ++st) {
      T uu = map[a * NST + st];
      v[(int64_t)e * NST + st] = acc[st] + hd * uu;
      s[(int64_t)e * NST + st] = sacc[st] + sd * uu;
    }
  }
}

// same-spin beta part: workgroup = (alpha row a, chunk of 1024), one
// determinant per thread; v += over the row's beta partners
template <typename T, int NI, int NST>
__global__ __launch_bounds__(1024) void sigma_dense_row_kernel(DenseArgs A, int nchunk, const T *__restrict__ V,
                                                               const T *__restrict__ bs_val,
                                                               const T *__restrict__ bd_val,
                                                               const T *__restrict__ u, T *__restrict__ v) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *map = reinterpret_cast<T *>(lds_raw);
  T *slice = map + (size_t)A.n_b * NST;
  const int per_xcd = gridDim.x >> 3;
  const int logical = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (logical >= A.n_a * nchunk) return;
  // chunk-major: the workgroups an XCD runs together share a chunk of beta
  // strings, whose same-spin connection lists then stay in its L2
  const int chunk = logical / A.n_a, a = logical - chunk * A.n_a;
  const int r0 = A.row_ptr[a], r1 = A.row_ptr[a + 1];
  const int n_row = r1 - r0;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int M = A.M;
  if (n_row != A.n_b) {
    for (int x = tid; x < A.n_b * NST; x += nth) map[x] = T(0);
    __syncthreads();
  }
  for (int t = r0 + tid; t < r1; t += nth) {
    int bi = A.csr_b[t];
    int e = A.csr2csc[t];
#pragma unroll
    for (int st = 0; st < NST; ++st) map[bi * NST + st] = u[(int64_t)e * NST + st];
  }
  uint64_t aw[NI];
#pragma unroll
  for (int w = 0; w < NI; ++w) aw[w] = A.astr[(int64_t)a * NI + w];
  coulomb_slice<T, NI>(aw, V, A.npair, M, slice);
  __syncthreads();
  const int t = r0 + chunk * nth + tid;
  if (r0 + chunk * nth + (tid & ~63) >= r1) return;  // whole wavefront past the row
  const bool mine = t < r1;
  const int b = mine ? A.csr_b[t] : 0;
  T acc[NST];
#pragma unroll
  for (int st = 0; st < NST; ++st) acc[st] = T(0);
  same_spin_x16<T, NST>(A.bs_idx, A.bs_info, bs_val, mine ? A.bs_ptr[b] : 0, mine ? A.bs_ptr[b + 1] : 0, A.bd_idx,
                        bd_val, mine ? A.bd_ptr[b] : 0, mine ? A.bd_ptr[b + 1] : 0, map, slice, M, acc);
  if (!mine) return;
  const int e = A.csr2csc[t];
#pragma unroll
  for (int st = 0; st < NST; ++st) v[(int64_t)e * NST + st] += acc[st];
}

// ---------------------------------------------------------------------------
// generic path: one wavefront per determinant
// ---------------------------------------------------------------------------
struct GenArgs {
  const uint64_t *astr, *bstr;
  const int *col_ptr, *csc_a, *csc_b, *row_ptr, *csr_b, *csr2csc;
  const int *as_ptr, *as_idx; const uint32_t *as_info;
  const int *bs_ptr, *bs_idx; const uint32_t *bs_info;
  const int64_t *cost_a, *cost_b;
  const double *hdiag, *sdiag;
  int64_t n_det;
  int M, npair, nst;
};

constexpr int GMAXST = 8;

template <typename T, int NI>
__device__ inline T single_elem(const uint64_t *same, const uint64_t *other, int h, int p, const T *V,
                                const T *hcore, int M, int npair) {
  const int64_t P = pair_index(h, p);
  T e = hcore[h * M + p];
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    uint64_t y = same[w];
    while (y) {
      int k = 64 * w + __ffsll((unsigned long long)y) - 1;
      y &= y - 1;
      e += V[P * npair + pair_index(k, k)] - V[pair_index(h, k) * npair + pair_index(k, p)];
    }
    y = other[w];
    while (y) {
      int k = 64 * w + __ffsll((unsigned long long)y) - 1;
      y &= y - 1;
      e += V[P * npair + pair_index(k, k)];
    }
  }
  return e;
}

template <typename T, int NI>
__global__ __launch_bounds__(256) void sigma_generic_kernel(GenArgs A, const T *__restrict__ V,
                                                            const T *__restrict__ hcore,
                                                            const T *__restrict__ u, T *__restrict__ v,
                                                            T *__restrict__ s) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (k >= A.n_det) return;
  const int nst = A.nst;
  const int M = A.M, npair = A.npair;
  const int a = A.csc_a[k], b = A.csc_b[k];
  uint64_t aw[NI], bw[NI];
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    aw[w] = A.astr[(int64_t)a * NI + w];
    bw[w] = A.bstr[(int64_t)b * NI + w];
  }
  T acc[GMAXST], accs[GMAXST];
#pragma unroll
  for (int st = 0; st < GMAXST; ++st) acc[st] = accs[st] = T(0);

  // same-spin alpha: scan column b
  for (int e = A.col_ptr[b] + lane; e < A.col_ptr[b + 1]; e += 64) {
    const int ai = A.csc_a[e];
    if (ai == a) continue;
    uint64_t xw[NI];
    int d = 0;
#pragma unroll
    for (int w = 0; w < NI; ++w) {
      xw[w] = A.astr[(int64_t)ai * NI + w];
      d += __popcll(aw[w] ^ xw[w]);
    }
    T hij;
    if (d == 2) {
      int h, p;
      single_hp<NI>(aw, xw, &h, &p);
      hij = single_elem<T, NI>(aw, bw, h, p, V, hcore, M, npair);
      if (between_parity<NI>(aw, h, p)) hij = -hij;
    } else if (d == 4) {
      int h1, h2, p1, p2;
      double_hp<NI>(aw, xw, &h1, &h2, &p1, &p2);
      hij = V[pair_index(h1, p1) * npair + pair_index(h2, p2)] - V[pair_index(h1, p2) * npair + pair_index(h2, p1)];
      if (double_parity<NI>(aw, h1, h2, p1, p2)) hij = -hij;
    } else {
      continue;
    }
#pragma unroll
    for (int st = 0; st < GMAXST; ++st)
      if (st < nst) acc[st] += hij * u[(int64_t)e * nst + st];
  }
  // same-spin beta: scan row a
  for (int t = A.row_ptr[a] + lane; t < A.row_ptr[a + 1]; t += 64) {
    const int bi = A.csr_b[t];
    if (bi == b) continue;
    uint64_t xw[NI];
    int d = 0;
#pragma unroll
    for (int w = 0; w < NI; ++w) {
      xw[w] = A.bstr[(int64_t)bi * NI + w];
      d += __popcll(bw[w] ^ xw[w]);
    }
    T hij;
    if (d == 2) {
      int h, p;
      single_hp<NI>(bw, xw, &h, &p);
      hij = single_elem<T, NI>(bw, aw, h, p, V, hcore, M, npair);
      if (between_parity<NI>(bw, h, p)) hij = -hij;
    } else if (d == 4) {
      int h1, h2, p1, p2;
      double_hp<NI>(bw, xw, &h1, &h2, &p1, &p2);
      hij = V[pair_index(h1, p1) * npair + pair_index(h2, p2)] - V[pair_index(h1, p2) * npair + pair_index(h2, p1)];
      if (double_parity<NI>(bw, h1, h2, p1, p2)) hij = -hij;
    } else {
      continue;
    }
    const int64_t e = A.csr2csc[t];
#pragma unroll
    for (int st = 0; st < GMAXST; ++st)
      if (st < nst) acc[st] += hij * u[e * nst + st];
  }
  // alpha/beta doubles.  The determinant's alpha-singles a' (or
  // beta-singles b') are taken 64 at a time; their rows (columns) are
  // concatenated and the wavefront's lanes stride over the concatenation,
  // each lane finding its partner by a binary search of the lanes' prefix
  // sums (shuffles).  Long and short rows are thus spread evenly.
  const bool by_rows = A.cost_a[a] <= A.cost_b[b];
  const int s0 = by_rows ? A.as_ptr[a] : A.bs_ptr[b], s1 = by_rows ? A.as_ptr[a + 1] : A.bs_ptr[b + 1];
  const int *sidx = by_rows ? A.as_idx : A.bs_idx;
  const uint32_t *sinfo = by_rows ? A.as_info : A.bs_info;
  const int *lptr = by_rows ? A.row_ptr : A.col_ptr;
  for (int es0 = s0; es0 < s1; es0 += 64) {
    const int es = es0 + lane;
    int len = 0, start = 0;
    uint32_t info = 0;
    if (es < s1) {
      const int xp = sidx[es];
      info = sinfo[es];
      start = lptr[xp];
      len = lptr[xp + 1] - start;
    }
    int pre = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(pre, off, 64);
      if (lane >= off) pre += y;
    }
    const int total = __shfl(pre, 63, 64);
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
      const int pre_p = __shfl(pre, lo, 64), len_p = __shfl(len, lo, 64), start_p = __shfl(start, lo, 64);
      const uint32_t inf = (uint32_t)__shfl((int)info, lo, 64);
      if (f >= total) continue;
      const int t = start_p + f - (pre_p - len_p);
      const int hs = inf & 0xff, ps = (inf >> 8) & 0xff, phs = (inf >> 16) & 1;
      uint64_t xw[NI];
      int d = 0;
      if (by_rows) {
        const int bi = A.csr_b[t];
#pragma unroll
        for (int w = 0; w < NI; ++w) {
          xw[w] = A.bstr[(int64_t)bi * NI + w];
          d += __popcll(bw[w] ^ xw[w]);
        }
      } else {
        const int ai = A.csc_a[t];
#pragma unroll
        for (int w = 0; w < NI; ++w) {
          xw[w] = A.astr[(int64_t)ai * NI + w];
          d += __popcll(aw[w] ^ xw[w]);
        }
      }
      if (d != 2) continue;
      int ha, pa, hb, pb;
      bool neg;
      int64_t e;
      if (by_rows) {
        ha = hs; pa = ps;
        single_hp<NI>(bw, xw, &hb, &pb);
        neg = phs ^ between_parity<NI>(bw, hb, pb);
        e = A.csr2csc[t];
      } else {
        hb = hs; pb = ps;
        single_hp<NI>(aw, xw, &ha, &pa);
        neg = phs ^ between_parity<NI>(aw, ha, pa);
        e = t;
      }
      T hij = V[pair_index(ha, pa) * npair + pair_index(hb, pb)];
      if (neg) hij = -hij;
      const bool flip = (ha == pb) && (pa == hb);
#pragma unroll
      for (int st = 0; st < GMAXST; ++st)
        if (st < nst) {
          T uu = u[e * nst + st];
          acc[st] += hij * uu;
          if (flip) accs[st] += neg ? uu : -uu;
        }
    }
  }
  // wavefront reduction
#pragma unroll
  for (int st = 0; st < GMAXST; ++st) {
    if (st >= nst) break;
    T x = acc[st], y = accs[st];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      x += __shfl_xor(x, off, 64);
      y += __shfl_xor(y, off, 64);
    }
    acc[st] = x;
    accs[st] = y;
  }
  if (lane == 0) {
    const T hd = (T)A.hdiag[k], sd = (T)A.sdiag[k];
#pragma unroll
    for (int st = 0; st < GMAXST; ++st)
      if (st < nst) {
        T uu = u[k * nst + st];
        v[k * nst + st] = acc[st] + hd * uu;
        s[k * nst + st] = accs[st] + sd * uu;
      }
  }
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
static DenseArgs dense_args(qpx_ctx *c) {
  Wfn &W = c->W;
  DenseArgs A;
  A.astr = W.astr.p; A.bstr = W.bstr.p;
  A.col_ptr = W.col_ptr.p; A.csc_a = W.csc_a.p; A.csc_b = W.csc_b.p;
  A.row_ptr = W.row_ptr.p; A.csr_b = W.csr_b.p; A.csr2csc = W.csr2csc.p;
  A.as_ptr = W.ca.s_ptr.p; A.as_idx = W.ca.s_idx.p; A.as_info = W.ca.s_info.p;
  A.ad_ptr = W.ca.d_ptr.p; A.ad_idx = W.ca.d_idx.p;
  A.bs_ptr = W.cb.s_ptr.p; A.bs_idx = W.cb.s_idx.p; A.bs_info = W.cb.s_info.p;
  A.bd_ptr = W.cb.d_ptr.p; A.bd_idx = W.cb.d_idx.p;
  A.hdiag = W.hdiag.p; A.sdiag = W.sdiag.p;
  A.n_a = W.n_a; A.n_b = W.n_b; A.M = c->H.M; A.npair = c->H.npair;
  return A;
}

template <typename T, int NI, int NST, int MAXS, int BS>
static void launch_dense(qpx_ctx *c, const T *u, T *v, T *s, hipStream_t st) {
  DenseArgs A = dense_args(c);
  const Ham &H = c->H;
  const Wfn &W = c->W;
  size_t lds_col = ((size_t)(W.n_a + 1) * NST + (size_t)H.M * H.M + (size_t)BS * NST) * sizeof(T);
  size_t lds_row = ((size_t)W.n_b * NST + (size_t)H.M * H.M) * sizeof(T);
  auto kc = sigma_dense_col_kernel<T, NI, NST, MAXS, BS>;
  auto kr = sigma_dense_row_kernel<T, NI, NST>;
  QPX_HIP(hipFuncSetAttribute((const void *)kc, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_col));
  QPX_HIP(hipFuncSetAttribute((const void *)kr, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_row));
  const int ncc = (W.max_col + BS - 1) / BS, ncr = (W.max_row + 1023) / 1024;
  if (c->ktiming) QPX_HIP(hipEventRecord(c->kev[0], st));
  const int gc = ((W.n_b * ncc + 7) / 8) * 8, gr = ((W.n_a * ncr + 7) / 8) * 8;  // multiples of the 8 XCDs
  hipLaunchKernelGGL(kc, dim3(gc), dim3(BS), lds_col, st, A, ncc, Tab<T>::V(H), Tab<T>::sval(W.ca),
                     Tab<T>::dval(W.ca), u, v, s);
  QPX_HIP(hipGetLastError());
  if (c->ktiming) QPX_HIP(hipEventRecord(c->kev[1], st));
  hipLaunchKernelGGL(kr, dim3(gr), dim3(1024), lds_row, st, A, ncr, Tab<T>::V(H), Tab<T>::sval(W.cb),
                     Tab<T>::dval(W.cb), u, v);
  QPX_HIP(hipGetLastError());
  if (c->ktiming) QPX_HIP(hipEventRecord(c->kev[2], st));
}

template <typename T, int NI, int NST, int BS>
static void launch_dense_maxs_bs(qpx_ctx *c, int maxs, const T *u, T *v, T *s, hipStream_t st) {
  if (maxs <= 16) launch_dense<T, NI, NST, 16, BS>(c, u, v, s, st);
  else if (maxs <= 32) launch_dense<T, NI, NST, 32, BS>(c, u, v, s, st);
  else if (maxs <= 48) launch_dense<T, NI, NST, 48, BS>(c, u, v, s, st);
  else if (maxs <= 64) launch_dense<T, NI, NST, 64, BS>(c, u, v, s, st);
  else launch_dense<T, NI, NST, 96, BS>(c, u, v, s, st);
}

// workgroup size of the column kernel: 1024 (16 waves, 128 VGPRs; the
// LDS of one column allows one workgroup per CU)
template <typename T, int NI, int NST>
static void launch_dense_maxs(qpx_ctx *c, int maxs, const T *u, T *v, T *s, hipStream_t st) {
  launch_dense_maxs_bs<T, NI, NST, 1024>(c, maxs, u, v, s, st);
}

template <typename T, int NI>
static void launch_dense_nst(qpx_ctx *c, int nst, int maxs, const T *u, T *v, T *s, hipStream_t st) {
  switch (nst) {
    case 1: launch_dense_maxs<T, NI, 1>(c, maxs, u, v, s, st); break;
    case 2: launch_dense_maxs<T, NI, 2>(c, maxs, u, v, s, st); break;
    case 3: launch_dense_maxs<T, NI, 3>(c, maxs, u, v, s, st); break;
    default: launch_dense_maxs<T, NI, 4>(c, maxs, u, v, s, st); break;
  }
}

template <typename T>
static void launch_generic(qpx_ctx *c, int nst, const T *u, T *v, T *s, hipStream_t st) {
  Wfn &W = c->W;
  GenArgs A;
  A.astr = W.astr.p; A.bstr = W.bstr.p;
  A.col_ptr = W.col_ptr.p; A.csc_a = W.csc_a.p; A.csc_b = W.csc_b.p;
  A.row_ptr = W.row_ptr.p; A.csr_b = W.csr_b.p; A.csr2csc = W.csr2csc.p;
  A.as_ptr = W.ca.s_ptr.p; A.as_idx = W.ca.s_idx.p; A.as_info = W.ca.s_info.p;
  A.bs_ptr = W.cb.s_ptr.p; A.bs_idx = W.cb.s_idx.p; A.bs_info = W.cb.s_info.p;
  A.cost_a = W.cost_a.p; A.cost_b = W.cost_b.p;
  A.hdiag = W.hdiag.p; A.sdiag = W.sdiag.p;
  A.n_det = W.n_det; A.M = c->H.M; A.npair = c->H.npair; A.nst = nst;
  const int wpb = 4;
  int64_t blocks = (W.n_det + wpb - 1) / wpb;
  QPX_REQUIRE(blocks < (1LL << 31), "too many determinants for one launch");
  const Ham &H = c->H;
  if (c->ktiming) QPX_HIP(hipEventRecord(c->kev[0], st));
  switch (W.nint) {
    case 1: hipLaunchKernelGGL((sigma_generic_kernel<T, 1>), dim3((unsigned)blocks), dim3(64 * wpb), 0, st, A, Tab<T>::V(H), Tab<T>::hcore(H), u, v, s); break;
    case 2: hipLaunchKernelGGL((sigma_generic_kernel<T, 2>), dim3((unsigned)blocks), dim3(64 * wpb), 0, st, A, Tab<T>::V(H), Tab<T>::hcore(H), u, v, s); break;
    case 3: hipLaunchKernelGGL((sigma_generic_kernel<T, 3>), dim3((unsigned)blocks), dim3(64 * wpb), 0, st, A, Tab<T>::V(H), Tab<T>::hcore(H), u, v, s); break;
    default: hipLaunchKernelGGL((sigma_generic_kernel<T, 4>), dim3((unsigned)blocks), dim3(64 * wpb), 0, st, A, Tab<T>::V(H), Tab<T>::hcore(H), u, v, s); break;
  }
  QPX_HIP(hipGetLastError());
  if (c->ktiming) {
    QPX_HIP(hipEventRecord(c->kev[1], st));
    QPX_HIP(hipEventRecord(c->kev[2], st));
  }
}

namespace qpx {

// dense-map path is usable when one column (row) of u plus the M x M slice
// fits in the LDS of one workgroup and columns (rows) hold <= 4096 entries
static bool dense_ok(qpx_ctx *c, int nst, size_t tsize) {
  const Wfn &W = c->W;
  const Ham &H = c->H;
  if (W.nint > 2 || nst > 4) return false;
  if (W.n_a > 65534 || H.M > 128) return false;
  ensure_conn(c);   // at most 65534 alpha strings here: cheap
  if (W.ca.max_singles > 96) return false;
  size_t need_c = ((size_t)(W.n_a + 1) * nst + (size_t)H.M * H.M + (size_t)1024 * nst) * tsize;
  size_t need_r = ((size_t)W.n_b * nst + (size_t)H.M * H.M) * tsize;
  size_t lim = (size_t)c->lds_per_block;
  return need_c <= lim && need_r <= lim;
}

void sigma_device(qpx_ctx *c, int prec, int nst, const void *u, void *v, void *s, hipStream_t st) {
  QPX_REQUIRE(c->W.valid, "no wavefunction set");
  QPX_REQUIRE(nst >= 1 && nst <= GMAXST, "n_st must be in 1..8");
  QPX_REQUIRE(prec == QPX_FP64 || prec == QPX_FP32, "prec must be QPX_FP64 or QPX_FP32");
  size_t ts = prec == QPX_FP64 ? sizeof(double) : sizeof(float);
  bool dense = dense_ok(c, nst, ts);
  if (c->force_path >= 2) dense = false;
  if (c->force_path == 1) QPX_REQUIRE(dense, "dense-map sigma path not applicable to this wavefunction");
  if (!dense && c->force_path != 2 && sparse_ok(c, nst)) {
    c->sigma_path = 3;
    sigma_sparse(c, prec, nst, u, v, s, st);
    return;
  }
  QPX_REQUIRE(c->force_path != 3, "sparse sigma path not applicable to this wavefunction");
  const Wfn &W = c->W;
  if (dense) {
    c->sigma_path = 1;
    const int maxs = W.ca.max_singles;
    if (prec == QPX_FP64) {
      if (W.nint == 1) launch_dense_nst<double, 1>(c, nst, maxs, (const double *)u, (double *)v, (double *)s, st);
      else launch_dense_nst<double, 2>(c, nst, maxs, (const double *)u, (double *)v, (double *)s, st);
    } else {
      if (W.nint == 1) launch_dense_nst<float, 1>(c, nst, maxs, (const float *)u, (float *)v, (float *)s, st);
      else launch_dense_nst<float, 2>(c, nst, maxs, (const float *)u, (float *)v, (float *)s, st);
    }
  } else {
    c->sigma_path = 2;
    ensure_conn(c);
    if (prec == QPX_FP64) launch_generic<double>(c, nst, (const double *)u, (double *)v, (double *)s, st);
    else launch_generic<float>(c, nst, (const float *)u, (float *)v, (float *)s, st);
  }
}

}  // namespace qpx

// ---------------------------------------------------------------------------
// host-facing entry points
// ---------------------------------------------------------------------------
template <typename T>
__global__ void gather_u_kernel(const double *__restrict__ u, int64_t sze, const int64_t *__restrict__ perm,
                                int64_t n, int nst, T *__restrict__ ut) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n * nst; x += (int64_t)gridDim.x * blockDim.x) {
    int64_t k = x / nst;
    int st = (int)(x - k * nst);
    ut[x] = (T)u[perm[k] + sze * st];
  }
}

template <typename T>
__global__ void scatter_v_kernel(const T *__restrict__ vt, const int64_t *__restrict__ perm, int64_t n, int nst,
                                 int64_t sze, double *__restrict__ v) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n * nst; x += (int64_t)gridDim.x * blockDim.x) {
    int64_t k = x / nst;
    int st = (int)(x - k * nst);
    v[perm[k] + sze * st] = (double)vt[x];
  }
}

template <typename T>
static void sigma_host(qpx_ctx *c, int prec, int nst, int64_t sze, const double *u, double *v, double *s) {
  Wfn &W = c->W;
  const int64_t n = W.n_det;
  hipStream_t st = c->stream;
  char *io = c->ws_sigma_io.get(sizeof(double) * (size_t)sze * nst * 3);
  double *du = (double *)io, *dv = du + (size_t)sze * nst, *ds = dv + (size_t)sze * nst;
  char *tb = c->ws_sigma_t.get(sizeof(T) * (size_t)n * nst * 3);
  T *ut = (T *)tb, *vt = ut + (size_t)n * nst, *stt = vt + (size_t)n * nst;
  QPX_HIP(hipMemcpyAsync(du, u, sizeof(double) * (size_t)sze * nst, hipMemcpyHostToDevice, st));
  int grid = (int)std::min<int64_t>((n * nst + 255) / 256, 65536);
  hipLaunchKernelGGL(gather_u_kernel<T>, dim3(grid), dim3(256), 0, st, du, sze, W.perm_d.p, n, nst, ut);
  QPX_HIP(hipGetLastError());
  sigma_device(c, prec, nst, ut, vt, stt, st);
  QPX_HIP(hipMemsetAsync(dv, 0, sizeof(double) * (size_t)sze * nst * 2, st));
  hipLaunchKernelGGL(scatter_v_kernel<T>, dim3(grid), dim3(256), 0, st, vt, W.perm_d.p, n, nst, sze, dv);
  hipLaunchKernelGGL(scatter_v_kernel<T>, dim3(grid), dim3(256), 0, st, stt, W.perm_d.p, n, nst, sze, ds);
  QPX_HIP(hipGetLastError());
  QPX_HIP(hipMemcpyAsync(v, dv, sizeof(double) * (size_t)sze * nst, hipMemcpyDeviceToHost, st));
  QPX_HIP(hipMemcpyAsync(s, ds, sizeof(double) * (size_t)sze * nst, hipMemcpyDeviceToHost, st));
  QPX_HIP(hipStreamSynchronize(st));
}

extern "C" int qpx_sigma(qpx_ctx *c, int prec, int nst, int64_t sze, const double *u, double *v, double *s) {
  return qpx_guard(c, [&] {
    QPX_HIP(hipSetDevice(c->device));
    QPX_REQUIRE(c->W.valid, "no wavefunction set");
    QPX_REQUIRE(sze >= c->W.n_det, "sze must be >= n_det");
    QPX_REQUIRE(u && v && s, "null vector");
    if (prec == QPX_FP64) sigma_host<double>(c, prec, nst, sze, u, v, s);
    else sigma_host<float>(c, prec, nst, sze, u, v, s);
  });
}

extern "C" int qpx_sigma_device(qpx_ctx *c, int prec, int nst, const void *u_t, void *v_t, void *s_t,
                                void *stream) {
  return qpx_guard(c, [&] {
    QPX_HIP(hipSetDevice(c->device));
    QPX_REQUIRE(u_t && v_t && s_t, "null device vector");
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    sigma_device(c, prec, nst, u_t, v_t, s_t, st);
  });
}

// LDS read bandwidth of the device: every lane of 1024-thread workgroups
// (two per CU, 64 KiB of LDS each) reads 16 bytes per iteration from LDS, in
// the conflict-free pattern (consecutive lanes, consecutive 16-byte words),
// and sums them so that no read can be dropped.
__global__ __launch_bounds__(1024) void lds_read_kernel(int iters, float *out) {
  __shared__ float4 buf[4096];
  for (int i = threadIdx.x; i < 4096; i += blockDim.x) buf[i] = make_float4(i, i + 1, i + 2, i + 3);
  __syncthreads();
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
  const int base = threadIdx.x;
  for (int it = 0; it < iters; ++it) {
    const float4 x = buf[(base + 64 * it) & 4095];
    a0 += x.x;
    a1 += x.y;
    a2 += x.z;
    a3 += x.w;
  }
  const float r = a0 + a1 + a2 + a3;
  if (r == -1.f) out[blockIdx.x] = r;   // never true: keeps the reads
}

extern "C" int qpx_debug_lds_bandwidth(qpx_ctx *c, double *bytes_per_s) {
  return qpx_guard(c, [&] {
    QPX_REQUIRE(bytes_per_s, "null output");
    QPX_HIP(hipSetDevice(c->device));
    int cus = 0;
    QPX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const int grid = 2 * cus, iters = 16384;
    DBuf<float> out;
    out.alloc(grid);
    hipEvent_t e0, e1;
    QPX_HIP(hipEventCreate(&e0));
    QPX_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(lds_read_kernel, dim3(grid), dim3(1024), 0, c->stream, iters, out.p);   // warm-up
    QPX_HIP(hipEventRecord(e0, c->stream));
    hipLaunchKernelGGL(lds_read_kernel, dim3(grid), dim3(1024), 0, c->stream, iters, out.p);
    QPX_HIP(hipEventRecord(e1, c->stream));
    QPX_HIP(hipGetLastError());
    QPX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    QPX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *bytes_per_s = (double)grid * 1024.0 * iters * 16.0 / (ms * 1e-3);
  });
}

// test hook: 0 = automatic, 1 = dense-map only, 2 = generic only
extern "C" int qpx_debug_force_sigma_path(qpx_ctx *c, int path, int *last_path) {
  return qpx_guard(c, [&] {
    QPX_REQUIRE(path >= 0 && path <= 3, "path must be 0 (automatic), 1 (dense map), 2 (generic scan) or 3 (sparse buckets)");
    c->force_path = path;
    if (last_path) *last_path = c->sigma_path;
  });
}

// benchmark hook: average duration (ms) of the sigma kernels over `reps`
// calls, measured with events on the launch stream: ms[0] = alpha/beta +
// same-spin alpha kernel (or the generic kernel), ms[1] = same-spin beta.
extern "C" int qpx_debug_sigma_kernel_times(qpx_ctx *c, int prec, int nst, const void *u_t, void *v_t, void *s_t,
                                            void *stream, float *ms, int reps) {
  return qpx_guard(c, [&] {
    QPX_HIP(hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    for (int i = 0; i < 3; ++i)
      if (!c->kev[i]) QPX_HIP(hipEventCreate(&c->kev[i]));
    double t0 = 0.0, t1 = 0.0;
    c->ktiming = true;
    try {
      for (int r = 0; r < reps; ++r) {
        sigma_device(c, prec, nst, u_t, v_t, s_t, st);
        QPX_HIP(hipEventSynchronize(c->kev[2]));
        float a = 0.f, b = 0.f;
        QPX_HIP(hipEventElapsedTime(&a, c->kev[0], c->kev[1]));
        QPX_HIP(hipEventElapsedTime(&b, c->kev[1], c->kev[2]));
        t0 += a;
        t1 += b;
      }
    } catch (...) {
      c->ktiming = false;
      throw;
    }
    c->ktiming = false;
    ms[0] = (float)(t0 / reps);
    ms[1] = (float)(t1 / reps);
  });
}
