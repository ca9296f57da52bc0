// s2_complete.hip -- make_s2_eigenfunction on the GPU.
//
// Reference: src/Determinants/occ_pattern.irp.f:251-309.  Every determinant
// that shares an occupation pattern (doubly occupied orbitals, singly
// occupied orbitals) with a determinant of psi and is not in psi is added,
// so that psi spans whole spin multiplets.  The output order is the one of
// the host routine qpx_make_s2_eigenfunction (psi_update.hip), so both give
// the same array: patterns grouped by (open shells, alpha electrons among
// them) ascending; within a group the alpha subsets of the open shells in
// lexicographic order and, for each subset, the patterns in word order.
//
// Device steps: patterns of psi; psi and the patterns sorted in word order
// (stable LSD radix passes over the words, hipCUB); unique patterns; a
// stable sort of them by group; then one thread per pattern of a group walks
// the group's alpha subsets in lexicographic order, builds each determinant
// and looks it up in the sorted psi (binary search).  A first pass flags the
// missing determinants, an exclusive scan of the flags gives their output
// positions, and a second pass writes them there.
#include <algorithm>
#include <hipcub/hipcub.hpp>
#include <vector>
#include "qpx_internal.h"

using namespace qpx;

namespace {

__global__ void pattern_kernel(const uint64_t *__restrict__ psi, int64_t n, int nint, uint64_t *__restrict__ pat) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t *d = psi + i * 2 * nint;
    uint64_t *p = pat + i * 2 * nint;
    for (int w = 0; w < nint; ++w) {
      p[w] = d[w] & d[nint + w];          // doubly occupied
      p[nint + w] = d[w] ^ d[nint + w];   // singly occupied
    }
  }
}

__global__ void iota_kernel(int *__restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = (int)i;
}

// key[i] = word w of row perm[i]
__global__ void word_key_kernel(const uint64_t *__restrict__ rows, int64_t n, int nw, int w,
                                const int *__restrict__ perm, uint64_t *__restrict__ key) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    key[i] = rows[(int64_t)perm[i] * nw + w];
}

__global__ void gather_rows_kernel(const uint64_t *__restrict__ rows, int64_t n, int nw, const int *__restrict__ perm,
                                   uint64_t *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    for (int w = 0; w < nw; ++w) out[i * nw + w] = rows[(int64_t)perm[i] * nw + w];
}

// 1 where sorted row i differs from row i - 1
__global__ void head_kernel(const uint64_t *__restrict__ rows, int64_t n, int nw, int *__restrict__ head) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int h = (i == 0);
    for (int w = 0; w < nw && !h; ++w) h = rows[i * nw + w] != rows[(i - 1) * nw + w];
    head[i] = h;
  }
}

__global__ void unique_kernel(const uint64_t *__restrict__ rows, int64_t n, int nw, const int *__restrict__ head,
                              const int *__restrict__ pos, uint64_t *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (head[i])
      for (int w = 0; w < nw; ++w) out[(int64_t)pos[i] * nw + w] = rows[i * nw + w];
}

// group of a pattern: open shells ns and alpha electrons among them
// k = n_alpha - doubly occupied, as ns << 11 | (k + 1024)
__global__ void group_key_kernel(const uint64_t *__restrict__ pat, int64_t n, int nint, int n_alpha,
                                 uint32_t *__restrict__ key) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int nd = 0, ns = 0;
    for (int w = 0; w < nint; ++w) {
      nd += __popcll(pat[i * 2 * nint + w]);
      ns += __popcll(pat[i * 2 * nint + nint + w]);
    }
    key[i] = ((uint32_t)ns << 11) | (uint32_t)(n_alpha - nd + 1024);
  }
}

// word-order comparison of a determinant with row r of the sorted psi
__device__ inline int cmp_row(const uint64_t *d, const uint64_t *row, int nw) {
  for (int w = 0; w < nw; ++w)
    if (d[w] != row[w]) return d[w] < row[w] ? -1 : 1;
  return 0;
}

// One thread per pattern q of a group (ns open shells, k of them alpha):
// the alpha subsets in lexicographic order, r = 0 .. ncomb - 1.  Subset
// {c_0 < c_1 < ...} is the mask m with bit ns-1-c_j set: lexicographic order
// is decreasing m, i.e. increasing ~m, whose (ns - k)-bit masks Gosper's
// step enumerates in increasing order.  Item (r, q) sits at base + r * cnt
// + q.  pass 0 writes flag = missing from psi; pass 1 writes the missing
// determinants at their scanned positions.
template <int NI>
__global__ __launch_bounds__(256) void enum_kernel(int pass, const uint64_t *__restrict__ gpat, int64_t cnt, int ns,
                                                   int k, int64_t ncomb, const uint64_t *__restrict__ sdet,
                                                   int64_t n_det, int64_t base, int *__restrict__ flag,
                                                   const int *__restrict__ pos, uint64_t *__restrict__ out) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= cnt) return;
  constexpr int NW = 2 * NI;
  uint64_t docc[NI], socc[NI];
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    docc[w] = gpat[q * NW + w];
    socc[w] = gpat[q * NW + NI + w];
  }
  const int nk = ns - k;
  uint64_t mc = (nk == 0) ? 0ULL : ((nk == 64) ? ~0ULL : ((1ULL << nk) - 1));   // ~m, nk bits
  const uint64_t full = (ns == 64) ? ~0ULL : ((1ULL << ns) - 1);
  for (int64_t r = 0; r < ncomb; ++r) {
    const uint64_t m = ~mc & full;
    uint64_t d[NW];
#pragma unroll
    for (int w = 0; w < NI; ++w) d[w] = d[NI + w] = docc[w];
    int i = 0;   // open shell number, ascending orbitals
#pragma unroll
    for (int w = 0; w < NI; ++w) {
      uint64_t y = socc[w];
      uint64_t a = 0, b = 0;
      while (y) {
        const uint64_t bitv = y & (~y + 1);
        y ^= bitv;
        if ((m >> (ns - 1 - i)) & 1ULL) a |= bitv;
        else b |= bitv;
        ++i;
      }
      d[w] |= a;
      d[NI + w] |= b;
    }
    // binary search in the sorted psi
    int64_t lo = 0, hi = n_det;
    bool found = false;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      const int c = cmp_row(d, sdet + mid * NW, NW);
      if (c == 0) {
        found = true;
        break;
      }
      if (c > 0) lo = mid + 1;
      else hi = mid;
    }
    const int64_t item = base + r * cnt + q;
    if (pass == 0) {
      flag[item] = found ? 0 : 1;
    } else if (!found) {
      const int64_t p = pos[item];
#pragma unroll
      for (int w = 0; w < NW; ++w) out[p * NW + w] = d[w];
    }
    // next (ns - k)-bit mask in increasing order (Gosper)
    if (mc != 0 && r + 1 < ncomb) {
      const uint64_t c = mc & (~mc + 1);
      const uint64_t rr = mc + c;
      mc = (((rr ^ mc) >> 2) >> (__ffsll((unsigned long long)mc) - 1)) | rr;   // / c, c a power of two
    }
  }
}

// binomial coefficient, or -1 above 2^40
int64_t binom_capped(int n, int k) {
  if (k < 0 || k > n) return 0;
  k = std::min(k, n - k);
  double v = 1.0;
  int64_t r = 1;
  for (int j = 1; j <= k; ++j) {
    v = v * (n - k + j) / j;
    if (v > 1.1e12) return -1;
    r = r * (n - k + j) / j;   // exact: r * (n-k+j) is divisible by j at each step
  }
  return r;
}

// stable LSD sort of n rows of nw words in word order (word 0 most
// significant): perm receives the sorted order of the row indices
void sort_rows(const uint64_t *rows, int64_t n, int nw, DBuf<int> &perm, DBuf<int> &perm2, DBuf<uint64_t> &k1,
               DBuf<uint64_t> &k2, DBuf<unsigned char> &temp, hipStream_t s) {
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(iota_kernel, dim3(grid), dim3(256), 0, s, perm.p, n);
  for (int w = nw - 1; w >= 0; --w) {
    hipLaunchKernelGGL(word_key_kernel, dim3(grid), dim3(256), 0, s, rows, n, nw, w, perm.p, k1.p);
    size_t tb = temp.n;
    QPX_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, tb, k1.p, k2.p, perm.p, perm2.p, (int)n, 0, 64, s));
    std::swap(perm.p, perm2.p);
  }
  QPX_HIP(hipGetLastError());
}

}  // namespace

extern "C" int qpx_s2_complete(qpx_ctx *c, int nint, int64_t ndet, const uint64_t *psi_det, int n_alpha,
                               int64_t max_out, uint64_t *out_dets, int64_t *n_out) {
  int rc = 0;
  const int g = qpx_guard(c, [&] {
    QPX_REQUIRE(n_out != nullptr, "null n_out");
    QPX_REQUIRE(nint >= 1 && nint <= 4, "n_int out of range (1..4)");
    QPX_REQUIRE(ndet >= 0 && ndet < (1LL << 31), "n_det out of range");
    QPX_REQUIRE(ndet == 0 || psi_det != nullptr, "null psi_det");
    QPX_REQUIRE(n_alpha >= 0 && n_alpha <= 64 * nint, "n_alpha out of range");
    *n_out = 0;
    if (ndet == 0) return;
    QPX_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int nw = 2 * nint;
    const int64_t n = ndet;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
    DBuf<uint64_t> psi, pat, sdet, spat, k1, k2;
    DBuf<int> perm, perm2, head, hpos;
    psi.upload(psi_det, (size_t)n * nw, s);
    pat.alloc((size_t)n * nw);
    sdet.alloc((size_t)n * nw);
    spat.alloc((size_t)n * nw);
    k1.alloc(n); k2.alloc(n); perm.alloc(n); perm2.alloc(n); head.alloc(n); hpos.alloc(n);
    size_t tmp_sort = 0, tmp_scan = 0;
    QPX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sort, k1.p, k2.p, perm.p, perm2.p, (int)n, 0, 64, s));
    QPX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_scan, head.p, hpos.p, (int)n, s));
    DBuf<unsigned char> temp;
    temp.alloc(std::max({tmp_sort, tmp_scan, (size_t)1}));
    hipLaunchKernelGGL(pattern_kernel, dim3(grid), dim3(256), 0, s, psi.p, n, nint, pat.p);
    QPX_HIP(hipGetLastError());
    // psi in word order, for the membership test
    sort_rows(psi.p, n, nw, perm, perm2, k1, k2, temp, s);
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid), dim3(256), 0, s, psi.p, n, nw, perm.p, sdet.p);
    // unique patterns in word order
    sort_rows(pat.p, n, nw, perm, perm2, k1, k2, temp, s);
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid), dim3(256), 0, s, pat.p, n, nw, perm.p, spat.p);
    hipLaunchKernelGGL(head_kernel, dim3(grid), dim3(256), 0, s, spat.p, n, nw, head.p);
    {
      size_t tb = temp.n;
      QPX_HIP(hipcub::DeviceScan::ExclusiveSum(temp.p, tb, head.p, hpos.p, (int)n, s));
    }
    int last_pos = 0, last_head = 0;
    QPX_HIP(hipMemcpyAsync(&last_pos, hpos.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    QPX_HIP(hipMemcpyAsync(&last_head, head.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    QPX_HIP(hipStreamSynchronize(s));
    const int64_t nu = (int64_t)last_pos + last_head;
    DBuf<uint64_t> upat, gpat;
    upat.alloc((size_t)nu * nw);
    gpat.alloc((size_t)nu * nw);
    hipLaunchKernelGGL(unique_kernel, dim3(grid), dim3(256), 0, s, spat.p, n, nw, head.p, hpos.p, upat.p);
    QPX_HIP(hipGetLastError());
    // stable sort of the unique patterns by group
    const int gridu = (int)std::min<int64_t>((nu + 255) / 256, 65536);
    DBuf<uint32_t> gk1, gk2;
    gk1.alloc(nu);
    gk2.alloc(nu);
    hipLaunchKernelGGL(group_key_kernel, dim3(gridu), dim3(256), 0, s, upat.p, nu, nint, n_alpha, gk1.p);
    hipLaunchKernelGGL(iota_kernel, dim3(gridu), dim3(256), 0, s, perm.p, nu);
    {
      size_t tb = 0;
      QPX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, gk1.p, gk2.p, perm.p, perm2.p, (int)nu, 0, 32, s));
      DBuf<unsigned char> t2;
      t2.alloc(std::max(tb, (size_t)1));
      QPX_HIP(hipcub::DeviceRadixSort::SortPairs(t2.p, tb, gk1.p, gk2.p, perm.p, perm2.p, (int)nu, 0, 32, s));
    }
    hipLaunchKernelGGL(gather_rows_kernel, dim3(gridu), dim3(256), 0, s, upat.p, nu, nw, perm2.p, gpat.p);
    QPX_HIP(hipGetLastError());
    std::vector<uint32_t> gkey(nu);
    QPX_HIP(hipMemcpyAsync(gkey.data(), gk2.p, sizeof(uint32_t) * nu, hipMemcpyDeviceToHost, s));
    QPX_HIP(hipStreamSynchronize(s));
    // groups: [g0, g1) with the same key; items ncomb * count each
    struct Group {
      int64_t g0, cnt, ncomb, base;
      int ns, k;
    };
    std::vector<Group> groups;
    int64_t items = 0;
    for (int64_t g0 = 0; g0 < nu;) {
      int64_t g1 = g0;
      while (g1 < nu && gkey[g1] == gkey[g0]) ++g1;
      const int ns = (int)(gkey[g0] >> 11), k = (int)(gkey[g0] & 2047u) - 1024;
      if (ns > 0 && k >= 0 && k <= ns) {
        QPX_REQUIRE(ns <= 64, "more than 64 open shells in an occupation pattern");
        const int64_t nc = binom_capped(ns, k);
        QPX_REQUIRE(nc > 0, "too many alpha subsets of the open shells of a pattern");
        groups.push_back({g0, g1 - g0, nc, items, ns, k});
        items += nc * (g1 - g0);
        QPX_REQUIRE(items < (1LL << 31), "too many candidate determinants (2^31)");
      }
      g0 = g1;
    }
    if (items == 0) return;
    DBuf<int> flag, fpos;
    flag.alloc(items);
    fpos.alloc(items);
    auto launch = [&](int pass, uint64_t *outp) {
      for (const Group &G : groups) {
        const int gb = (int)((G.cnt + 255) / 256);
        const uint64_t *gp = gpat.p + G.g0 * nw;
        switch (nint) {
          case 1:
            hipLaunchKernelGGL(enum_kernel<1>, dim3(gb), dim3(256), 0, s, pass, gp, G.cnt, G.ns, G.k, G.ncomb,
                               sdet.p, n, G.base, flag.p, fpos.p, outp);
            break;
          case 2:
            hipLaunchKernelGGL(enum_kernel<2>, dim3(gb), dim3(256), 0, s, pass, gp, G.cnt, G.ns, G.k, G.ncomb,
                               sdet.p, n, G.base, flag.p, fpos.p, outp);
            break;
          case 3:
            hipLaunchKernelGGL(enum_kernel<3>, dim3(gb), dim3(256), 0, s, pass, gp, G.cnt, G.ns, G.k, G.ncomb,
                               sdet.p, n, G.base, flag.p, fpos.p, outp);
            break;
          default:
            hipLaunchKernelGGL(enum_kernel<4>, dim3(gb), dim3(256), 0, s, pass, gp, G.cnt, G.ns, G.k, G.ncomb,
                               sdet.p, n, G.base, flag.p, fpos.p, outp);
            break;
        }
      }
      QPX_HIP(hipGetLastError());
    };
    launch(0, nullptr);
    {
      size_t tb = 0;
      QPX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag.p, fpos.p, (int)items, s));
      DBuf<unsigned char> t3;
      t3.alloc(std::max(tb, (size_t)1));
      QPX_HIP(hipcub::DeviceScan::ExclusiveSum(t3.p, tb, flag.p, fpos.p, (int)items, s));
    }
    int lp = 0, lf = 0;
    QPX_HIP(hipMemcpyAsync(&lp, fpos.p + (items - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    QPX_HIP(hipMemcpyAsync(&lf, flag.p + (items - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    QPX_HIP(hipStreamSynchronize(s));
    const int64_t nadd = (int64_t)lp + lf;
    *n_out = nadd;
    if (!out_dets || nadd == 0) return;
    if (max_out < nadd) {
      rc = 4;
      c->err = "max_out is smaller than the number of determinants to add (*n_out)";
      return;
    }
    DBuf<uint64_t> out;
    out.alloc((size_t)nadd * nw);
    launch(1, out.p);
    QPX_HIP(hipMemcpyAsync(out_dets, out.p, sizeof(uint64_t) * nadd * nw, hipMemcpyDeviceToHost, s));
    QPX_HIP(hipStreamSynchronize(s));
  });
  return g ? g : rc;
}
