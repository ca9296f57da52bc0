// wavefunction.hip -- unique spin strings, alpha x beta bilinear structure,
// spin-string connectivity and the diagonal <I|H|I>, <I|S^2|I>.
//
// Reference: src/Determinants/spindeterminants.irp.f:56-136 (unique
// strings), :387-597 (psi_bilinear_matrix_* CSC by beta column and its
// transpose), :843-918 (singles_alpha_csc, declared but unused by the
// reference sigma: here every unique string gets its singles and doubles
// inside the unique set, with the excitation data precomputed), and
// src/Determinants/slater_rules.irp.f:1869-1929 / s2.irp.f:1-30 (diagonal).
//
// Layout: strings are (n, nint) uint64 rows; bilinear (CSC) order sorts
// determinants by (beta index, alpha index); CSR order by (alpha, beta).
#include <algorithm>
#include <numeric>
#include <cstring>
#include <hipcub/hipcub.hpp>
#include "qpx_internal.h"
#include "qpx_device.h"

using namespace qpx;

// ---------------------------------------------------------------------------
// connectivity of unique strings: all-pairs popcount, LDS-tiled
// ---------------------------------------------------------------------------
template <int NI>
__global__ __launch_bounds__(256) void conn_count_kernel(const uint64_t *__restrict__ str, int n,
                                                         int *__restrict__ ns, int *__restrict__ nd) {
  __shared__ uint64_t tile[1024 * NI];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t me[NI];
#pragma unroll
  for (int w = 0; w < NI; ++w) me[w] = (i < n) ? str[(int64_t)i * NI + w] : 0;
  int cs = 0, cd = 0;
  for (int t0 = 0; t0 < n; t0 += 1024) {
    int m = min(1024, n - t0);
    __syncthreads();
    for (int x = threadIdx.x; x < m * NI; x += blockDim.x) tile[x] = str[(int64_t)t0 * NI + x];
    __syncthreads();
    if (i < n) {
      for (int j = 0; j < m; ++j) {
        int d = 0;
#pragma unroll
        for (int w = 0; w < NI; ++w) d += __popcll(me[w] ^ tile[j * NI + w]);
        cs += (d == 2);
        cd += (d == 4);
      }
    }
  }
  if (i < n) {
    ns[i] = cs;
    nd[i] = cd;
  }
}

template <int NI>
__global__ __launch_bounds__(256) void conn_fill_kernel(const uint64_t *__restrict__ str, int n, int M,
                                                        const double *__restrict__ V, int npair,
                                                        const double *__restrict__ hcore,
                                                        const int *__restrict__ sptr, const int *__restrict__ dptr,
                                                        int *__restrict__ sidx, uint32_t *__restrict__ sinfo,
                                                        double *__restrict__ sval, float *__restrict__ svalf,
                                                        int *__restrict__ didx, double *__restrict__ dval,
                                                        float *__restrict__ dvalf) {
  __shared__ uint64_t tile[1024 * NI];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t me[NI];
#pragma unroll
  for (int w = 0; w < NI; ++w) me[w] = (i < n) ? str[(int64_t)i * NI + w] : 0;
  int os = (i < n) ? sptr[i] : 0, od = (i < n) ? dptr[i] : 0;
  for (int t0 = 0; t0 < n; t0 += 1024) {
    int m = min(1024, n - t0);
    __syncthreads();
    for (int x = threadIdx.x; x < m * NI; x += blockDim.x) tile[x] = str[(int64_t)t0 * NI + x];
    __syncthreads();
    if (i >= n) continue;
    for (int j = 0; j < m; ++j) {
      int d = 0;
#pragma unroll
      for (int w = 0; w < NI; ++w) d += __popcll(me[w] ^ tile[j * NI + w]);
      if (d == 2) {
        int h, p;
        single_hp<NI>(me, &tile[j * NI], &h, &p);
        int ph = between_parity<NI>(me, h, p);
        double a = hcore[(size_t)h * M + p];
        int64_t Php = pair_index(h, p);
#pragma unroll
        for (int w = 0; w < NI; ++w) {
          uint64_t x = me[w];
          while (x) {
            int k = w * 64 + __ffsll((unsigned long long)x) - 1;
            x &= x - 1;
            a += V[Php * npair + pair_index(k, k)] - V[pair_index(h, k) * npair + pair_index(k, p)];
          }
        }
        if (ph) a = -a;
        sidx[os] = t0 + j;
        sinfo[os] = (uint32_t)h | ((uint32_t)p << 8) | ((uint32_t)ph << 16);
        sval[os] = a;
        svalf[os] = (float)a;
        ++os;
      } else if (d == 4) {
        int h1, h2, p1, p2;
        double_hp<NI>(me, &tile[j * NI], &h1, &h2, &p1, &p2);
        int ph = double_parity<NI>(me, h1, h2, p1, p2);
        double v = V[pair_index(h1, p1) * npair + pair_index(h2, p2)] -
                   V[pair_index(h1, p2) * npair + pair_index(h2, p1)];
        if (ph) v = -v;
        didx[od] = t0 + j;
        dval[od] = v;
        dvalf[od] = (float)v;
        ++od;
      }
    }
  }
}

// diagonal elements in bilinear order
template <int NI>
__global__ void diag_kernel(const uint64_t *__restrict__ astr, const uint64_t *__restrict__ bstr,
                            const int *__restrict__ csc_a, const int *__restrict__ csc_b, int64_t n_det, int M,
                            const double *__restrict__ hcore, const double *__restrict__ J,
                            const double *__restrict__ K, double *__restrict__ hdiag, double *__restrict__ sdiag) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n_det; k += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t *a = astr + (int64_t)csc_a[k] * NI;
    const uint64_t *b = bstr + (int64_t)csc_b[k] * NI;
    uint64_t aw[NI], bw[NI];
#pragma unroll
    for (int w = 0; w < NI; ++w) { aw[w] = a[w]; bw[w] = b[w]; }
    hdiag[k] = diag_energy<NI>(aw, bw, M, hcore, J, K);
    int nup = 0;
#pragma unroll
    for (int w = 0; w < NI; ++w) nup += __popcll((aw[w] ^ bw[w]) & aw[w]);
    sdiag[k] = (double)nup;
  }
}

// per-string scan cost for the generic kernel: sum over the string's singles
// of the length of their row/column
__global__ void scan_cost_kernel(const int *__restrict__ sptr, const int *__restrict__ sidx,
                                 const int *__restrict__ other_ptr, int n, int64_t *__restrict__ cost) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int64_t c = 0;
    for (int e = sptr[i]; e < sptr[i + 1]; ++e) {
      int j = sidx[e];
      c += other_ptr[j + 1] - other_ptr[j];
    }
    cost[i] = c;
  }
}

// ---------------------------------------------------------------------------
static void build_conn(qpx_ctx *c, const DBuf<uint64_t> &str, int n, int nint, SpinConn &cc) {
  hipStream_t s = c->stream;
  DBuf<int> ns, nd;
  ns.alloc(n);
  nd.alloc(n);
  int grid = (n + 255) / 256;
  switch (nint) {
    case 1: hipLaunchKernelGGL(conn_count_kernel<1>, dim3(grid), dim3(256), 0, s, str.p, n, ns.p, nd.p); break;
    case 2: hipLaunchKernelGGL(conn_count_kernel<2>, dim3(grid), dim3(256), 0, s, str.p, n, ns.p, nd.p); break;
    case 3: hipLaunchKernelGGL(conn_count_kernel<3>, dim3(grid), dim3(256), 0, s, str.p, n, ns.p, nd.p); break;
    default: hipLaunchKernelGGL(conn_count_kernel<4>, dim3(grid), dim3(256), 0, s, str.p, n, ns.p, nd.p); break;
  }
  QPX_HIP(hipGetLastError());
  std::vector<int> hs(n), hd(n);
  QPX_HIP(hipMemcpyAsync(hs.data(), ns.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  QPX_HIP(hipMemcpyAsync(hd.data(), nd.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  QPX_HIP(hipStreamSynchronize(s));
  std::vector<int> ps(n + 1), pd(n + 1);
  ps[0] = pd[0] = 0;
  int64_t ts = 0, td = 0;
  cc.max_singles = cc.max_doubles = 0;
  for (int i = 0; i < n; ++i) {
    ts += hs[i];
    td += hd[i];
    QPX_REQUIRE(ts < (1LL << 31) && td < (1LL << 31), "spin-string connectivity exceeds 2^31 entries");
    ps[i + 1] = (int)ts;
    pd[i + 1] = (int)td;
    cc.max_singles = std::max(cc.max_singles, hs[i]);
    cc.max_doubles = std::max(cc.max_doubles, hd[i]);
  }
  cc.n_singles = ts;
  cc.n_doubles = td;
  cc.s_ptr.upload(ps, s);
  cc.d_ptr.upload(pd, s);
  cc.s_idx.alloc(std::max<int64_t>(ts, 1));
  cc.s_info.alloc(std::max<int64_t>(ts, 1));
  cc.s_val.alloc(std::max<int64_t>(ts, 1));
  cc.s_valf.alloc(std::max<int64_t>(ts, 1));
  cc.d_idx.alloc(std::max<int64_t>(td, 1));
  cc.d_val.alloc(std::max<int64_t>(td, 1));
  cc.d_valf.alloc(std::max<int64_t>(td, 1));
  Ham &H = c->H;
#define QPX_FILL(NI)                                                                                        \
  hipLaunchKernelGGL(conn_fill_kernel<NI>, dim3(grid), dim3(256), 0, s, str.p, n, H.M, H.V.p, H.npair,     \
                     H.hcore.p, cc.s_ptr.p, cc.d_ptr.p, cc.s_idx.p, cc.s_info.p, cc.s_val.p, cc.s_valf.p,  \
                     cc.d_idx.p, cc.d_val.p, cc.d_valf.p)
  switch (nint) {
    case 1: QPX_FILL(1); break;
    case 2: QPX_FILL(2); break;
    case 3: QPX_FILL(3); break;
    default: QPX_FILL(4); break;
  }
#undef QPX_FILL
  QPX_HIP(hipGetLastError());
}

namespace qpx {

// ---------------------------------------------------------------------------
// set-up of a wavefunction on the device: unique spin strings (stable sort of
// the determinants by their string, then run heads), the CSC order by (beta
// id, alpha id) and the CSR order by (alpha id, beta id), all with hipCUB
// radix sorts.  The unique strings come in the reference's order: by
// spin_det_search_key, the exclusive-or of the words compared as unsigned
// (src/Determinants/spindeterminants.irp.f:9-24,74-77), and among equal keys
// (possible for N_int > 1) by the words, first word first.
// ---------------------------------------------------------------------------
}  // namespace qpx

namespace {

// electron counts and orbital range of every determinant: bad[0] = 1 when
// one has the wrong number of alpha / beta electrons or an orbital beyond M
__global__ void check_dets_kernel(const uint64_t *__restrict__ psi, int64_t n, int nint, int M, int na, int nb,
                                  int *__restrict__ bad) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    int ca = 0, cb = 0;
    bool out = false;
    for (int w = 0; w < nint; ++w) {
      const uint64_t a = psi[(2 * k) * nint + w], b = psi[(2 * k + 1) * nint + w];
      ca += __popcll(a);
      cb += __popcll(b);
      if (w == nint - 1 && (M & 63)) {
        const uint64_t hi = ~((1ULL << (M & 63)) - 1);
        out = out || (a & hi) || (b & hi);
      }
    }
    if (ca != na || cb != nb || out) atomicOr(bad, 1);
  }
}

__global__ void iota_kernel(int *__restrict__ x, int64_t n) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    x[k] = (int)k;
}

// key[t] = word w of the spin string of determinant ord[t]
__global__ void spin_word_kernel(const uint64_t *__restrict__ psi, int64_t n, int nint, int spin, int w,
                                 const int *__restrict__ ord, uint64_t *__restrict__ key) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    key[t] = psi[(2 * (int64_t)ord[t] + spin) * nint + w];
}

// key[t] = spin_det_search_key of the spin string of determinant ord[t]: the
// exclusive-or of its words
__global__ void spin_xor_kernel(const uint64_t *__restrict__ psi, int64_t n, int nint, int spin,
                                const int *__restrict__ ord, uint64_t *__restrict__ key) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t *x = psi + (2 * (int64_t)ord[t] + spin) * nint;
    uint64_t k = x[0];
    for (int w = 1; w < nint; ++w) k ^= x[w];
    key[t] = k;
  }
}

// head[t] = 1 where the string of ord[t] differs from that of ord[t-1]
__global__ void run_head_kernel(const uint64_t *__restrict__ psi, int64_t n, int nint, int spin,
                                const int *__restrict__ ord, int *__restrict__ head) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    int h = 1;
    if (t > 0) {
      const uint64_t *a = psi + (2 * (int64_t)ord[t] + spin) * nint, *b = psi + (2 * (int64_t)ord[t - 1] + spin) * nint;
      h = 0;
      for (int w = 0; w < nint; ++w) h |= (a[w] != b[w]);
    }
    head[t] = h;
  }
}

// id_of_det[ord[t]] = sid[t] - 1; the string of each run head goes to
// ustr[sid[t] - 1]
__global__ void unique_scatter_kernel(const uint64_t *__restrict__ psi, int64_t n, int nint, int spin,
                                      const int *__restrict__ ord, const int *__restrict__ head,
                                      const int *__restrict__ sid, int *__restrict__ id_of_det,
                                      uint64_t *__restrict__ ustr) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int u = sid[t] - 1;
    id_of_det[ord[t]] = u;
    if (head[t])
      for (int w = 0; w < nint; ++w) ustr[(int64_t)u * nint + w] = psi[(2 * (int64_t)ord[t] + spin) * nint + w];
  }
}

__global__ void csc_key_kernel(const int *__restrict__ aid, const int *__restrict__ bid, int64_t n, int64_t n_a,
                               uint64_t *__restrict__ key) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    key[k] = (uint64_t)bid[k] * (uint64_t)n_a + (uint64_t)aid[k];
}

// bilinear (CSC) position t: its alpha / beta ids, the psi index, the
// duplicate test against position t-1 and the column / row counts
__global__ void csc_fill_kernel(const uint64_t *__restrict__ key, const int *__restrict__ perm, int64_t n,
                                const int *__restrict__ aid, const int *__restrict__ bid, int *__restrict__ csc_a,
                                int *__restrict__ csc_b, int64_t *__restrict__ perm64, int *__restrict__ col_cnt,
                                int *__restrict__ row_cnt, int *__restrict__ dup) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int k = perm[t];
    const int a = aid[k], b = bid[k];
    csc_a[t] = a;
    csc_b[t] = b;
    perm64[t] = k;
    atomicAdd(&col_cnt[b], 1);
    atomicAdd(&row_cnt[a], 1);
    if (t > 0 && key[t] == key[t - 1]) atomicOr(dup, 1);
  }
}

__global__ void csr_key_kernel(const int *__restrict__ csc_a, const int *__restrict__ csc_b, int64_t n, int64_t n_b,
                               uint64_t *__restrict__ key) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    key[t] = (uint64_t)csc_a[t] * (uint64_t)n_b + (uint64_t)csc_b[t];
}

__global__ void csr_fill_kernel(const int *__restrict__ ord2, const int *__restrict__ csc_b, int64_t n,
                                int *__restrict__ csr_b, int *__restrict__ csr2csc) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int q = ord2[t];
    csr_b[t] = csc_b[q];
    csr2csc[t] = q;
  }
}

int bits_for(uint64_t x) {   // bits needed to represent the values 0..x
  int b = 1;
  while (b < 64 && (x >> b) != 0) ++b;
  return b;
}

}  // namespace

namespace qpx {

void build_wavefunction(qpx_ctx *c, int nint, int64_t n_det, const uint64_t *psi) {
  QPX_REQUIRE(c->ham_set, "qpx_set_hamiltonian must be called before qpx_set_wavefunction");
  QPX_REQUIRE(nint == c->H.nint, "n_int does not match mo_num");
  QPX_REQUIRE(n_det > 0 && n_det < (1LL << 31), "n_det out of range");
  QPX_REQUIRE(psi != nullptr, "null psi_det");
  Wfn &W = c->W;
  W.valid = false;
  W.sp.clear();
  W.perm.clear();
  W.n_det = n_det;
  W.nint = nint;
  hipStream_t s = c->stream;
  const int M = c->H.M;
  const int64_t n = n_det;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
  W.psi.upload(psi, (size_t)n * 2 * nint, s);
  DBuf<int> flags;
  flags.alloc(2);
  QPX_HIP(hipMemsetAsync(flags.p, 0, sizeof(int) * 2, s));
  hipLaunchKernelGGL(check_dets_kernel, dim3(grid), dim3(256), 0, s, W.psi.p, n, nint, M, c->H.na, c->H.nb, flags.p);
  QPX_HIP(hipGetLastError());
  // sort buffers: keys and int values, double-buffered, plus hipCUB's
  // temporary storage sized for the largest call
  DBuf<uint64_t> k1, k2;
  DBuf<int> v1, v2, head, sid, aid, bid;
  k1.alloc(n); k2.alloc(n); v1.alloc(n); v2.alloc(n); head.alloc(n); sid.alloc(n); aid.alloc(n); bid.alloc(n);
  size_t tmp_sort = 0, tmp_scan = 0;
  QPX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sort, k1.p, k2.p, v1.p, v2.p, (int)n, 0, 64, s));
  QPX_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_scan, head.p, sid.p, (int)n, s));
  const size_t n_cnt = (size_t)n + 1;   // the column / row counts are at most n_det + 1 entries
  size_t tmp_scan2 = 0;
  {
    int *dummy = nullptr;
    QPX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_scan2, dummy, dummy, (int)n_cnt, s));
  }
  DBuf<unsigned char> temp;
  temp.alloc(std::max({tmp_sort, tmp_scan, tmp_scan2, (size_t)1}));
  auto unique_spin = [&](int spin, DBuf<uint64_t> &ustr, DBuf<int> &id_of_det) -> int {
    // stable LSD passes, least significant key first: the words (last word
    // first) when N_int > 1, then the search key (ties by psi index)
    hipLaunchKernelGGL(iota_kernel, dim3(grid), dim3(256), 0, s, v1.p, n);
    for (int w = (nint > 1 ? nint - 1 : -1); w >= -1; --w) {
      if (w >= 0)
        hipLaunchKernelGGL(spin_word_kernel, dim3(grid), dim3(256), 0, s, W.psi.p, n, nint, spin, w, v1.p, k1.p);
      else
        hipLaunchKernelGGL(spin_xor_kernel, dim3(grid), dim3(256), 0, s, W.psi.p, n, nint, spin, v1.p, k1.p);
      size_t tb = temp.n;
      QPX_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, tb, k1.p, k2.p, v1.p, v2.p, (int)n, 0, 64, s));
      std::swap(v1.p, v2.p);
    }
    hipLaunchKernelGGL(run_head_kernel, dim3(grid), dim3(256), 0, s, W.psi.p, n, nint, spin, v1.p, head.p);
    size_t tb = temp.n;
    QPX_HIP(hipcub::DeviceScan::InclusiveSum(temp.p, tb, head.p, sid.p, (int)n, s));
    int n_u = 0;
    QPX_HIP(hipMemcpyAsync(&n_u, sid.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    QPX_HIP(hipStreamSynchronize(s));
    ustr.alloc((size_t)n_u * nint);
    hipLaunchKernelGGL(unique_scatter_kernel, dim3(grid), dim3(256), 0, s, W.psi.p, n, nint, spin, v1.p, head.p,
                       sid.p, id_of_det.p, ustr.p);
    QPX_HIP(hipGetLastError());
    return n_u;
  };
  int bad[2] = {0, 0};
  QPX_HIP(hipMemcpyAsync(bad, flags.p, sizeof(int), hipMemcpyDeviceToHost, s));
  QPX_HIP(hipStreamSynchronize(s));
  QPX_REQUIRE(!bad[0], "determinant electron count differs from n_alpha/n_beta, or an orbital beyond mo_num");
  W.n_a = unique_spin(0, W.astr, aid);
  W.n_b = unique_spin(1, W.bstr, bid);
  // CSC order: by (beta id, alpha id); duplicates of a determinant are rejected
  hipLaunchKernelGGL(csc_key_kernel, dim3(grid), dim3(256), 0, s, aid.p, bid.p, n, (int64_t)W.n_a, k1.p);
  hipLaunchKernelGGL(iota_kernel, dim3(grid), dim3(256), 0, s, v1.p, n);
  {
    size_t tb = temp.n;
    const int nbits = bits_for((uint64_t)W.n_a * (uint64_t)W.n_b);
    QPX_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, tb, k1.p, k2.p, v1.p, v2.p, (int)n, 0, nbits, s));
  }
  W.csc_a.alloc(n);
  W.csc_b.alloc(n);
  W.perm_d.alloc(n);
  DBuf<int> col_cnt, row_cnt;
  col_cnt.alloc((size_t)W.n_b + 1);
  row_cnt.alloc((size_t)W.n_a + 1);
  QPX_HIP(hipMemsetAsync(col_cnt.p, 0, sizeof(int) * ((size_t)W.n_b + 1), s));
  QPX_HIP(hipMemsetAsync(row_cnt.p, 0, sizeof(int) * ((size_t)W.n_a + 1), s));
  hipLaunchKernelGGL(csc_fill_kernel, dim3(grid), dim3(256), 0, s, k2.p, v2.p, n, aid.p, bid.p, W.csc_a.p, W.csc_b.p,
                     W.perm_d.p, col_cnt.p, row_cnt.p, flags.p + 1);
  QPX_HIP(hipGetLastError());
  W.col_ptr.alloc((size_t)W.n_b + 1);
  W.row_ptr.alloc((size_t)W.n_a + 1);
  {
    size_t tb = temp.n;
    QPX_HIP(hipcub::DeviceScan::ExclusiveSum(temp.p, tb, col_cnt.p, W.col_ptr.p, W.n_b + 1, s));
    tb = temp.n;
    QPX_HIP(hipcub::DeviceScan::ExclusiveSum(temp.p, tb, row_cnt.p, W.row_ptr.p, W.n_a + 1, s));
  }
  // CSR order: by (alpha id, beta id), pointing back to CSC positions
  hipLaunchKernelGGL(csr_key_kernel, dim3(grid), dim3(256), 0, s, W.csc_a.p, W.csc_b.p, n, (int64_t)W.n_b, k1.p);
  hipLaunchKernelGGL(iota_kernel, dim3(grid), dim3(256), 0, s, v1.p, n);
  {
    size_t tb = temp.n;
    const int nbits = bits_for((uint64_t)W.n_a * (uint64_t)W.n_b);
    QPX_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, tb, k1.p, k2.p, v1.p, v2.p, (int)n, 0, nbits, s));
  }
  W.csr_b.alloc(n);
  W.csr2csc.alloc(n);
  hipLaunchKernelGGL(csr_fill_kernel, dim3(grid), dim3(256), 0, s, v2.p, W.csc_b.p, n, W.csr_b.p, W.csr2csc.p);
  QPX_HIP(hipGetLastError());
  // host copies: the unique strings (sub-string tables of the sparse path)
  // and the column / row lengths
  std::vector<int> col_ptr(W.n_b + 1), row_ptr(W.n_a + 1);
  W.astr_h.resize((size_t)W.n_a * nint);
  W.bstr_h.resize((size_t)W.n_b * nint);
  QPX_HIP(hipMemcpyAsync(W.astr_h.data(), W.astr.p, sizeof(uint64_t) * W.astr_h.size(), hipMemcpyDeviceToHost, s));
  QPX_HIP(hipMemcpyAsync(W.bstr_h.data(), W.bstr.p, sizeof(uint64_t) * W.bstr_h.size(), hipMemcpyDeviceToHost, s));
  QPX_HIP(hipMemcpyAsync(col_ptr.data(), W.col_ptr.p, sizeof(int) * col_ptr.size(), hipMemcpyDeviceToHost, s));
  QPX_HIP(hipMemcpyAsync(row_ptr.data(), W.row_ptr.p, sizeof(int) * row_ptr.size(), hipMemcpyDeviceToHost, s));
  QPX_HIP(hipMemcpyAsync(bad + 1, flags.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  QPX_HIP(hipStreamSynchronize(s));
  QPX_REQUIRE(!bad[1], "duplicate determinant in psi_det");
  W.col_len_h.resize(W.n_b);
  W.row_len_h.resize(W.n_a);
  W.max_col = W.max_row = 0;
  for (int j = 0; j < W.n_b; ++j) {
    W.col_len_h[j] = col_ptr[j + 1] - col_ptr[j];
    W.max_col = std::max(W.max_col, W.col_len_h[j]);
  }
  for (int i = 0; i < W.n_a; ++i) {
    W.row_len_h[i] = row_ptr[i + 1] - row_ptr[i];
    W.max_row = std::max(W.max_row, W.row_len_h[i]);
  }
  // the spin-string connectivity (all pairs of unique strings) is built on
  // first use: the sparse sigma path and the selection do not need it
  W.conn_valid = false;
  W.hdiag.alloc(n_det);
  W.sdiag.alloc(n_det);
  Ham &H = c->H;
#define QPX_DIAG(NI)                                                                                         \
  hipLaunchKernelGGL(diag_kernel<NI>, dim3(grid), dim3(256), 0, s, W.astr.p, W.bstr.p, W.csc_a.p, W.csc_b.p, \
                     n_det, M, H.hcore.p, H.Jd.p, H.Kd.p, W.hdiag.p, W.sdiag.p)
  switch (nint) {
    case 1: QPX_DIAG(1); break;
    case 2: QPX_DIAG(2); break;
    case 3: QPX_DIAG(3); break;
    default: QPX_DIAG(4); break;
  }
#undef QPX_DIAG
  QPX_HIP(hipGetLastError());
  QPX_HIP(hipStreamSynchronize(s));
  W.valid = true;
}

void ensure_conn(qpx_ctx *c) {
  Wfn &W = c->W;
  QPX_REQUIRE(W.valid, "no wavefunction set");
  if (W.conn_valid) return;
  hipStream_t s = c->stream;
  build_conn(c, W.astr, W.n_a, W.nint, W.ca);
  build_conn(c, W.bstr, W.n_b, W.nint, W.cb);
  // scan costs: alpha string a -> sum over its alpha singles a' of row length
  W.cost_a.alloc(W.n_a);
  W.cost_b.alloc(W.n_b);
  hipLaunchKernelGGL(scan_cost_kernel, dim3((W.n_a + 255) / 256), dim3(256), 0, s, W.ca.s_ptr.p, W.ca.s_idx.p,
                     W.row_ptr.p, W.n_a, W.cost_a.p);
  hipLaunchKernelGGL(scan_cost_kernel, dim3((W.n_b + 255) / 256), dim3(256), 0, s, W.cb.s_ptr.p, W.cb.s_idx.p,
                     W.col_ptr.p, W.n_b, W.cost_b.p);
  QPX_HIP(hipGetLastError());
  QPX_HIP(hipStreamSynchronize(s));
  W.conn_valid = true;
}

}  // namespace qpx

extern "C" int qpx_set_wavefunction(qpx_ctx *c, int n_int, int64_t n_det, const uint64_t *psi_det) {
  return qpx_guard(c, [&] {
    QPX_HIP(hipSetDevice(c->device));
    build_wavefunction(c, n_int, n_det, psi_det);
  });
}

extern "C" int qpx_get_bilinear_order(qpx_ctx *c, int64_t *perm) {
  return qpx_guard(c, [&] {
    QPX_REQUIRE(c->W.valid, "no wavefunction set");
    QPX_REQUIRE(perm, "null output");
    QPX_HIP(hipSetDevice(c->device));
    QPX_HIP(hipMemcpyAsync(perm, c->W.perm_d.p, sizeof(int64_t) * c->W.n_det, hipMemcpyDeviceToHost, c->stream));
    QPX_HIP(hipStreamSynchronize(c->stream));
  });
}

__global__ void scatter_diag(const double *__restrict__ hd, const int64_t *__restrict__ perm, int64_t n,
                             double *__restrict__ out) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    out[perm[k]] = hd[k];
}

extern "C" int qpx_diag_h(qpx_ctx *c, double *h_jj) {
  return qpx_guard(c, [&] {
    QPX_HIP(hipSetDevice(c->device));
    QPX_REQUIRE(c->W.valid, "no wavefunction set");
    int64_t n = c->W.n_det;
    DBuf<double> out;
    out.alloc(n);
    hipLaunchKernelGGL(scatter_diag, dim3((int)std::min<int64_t>((n + 255) / 256, 65536)), dim3(256), 0, c->stream,
                       c->W.hdiag.p, c->W.perm_d.p, n, out.p);
    QPX_HIP(hipGetLastError());
    QPX_HIP(hipMemcpyAsync(h_jj, out.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    QPX_HIP(hipStreamSynchronize(c->stream));
  });
}
