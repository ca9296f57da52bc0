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

void build_wavefunction(qpx_ctx *c, int nint, int64_t n_det, const uint64_t *psi) {
  QPX_REQUIRE(c->ham_set, "qpx_set_hamiltonian must be called before qpx_set_wavefunction");
  QPX_REQUIRE(nint == c->H.nint, "n_int does not match mo_num");
  QPX_REQUIRE(n_det > 0 && n_det < (1LL << 31), "n_det out of range");
  QPX_REQUIRE(psi != nullptr, "null psi_det");
  Wfn &W = c->W;
  W.valid = false;
  W.sp.clear();
  W.n_det = n_det;
  W.nint = nint;
  hipStream_t s = c->stream;
  const int M = c->H.M;
  // electron counts
  for (int64_t k = 0; k < n_det; ++k) {
    int na = 0, nb = 0;
    for (int w = 0; w < nint; ++w) {
      na += __builtin_popcountll(psi[(2 * k) * nint + w]);
      nb += __builtin_popcountll(psi[(2 * k + 1) * nint + w]);
      if (w == nint - 1 && M % 64) {
        uint64_t hi = ~((1ULL << (M % 64)) - 1);
        QPX_REQUIRE(!(psi[(2 * k) * nint + w] & hi) && !(psi[(2 * k + 1) * nint + w] & hi),
                    "determinant occupies an orbital beyond mo_num");
      }
    }
    QPX_REQUIRE(na == c->H.na && nb == c->H.nb, "determinant electron count differs from n_alpha/n_beta");
  }
  // unique strings per spin (lexicographic on words, high word first)
  auto unique_spin = [&](int spin, std::vector<uint64_t> &out, std::vector<int> &idx_of_det) {
    std::vector<int64_t> ord(n_det);
    std::iota(ord.begin(), ord.end(), 0);
    auto str = [&](int64_t k) { return psi + (2 * k + spin) * nint; };
    std::sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) {
      const uint64_t *a = str(x), *b = str(y);
      for (int w = nint - 1; w >= 0; --w)
        if (a[w] != b[w]) return a[w] < b[w];
      return x < y;
    });
    out.clear();
    idx_of_det.assign(n_det, -1);
    int n = -1;
    const uint64_t *prev = nullptr;
    for (int64_t t = 0; t < n_det; ++t) {
      const uint64_t *a = str(ord[t]);
      if (!prev || std::memcmp(a, prev, sizeof(uint64_t) * nint) != 0) {
        ++n;
        out.insert(out.end(), a, a + nint);
        prev = a;
      }
      idx_of_det[ord[t]] = n;
    }
    return n + 1;
  };
  std::vector<int> aid, bid;
  W.n_a = unique_spin(0, W.astr_h, aid);
  W.n_b = unique_spin(1, W.bstr_h, bid);
  // CSC order: (beta, alpha); duplicates of a determinant are rejected
  std::vector<int64_t> key(n_det);
  W.perm.resize(n_det);
  std::iota(W.perm.begin(), W.perm.end(), 0);
  for (int64_t k = 0; k < n_det; ++k) key[k] = (int64_t)bid[k] * W.n_a + aid[k];
  std::sort(W.perm.begin(), W.perm.end(), [&](int64_t x, int64_t y) { return key[x] < key[y]; });
  std::vector<int> csc_a(n_det), csc_b(n_det), col_ptr(W.n_b + 1, 0);
  for (int64_t t = 0; t < n_det; ++t) {
    int64_t k = W.perm[t];
    if (t > 0) QPX_REQUIRE(key[k] != key[W.perm[t - 1]], "duplicate determinant in psi_det");
    csc_a[t] = aid[k];
    csc_b[t] = bid[k];
    col_ptr[bid[k] + 1]++;
  }
  for (int j = 0; j < W.n_b; ++j) col_ptr[j + 1] += col_ptr[j];
  // CSR order: (alpha, beta), pointing back to CSC positions
  std::vector<int64_t> ord2(n_det);
  std::iota(ord2.begin(), ord2.end(), 0);
  std::sort(ord2.begin(), ord2.end(), [&](int64_t x, int64_t y) {
    if (csc_a[x] != csc_a[y]) return csc_a[x] < csc_a[y];
    return csc_b[x] < csc_b[y];
  });
  std::vector<int> csr_b(n_det), csr2csc(n_det), row_ptr(W.n_a + 1, 0);
  for (int64_t t = 0; t < n_det; ++t) {
    csr_b[t] = csc_b[ord2[t]];
    csr2csc[t] = (int)ord2[t];
    row_ptr[csc_a[ord2[t]] + 1]++;
  }
  for (int i = 0; i < W.n_a; ++i) row_ptr[i + 1] += row_ptr[i];
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
  W.psi.upload(psi, (size_t)n_det * 2 * nint, s);
  W.astr.upload(W.astr_h, s);
  W.bstr.upload(W.bstr_h, s);
  W.col_ptr.upload(col_ptr, s);
  W.csc_a.upload(csc_a, s);
  W.csc_b.upload(csc_b, s);
  W.row_ptr.upload(row_ptr, s);
  W.csr_b.upload(csr_b, s);
  W.csr2csc.upload(csr2csc, s);
  W.perm_d.upload(W.perm, s);
  // the spin-string connectivity (all pairs of unique strings) is built on
  // first use: the sparse sigma path and the selection do not need it
  W.conn_valid = false;
  W.hdiag.alloc(n_det);
  W.sdiag.alloc(n_det);
  Ham &H = c->H;
  int grid = (int)std::min<int64_t>((n_det + 255) / 256, 65536);
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
    std::memcpy(perm, c->W.perm.data(), sizeof(int64_t) * c->W.n_det);
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
