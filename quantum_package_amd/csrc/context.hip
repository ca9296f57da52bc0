// context.hip -- context lifetime, errors and Hamiltonian upload.
//
// The Hamiltonian providers of the reference (mo_mono_elec_integral,
// mo_bielec_integral via the packed map, mo_bielec_integral_jj{,_anti},
// src/Integrals_Bielec/mo_bi_integrals.irp.f:1344-1369) become dense device
// tables: the packed (pq|rs) array is expanded once into the symmetric
// pair matrix V[P][R] (P, R = packed pair indices), so every kernel reads an
// integral with one load and no index folding.
#include <cstring>
#include <cstdlib>
#include <cmath>
#include "qpx_internal.h"

using namespace qpx;

__host__ __device__ static inline int64_t pair0(int p, int q) {  // 0-based orbitals
  int lo = p < q ? p : q, hi = p < q ? q : p;
  return (int64_t)hi * (hi + 1) / 2 + lo;
}

__global__ void expand_pair_matrix(const double *__restrict__ eri, double *__restrict__ V, float *__restrict__ Vf,
                                   int npair) {
  int64_t n = (int64_t)npair * npair;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    int64_t P = x / npair, R = x - P * npair;
    int64_t lo = P < R ? P : R, hi = P < R ? R : P;
    double v = eri[hi * (hi + 1) / 2 + lo];
    V[x] = v;
    Vf[x] = (float)v;
  }
}

// single-excitation tables: for the pair (h, p) and every orbital k,
// GS[(h*M+p)*M+k] = (hp|kk) - (hk|kp) (k of the excited spin) and
// GO[(h*M+p)*M+k] = (hp|kk) (k of the other spin), so that a single
// excitation element is h_hp + sum over the two strings' occupations
__global__ void single_tables_kernel(const double *__restrict__ V, int M, int npair, double *__restrict__ GS,
                                     double *__restrict__ GO, float *__restrict__ GSf, float *__restrict__ GOf) {
  const int64_t n = (int64_t)M * M * M;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(x % M);
    const int64_t hp = x / M;
    const int h = (int)(hp / M), p = (int)(hp % M);
    const double go = V[pair0(h, p) * npair + pair0(k, k)];
    const double gs = go - V[pair0(h, k) * npair + pair0(k, p)];
    GS[x] = gs;
    GO[x] = go;
    GSf[x] = (float)gs;
    GOf[x] = (float)go;
  }
}

// G4[((i*M + j)*M + y)*M + x] = (ix|jy), 0-based orbitals: every integral
// (ab|cd) sits in it with any one chosen orbital in the contiguous x slot,
// so a kernel whose lanes vary one orbital reads consecutive addresses
__global__ void g4_fill_kernel(const double *__restrict__ V, int M, int npair, double *__restrict__ G,
                               float *__restrict__ Gf) {
  const int64_t n = (int64_t)M * M * M * M;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(q % M);
    int64_t t = q / M;
    const int y = (int)(t % M);
    t /= M;
    const int j = (int)(t % M), i = (int)(t / M);
    const double v = V[pair0(i, x) * npair + pair0(j, y)];
    G[q] = v;
    Gf[q] = (float)v;
  }
}

namespace qpx {
bool ensure_g4(qpx_ctx *c) {
  Ham &H = c->H;
  if (H.M > kG4MaxOrb) return false;
  if (H.G4.p) return true;
  const int64_t n4 = (int64_t)H.M * H.M * H.M * H.M;
  H.G4.alloc(n4);
  H.G4f.alloc(n4);
  hipLaunchKernelGGL(g4_fill_kernel, dim3((unsigned)std::min<int64_t>((n4 + 255) / 256, 1 << 16)), dim3(256), 0,
                     c->stream, H.V.p, H.M, H.npair, H.G4.p, H.G4f.p);
  QPX_HIP(hipGetLastError());
  QPX_HIP(hipStreamSynchronize(c->stream));
  return true;
}
}  // namespace qpx

extern "C" int qpx_create(int device, qpx_ctx **out) {
  if (!out) return 1;
  *out = nullptr;
  qpx_ctx *c = new qpx_ctx();
  int rc = qpx_guard(c, [&] {
    int n = 0;
    QPX_HIP(hipGetDeviceCount(&n));
    QPX_REQUIRE(n > 0, "no HIP device visible");
    QPX_REQUIRE(device >= 0 && device < n, "device index out of range");
    c->device = device;
    QPX_HIP(hipSetDevice(device));
    QPX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    QPX_HIP(hipGetDeviceProperties(&prop, device));
    c->lds_per_block = (int)prop.sharedMemPerBlock;
    c->n_cu = prop.multiProcessorCount;
  });
  if (rc) {
    std::string m = c->err;
    delete c;
    (void)m;
    return rc;
  }
  *out = c;
  return 0;
}

extern "C" int qpx_destroy(qpx_ctx *c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
  for (int i = 0; i < 3; ++i)
    if (c->kev[i]) (void)hipEventDestroy(c->kev[i]);
  delete c;  // DBuf destructors free device memory
  return 0;
}

extern "C" int qpx_set_verbosity(qpx_ctx *c, int level) {
  return qpx_guard(c, [&] { c->verbosity = level; });
}

extern "C" int qpx_last_error(qpx_ctx *c, char *buf, int len) {
  if (!c || !buf || len <= 0) return 1;
  std::strncpy(buf, c->err.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return 0;
}

extern "C" int qpx_set_hamiltonian(qpx_ctx *c, int M, int na, int nb, const double *h_core, const double *eri,
                                   int64_t n8) {
  return qpx_guard(c, [&] {
    QPX_HIP(hipSetDevice(c->device));
    QPX_REQUIRE(M > 0 && M <= 64 * NINT_MAX, "mo_num out of range (1..256)");
    QPX_REQUIRE(na >= 0 && nb >= 0 && na <= M && nb <= M, "electron numbers out of range");
    QPX_REQUIRE(na >= nb, "elec_alpha_num must be >= elec_beta_num");
    QPX_REQUIRE(h_core && eri, "null Hamiltonian array");
    int64_t npair = (int64_t)M * (M + 1) / 2;
    QPX_REQUIRE(n8 == npair * (npair + 1) / 2, "eri_packed length must be npair*(npair+1)/2");
    Ham &H = c->H;
    H.M = M;
    H.na = na;
    H.nb = nb;
    H.nint = (M + 63) / 64;
    H.npair = (int)npair;
    H.h.assign(h_core, h_core + (size_t)M * M);
    H.eri.assign(eri, eri + n8);
    std::vector<float> hf(H.h.begin(), H.h.end());
    H.hcore.upload(H.h, c->stream);
    H.hcoref.upload(hf, c->stream);
    std::vector<int> pt((size_t)M * M);
    std::vector<double> J((size_t)M * M), K((size_t)M * M);
    for (int p = 0; p < M; ++p)
      for (int q = 0; q < M; ++q) {
        pt[(size_t)p * M + q] = (int)pair0(p, q);
        int64_t Pii = pair0(p, p), Pjj = pair0(q, q), Pij = pair0(p, q);
        auto pk = [&](int64_t a, int64_t b) {
          int64_t lo = a < b ? a : b, hi = a < b ? b : a;
          return eri[hi * (hi + 1) / 2 + lo];
        };
        J[(size_t)p * M + q] = pk(Pii, Pjj);
        K[(size_t)p * M + q] = pk(Pij, Pij);
      }
    H.ptab.upload(pt, c->stream);
    H.Jd.upload(J, c->stream);
    H.Kd.upload(K, c->stream);
    // jj_anti = jj - jj_exchange (mo_bi_integrals.irp.f:1363-1365)
    std::vector<double> JA((size_t)M * M);
    for (size_t x = 0; x < JA.size(); ++x) JA[x] = J[x] - K[x];
    H.JAd.upload(JA, c->stream);
    // closed-shell reference (mono_excitations.irp.f:2-19): orbitals 1..nb
    // doubly occupied.  fock(i,j) = h_ij + sum_k 2<ik|jk> - <ik|kj>
    // (mono_excitations.irp.f:22-95), summed in ascending k.
    auto g4 = [&](int i, int j, int k, int l) {  // <ij|kl>, 0-based orbitals
      int64_t P = pair0(i, k), R = pair0(j, l);
      int64_t lo = P < R ? P : R, hi = P < R ? R : P;
      return eri[hi * (hi + 1) / 2 + lo];
    };
    std::vector<double> F((size_t)M * M);
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < M; ++j) {
        double accu = 0.0;
        for (int k = 0; k < nb; ++k) accu += 2.0 * g4(i, k, j, k) - g4(i, k, k, j);
        F[(size_t)i + (size_t)M * j] = accu + h_core[(size_t)i + (size_t)M * j];
      }
    H.fock_cs.upload(F, c->stream);
    // ref_bitmask_energy (ref_bitmask.irp.f:1-57) in the reference's order
    {
      auto hh = [&](int i) { return h_core[(size_t)i * (M + 1)]; };
      double e = 0.0;
      for (int i = 0; i < nb; ++i) e += hh(i) + hh(i);
      for (int i = nb; i < na; ++i) e += hh(i);
      for (int j = 0; j < na; ++j)
        for (int i = j + 1; i < na; ++i) e += JA[(size_t)i + (size_t)M * j];
      for (int j = 0; j < nb; ++j) {
        for (int i = j + 1; i < nb; ++i) e += JA[(size_t)i + (size_t)M * j];
        for (int i = 0; i < na; ++i) e += J[(size_t)i + (size_t)M * j];
      }
      H.ref_energy = e;
    }
    DBuf<double> eri_d;
    eri_d.upload(eri, (size_t)n8, c->stream);
    H.V.alloc((size_t)npair * npair);
    H.Vf.alloc((size_t)npair * npair);
    int grid = (int)std::min<int64_t>((npair * npair + 255) / 256, 65536);
    hipLaunchKernelGGL(expand_pair_matrix, dim3(grid), dim3(256), 0, c->stream, eri_d.p, H.V.p, H.Vf.p,
                       (int)npair);
    QPX_HIP(hipGetLastError());
    const size_t m3 = (size_t)M * M * M;
    H.GS.alloc(m3);
    H.GO.alloc(m3);
    H.GSf.alloc(m3);
    H.GOf.alloc(m3);
    hipLaunchKernelGGL(single_tables_kernel, dim3((int)std::min<size_t>((m3 + 255) / 256, 65536)), dim3(256), 0,
                       c->stream, H.V.p, M, (int)npair, H.GS.p, H.GO.p, H.GSf.p, H.GOf.p);
    QPX_HIP(hipGetLastError());
    QPX_HIP(hipStreamSynchronize(c->stream));
    H.G4.release();   // rebuilt for the new integrals on the next sparse call
    H.G4f.release();
    c->ham_set = true;
    c->W.valid = false;
  });
}
