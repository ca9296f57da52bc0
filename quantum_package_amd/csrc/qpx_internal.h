// qpx_internal.h -- shared state of libqpx (MI355X / gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include <stdexcept>
#include "../../include/qpx.h"

namespace qpx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define QPX_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      throw ::qpx::Error(2, std::string("HIP error ") + hipGetErrorString(e_) + " at " +       \
                                __FILE__ + ":" + std::to_string(__LINE__) + " (" #call ")");   \
  } while (0)

#define QPX_REQUIRE(cond, msg)                                                                 \
  do {                                                                                         \
    if (!(cond)) throw ::qpx::Error(1, std::string(msg));                                      \
  } while (0)

constexpr int NINT_MAX = 4;
// largest orbital count for which the M^4 integral table (Ham::G4) is built:
// 8 + 4 bytes per entry, 7.9 GB at 160 orbitals; above it the kernels read
// the pair matrix
constexpr int kG4MaxOrb = 160;

// Device buffer with RAII.
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    QPX_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void upload(const T *h, size_t count, hipStream_t s) {
    alloc(count);
    if (count) QPX_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

// Hamiltonian tables on the device.
//   V[P*npair + R] = (pq|rs) with P = pair(p,q), R = pair(r,s) (0-based
//   packed pair index P = max(max+1)/2 + min over 0-based p,q).
struct Ham {
  int M = 0, na = 0, nb = 0, nint = 0, npair = 0;
  std::vector<double> h;          // M*M column-major
  std::vector<double> eri;        // packed n8
  DBuf<double> V;                 // npair^2
  DBuf<float> Vf;
  DBuf<double> hcore;             // M*M
  DBuf<float> hcoref;
  DBuf<double> Jd, Kd;            // M*M: J_ij=(ii|jj), K_ij=(ij|ji)
  DBuf<double> JAd;               // M*M: J_ij - K_ij (mo_bielec_integral_jj_anti)
  DBuf<double> fock_cs;           // M*M fock_operator_closed_shell_ref_bitmask
  DBuf<int> ptab;                 // M*M pair index
  DBuf<double> GS, GO;            // M^3 single-excitation tables (context.hip)
  DBuf<float> GSf, GOf;
  // M^4 table G4[((i*M + j)*M + y)*M + x] = (ix|jy), built on first use by
  // the sparse sigma path when M <= kG4MaxOrb (ensure_g4)
  DBuf<double> G4;
  DBuf<float> G4f;
  double ref_energy = 0.0;        // ref_bitmask_energy of the HF determinant
};

// Connectivity of unique spin strings (within the unique set).
struct SpinConn {
  DBuf<int> s_ptr, s_idx;         // singles: CSR over strings
  DBuf<uint32_t> s_info;          // h | p<<8 | (phase<0)<<16 (0-based orbitals)
  DBuf<double> s_val;             // same-spin part of the single element, phase included
  DBuf<float> s_valf;
  DBuf<int> d_ptr, d_idx;         // doubles
  DBuf<double> d_val;             // full same-spin double element, phase included
  DBuf<float> d_valf;
  int max_singles = 0, max_doubles = 0;
  int64_t n_singles = 0, n_doubles = 0;
};

// Sub-string buckets of the sparse sigma path (sigma_sparse.hip): per
// excitation class, the sorted entries (determinant | removed orbitals) and,
// per (determinant, slot), the range of its bucket.
struct Buckets {
  int nslot = 0;
  int64_t visited = 0;   // sum of squared bucket sizes
  DBuf<uint64_t> pay;
  DBuf<int2> range;
};
struct SparseIdx {
  bool valid = false;
  DBuf<int> a1, a2, b1, b2;   // sub-string ids per unique string
  Buckets cls[5];             // ab, a1, a2, b1, b2
  void clear() {
    valid = false;
    a1.release(); a2.release(); b1.release(); b2.release();
    for (auto &b : cls) { b.pay.release(); b.range.release(); b.nslot = 0; }
  }
};

struct Wfn {
  int64_t n_det = 0;
  int nint = 0;
  int n_a = 0, n_b = 0;
  bool valid = false;
  std::vector<uint64_t> astr_h, bstr_h;   // unique strings (n*nint)
  std::vector<int64_t> perm;              // bilinear position -> psi index
  DBuf<uint64_t> psi;                     // determinants in psi order (n_det, 2, nint)
  DBuf<uint64_t> astr, bstr;
  DBuf<int> col_ptr, csc_a, csc_b;        // CSC by beta column; entries sorted by alpha
  DBuf<int> row_ptr, csr_b, csr2csc;      // CSR by alpha row; entries sorted by beta
  DBuf<int64_t> perm_d;                   // bilinear position -> psi index
  SpinConn ca, cb;               // built on first use by the dense-map or scan path (ensure_conn)
  bool conn_valid = false;
  DBuf<double> hdiag, sdiag;              // bilinear order
  // per-string scan costs for the generic alpha/beta kernel
  DBuf<int64_t> cost_a, cost_b;
  std::vector<int> col_len_h, row_len_h;
  int max_col = 0, max_row = 0;
  SparseIdx sp;                           // built on first use of the sparse path
};

struct Workspace {
  DBuf<char> buf;
  char *get(size_t bytes) {
    if (buf.n < bytes) buf.alloc(bytes + (bytes >> 3));
    return buf.p;
  }
};

}  // namespace qpx

struct qpx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  qpx::Ham H;
  qpx::Wfn W;
  bool ham_set = false;
  int lds_per_block = 65536;
  int n_cu = 256;
  qpx::Workspace ws_sigma_io;   // host-order u/v/s staging for qpx_sigma
  qpx::Workspace ws_sigma_t;    // transposed u_t/v_t/s_t
  qpx::Workspace ws_misc;
  int64_t sel_stats[4] = {0, 0, 0, 0};
  std::vector<int64_t> sel_ticks;   // per-block durations of the last selection (benchmark hook)
  int sigma_path = 0;           // last sigma path: 1 = dense-map, 2 = generic scan, 3 = sparse buckets
  int force_path = 0;           // 0 auto, else the path to use (tests)
  int verbosity = 0;            // 1: Davidson iteration table on stderr
  hipEvent_t kev[3] = {nullptr, nullptr, nullptr};  // per-kernel timing (benchmark hook)
  bool ktiming = false;
};

namespace qpx {
// implemented in context.hip: builds Ham::G4/G4f if M <= kG4MaxOrb and
// returns whether the table exists
bool ensure_g4(qpx_ctx *c);
// implemented in wavefunction.hip; ensure_conn builds the spin-string
// connectivity and scan costs of the current wavefunction if needed
void build_wavefunction(qpx_ctx *c, int n_int, int64_t n_det, const uint64_t *psi_det);
void ensure_conn(qpx_ctx *c);
// implemented in sigma.hip
void sigma_device(qpx_ctx *c, int prec, int n_st, const void *u_t, void *v_t, void *s_t, hipStream_t s);
// implemented in sigma_sparse.hip
bool sparse_ok(qpx_ctx *c, int n_st);
void sigma_sparse(qpx_ctx *c, int prec, int n_st, const void *u_t, void *v_t, void *s_t, hipStream_t s);
}  // namespace qpx

// Run a body, translating exceptions into return codes.
template <typename F>
static inline int qpx_guard(qpx_ctx *c, F &&f) {
  if (!c) return 1;
  try {
    f();
    return 0;
  } catch (const qpx::Error &e) {
    c->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    c->err = e.what();
    return 3;
  }
}
