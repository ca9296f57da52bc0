// mlp.hip -- the block-proposal network: a small MLP trained and run with
// hand-written BF16 MFMA kernels (FP32 accumulation) on gfx950.
//
// No counterpart in the reference: it ranks the (generator, hole-pair)
// blocks of select_singles_and_doubles (plugins/Full_CI_ZMQ/selection.irp.f:
// 412-575) -- the same unit the reference's `subset` / fragment_count filter
// acts on (:561-562) -- by predicted log|PT2 contribution|.
//
// Network: x (D_IN features) -> 256 -> 256 -> 256 (ReLU) -> 1.  Features of a
// block: learned 32-d embeddings of h1 and h2 (one table over orbitals),
// h_11 and h_22 orbital energies, log|c| of the generator, one-hot spin
// class (aa, bb, ab), one-hot excitation degree of the generator from HF
// (0..4), zero padding to D_IN = 80.
//
// GEMM: C(MxN) = A(MxK) B(KxN), A and B bf16 with arbitrary row/column
// strides (so A^T, B^T need no copies), C fp32; workgroup tile 64x64, four
// wavefronts each owning 32x32 = 2x2 tiles of v_mfma_f32_16x16x32_bf16
// (CDNA4's bf16 MFMA with K = 32, one per tile and K-tile of 32).
// Lane layout of the 16x16x32 MFMA: lane l holds A[l%16][8(l/16)+0..7],
// B[8(l/16)+0..7][l%16] and D[4(l/16)+0..3][l%16].  Both tiles sit in LDS
// with k contiguous (A as [i][k], B as [j][k]), so each lane's MFMA operand
// is one 16-byte read; they are staged in runs of 8 elements (16-byte loads)
// along whichever index of the operand is contiguous in memory.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <mutex>
#include <vector>
#include <hip/hip_bf16.h>
#include "qpx_internal.h"

using namespace qpx;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffff) > 0x7f800000) return 0x7fc0;  // NaN
  u += 0x7fff + ((u >> 16) & 1);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

struct GemmArgs {
  const uint16_t *A;  // bf16
  int64_t a_rs, a_cs; // A(i,k) = A[i*a_rs + k*a_cs]
  const uint16_t *B;
  int64_t b_rs, b_cs; // B(k,j) = B[k*b_rs + j*b_cs]
  float *C;           // C(i,j) = C[i*c_rs + j]
  int64_t c_rs;
  int M, N, K;
  const float *bias;  // per column j (nullable)
  int relu;           // epilogue: C = relu(C + bias)
  uint16_t *Cb;       // optional bf16 copy of C (same layout)
  float beta;         // C = acc + beta * C_old (gradient accumulation)
  float *part;        // split-K partial products (ksplit x M x N), or null
  int ksplit;         // K is cut into ksplit chunks (blockIdx.z); 1 = no split
  int a_run, b_run;   // staging: 1 = runs along k, 2 = along i (A) / j (B), 0 = element by element
  const float *mask;  // optional: C(i,j) = 0 where mask(i,j) <= 0 (same layout as C; a ReLU's backward)
};

// LDK pads a tile row to 80 bytes: rows stay 16-byte aligned for the staging
// stores and 8-byte aligned for the operand reads
constexpr int TM = 64, TN = 64, TK = 32, LDK = TK + 8;

// Stages the rows x TK tile X(r, k) = P[r*rs + k*cs], r in [r0, r0+64),
// k in [k0, k0+TK), into T[r - r0][k - k0]; entries outside r < rmax,
// k < kmax are zero.  Each of the 256 threads moves one run of 8 elements:
// run = 1 reads it along k (cs == 1), run = 2 along r (rs == 1), run = 0
// element by element (any strides).  A run entirely in bounds is one
// 16-byte load; the host picks run 1 or 2 only for 16-byte-aligned runs.
__device__ inline void stage_tile(uint16_t (*T)[LDK], const uint16_t *__restrict__ P, int64_t rs, int64_t cs,
                                  int run, int r0, int k0, int rmax, int kmax, int tid) {
  u32x4 v = {0u, 0u, 0u, 0u};   // the run, two elements per word, element e in bits 16(e&1) of word e>>1
  if (run == 2) {
    const int kk = tid >> 3, rr = (tid & 7) * 8;
    const int gk = k0 + kk, gr = r0 + rr;
    if (gk < kmax) {
      if (gr + 8 <= rmax) {
        v = *(const u32x4 *)(P + (int64_t)gk * cs + gr);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (gr + e < rmax) v[e >> 1] |= (uint32_t)P[(int64_t)gk * cs + gr + e] << (16 * (e & 1));
      }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) T[rr + e][kk] = (uint16_t)(v[e >> 1] >> (16 * (e & 1)));
  } else {
    const int rr = tid >> 2, kk = (tid & 3) * 8;
    const int gr = r0 + rr, gk = k0 + kk;
    if (gr < rmax) {
      if (run == 1 && gk + 8 <= kmax) {
        v = *(const u32x4 *)(P + (int64_t)gr * rs + gk);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (gk + e < kmax) v[e >> 1] |= (uint32_t)P[(int64_t)gr * rs + (int64_t)(gk + e) * cs] << (16 * (e & 1));
      }
    }
    *(u32x4 *)&T[rr][kk] = v;
  }
}

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ __align__(16) uint16_t As[TM][LDK];   // As[i][k]
  __shared__ __align__(16) uint16_t Bs[TN][LDK];   // Bs[j][k]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
  const int wi = (wv >> 1) * 32, wj = (wv & 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
  // this workgroup's slice of K (all of it without a split); slices are
  // whole multiples of TK
  const int kc = ((g.K + g.ksplit - 1) / g.ksplit + TK - 1) / TK * TK;
  const int k_lo = (int)blockIdx.z * kc, k_hi = min(g.K, k_lo + kc);
  // A and B operands of v_mfma_f32_16x16x32_bf16: lane (r, g = lane / 16)
  // supplies row r of the tile at k = 8 g .. 8 g + 7; the result rows are
  // 4 g .. 4 g + 3 of column r
  const int r = lane & 15, q = (lane >> 4) * 4, qk = (lane >> 4) * 8;
  for (int k0 = k_lo; k0 < k_hi; k0 += TK) {
    __syncthreads();
    // B(k, j) = B[k*b_rs + j*b_cs] is staged as the tile X(j, k)
    stage_tile(As, g.A, g.a_rs, g.a_cs, g.a_run, i0, k0, g.M, k_hi, tid);
    stage_tile(Bs, g.B, g.b_cs, g.b_rs, g.b_run, j0, k0, g.N, k_hi, tid);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < TK; ks += 32) {
      bf16x8 av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = *(const bf16x8 *)&As[wi + 16 * a + r][ks + qk];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = *(const bf16x8 *)&Bs[wj + 16 * b + r][ks + qk];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
  }
  // epilogue: lane holds D[q+0..3][r] of each 16x16 tile
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        int gi = i0 + wi + 16 * a + q + t, gj = j0 + wj + 16 * b + r;
        if (gi < g.M && gj < g.N && g.ksplit > 1) {
          g.part[((int64_t)blockIdx.z * g.M + gi) * g.N + gj] = acc[a][b][t];
        } else if (gi < g.M && gj < g.N) {
          float v = acc[a][b][t];
          float *c = g.C + gi * g.c_rs + gj;
          if (g.beta != 0.f) v += g.beta * (*c);
          if (g.bias) v += g.bias[gj];
          if (g.relu) v = v > 0.f ? v : 0.f;
          if (g.mask) v = g.mask[gi * g.c_rs + gj] > 0.f ? v : 0.f;
          *c = v;
          if (g.Cb) g.Cb[gi * g.c_rs + gj] = f2bf(v);
        }
      }
}

// split-K: C = sum of the partial products in slice order, then the epilogue
__global__ void gemm_splitk_reduce_kernel(GemmArgs g) {
  const int64_t mn = (int64_t)g.M * g.N;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < mn; x += (int64_t)gridDim.x * blockDim.x) {
    const int gi = (int)(x / g.N), gj = (int)(x - (int64_t)gi * g.N);
    float v = 0.f;
    // unrolled so that the slice loads are in flight together; the adds stay in slice order
#pragma unroll 8
    for (int z = 0; z < g.ksplit; ++z) v += g.part[(int64_t)z * mn + x];
    float *c = g.C + gi * g.c_rs + gj;
    if (g.beta != 0.f) v += g.beta * (*c);
    if (g.bias) v += g.bias[gj];
    if (g.relu) v = v > 0.f ? v : 0.f;
    if (g.mask) v = g.mask[gi * g.c_rs + gj] > 0.f ? v : 0.f;
    *c = v;
    if (g.Cb) g.Cb[gi * g.c_rs + gj] = f2bf(v);
  }
}

__global__ void to_bf16_kernel(const float *__restrict__ x, uint16_t *__restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

// features of the blocks: desc = (h1, h2 (0-based), spin class 0..2,
// degree 0..4) int32 x4; fdesc = (eps1, eps2, log|c|) float x3
__global__ void features_kernel(const int *__restrict__ desc, const float *__restrict__ fdesc,
                                const float *__restrict__ emb, int n, int din, float *__restrict__ X,
                                uint16_t *__restrict__ Xb) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < (int64_t)n * din;
       x += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(x / din), f = (int)(x - (int64_t)i * din);
    const int *d = desc + 4 * i;
    float v = 0.f;
    if (f < 32) v = emb[d[0] * 32 + f];
    else if (f < 64) v = emb[d[1] * 32 + (f - 32)];
    else if (f < 67) v = fdesc[3 * i + (f - 64)];
    else if (f < 70) v = (d[2] == f - 67) ? 1.f : 0.f;
    else if (f < 75) v = (d[3] == f - 70) ? 1.f : 0.f;
    X[x] = v;
    Xb[x] = f2bf(v);
  }
}

// y = H3 w4 + b4 (one output), loss partials and dy = 2 (y - t) / n
// y = H w4 + b4 per row, one wavefront per row (the row read coalesced,
// reduced by shuffles in a fixed order); with targets, dy = 2 (y - t) / n and
// one partial sum of (y - t)^2 per workgroup
__global__ __launch_bounds__(256) void head_kernel(const float *__restrict__ H, const float *__restrict__ w4,
                                                   const float *__restrict__ b4p,
                                                   int n, int hdim, const float *__restrict__ t,
                                                   float *__restrict__ y, float *__restrict__ dy,
                                                   float *__restrict__ loss_part) {
  __shared__ float red[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float l = 0.f;
  for (int i = blockIdx.x * 4 + wv; i < n; i += gridDim.x * 4) {
    float s = 0.f;
    for (int k = lane; k < hdim; k += 64) s += H[(int64_t)i * hdim + k] * w4[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    s += *b4p;
    if (lane == 0) {
      y[i] = s;
      if (t) {
        const float e = s - t[i];
        l += e * e;
        if (dy) dy[i] = 2.f * e / n;
      }
    }
  }
  if (lane == 0) red[wv] = l;
  __syncthreads();
  if (threadIdx.x == 0 && loss_part) loss_part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

static inline int head_blocks(int n) { return std::max(1, std::min(1024, (n + 3) / 4)); }

// out[0] = (sum of the nb loss partials, in a fixed order) / n
__global__ __launch_bounds__(256) void loss_reduce_kernel(const float *__restrict__ part, int nb, int n,
                                                          float *__restrict__ out) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] / (float)n;
}

// dZ = (dy w4^T) masked by H > 0; bf16 copy; dw4 = H^T dy partial sums
__global__ void head_back_kernel(const float *__restrict__ H, const float *__restrict__ w4,
                                 const float *__restrict__ dy, int n, int hdim, float *__restrict__ dZ,
                                 uint16_t *__restrict__ dZb) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < (int64_t)n * hdim;
       x += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(x / hdim), k = (int)(x - (int64_t)i * hdim);
    float v = H[x] > 0.f ? dy[i] * w4[k] : 0.f;
    dZ[x] = v;
    dZb[x] = f2bf(v);
  }
}

// column sums of a (n x m) row-major fp32 matrix (bias gradients), and the
// head weight gradient dw4[k] = sum_i H[i,k] dy[i]
// out[k] = sum_i A[i][k] (w[i]): a workgroup of 256 threads per CB columns
// (CB = 32, or m when m < 32); CB consecutive threads read one row's
// columns (coalesced), 256 / CB row groups stride the rows, and the groups
// are summed in order (deterministic).  Launch with ceil(m / CB) groups.
__global__ __launch_bounds__(256) void colsum_kernel(const float *__restrict__ A, const float *__restrict__ w, int n,
                                                     int m, float *__restrict__ out) {
  __shared__ float red[256];
  const int cb = m < 32 ? m : 32, groups = 256 / cb;
  const int t = threadIdx.x, c = t % cb, g = t / cb;
  const int k = blockIdx.x * cb + c;
  float s = 0.f;
  if (g < groups && k < m)
    for (int i = g; i < n; i += groups) s += A[(int64_t)i * m + k] * (w ? w[i] : 1.f);
  red[t] = s;
  __syncthreads();
  if (t < cb && k < m) {
    float acc = 0.f;
    for (int q = 0; q < groups; ++q) acc += red[q * cb + t];
    out[k] = acc;
  }
}

static inline int colsum_grid(int m) { return m < 32 ? 1 : (m + 31) / 32; }

// Column sums of a wide matrix in two passes, each in a fixed order:
// workgroup (column block b, row chunk p) sums 64 columns over the rows of
// its chunk (4 row phases per column, combined in phase order) into
// part[p][k]; the second pass adds the chunks in chunk order.  Spreading the
// rows over many workgroups is what fills the GPU: one workgroup per 32
// columns left all but 8 CUs idle for m = 256.
__global__ __launch_bounds__(256) void colsum_part_kernel(const float *__restrict__ A, const float *__restrict__ w,
                                                          int n, int m, int chunk, float *__restrict__ part) {
  __shared__ float red[4][64];
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + c, p = blockIdx.y;
  const int r0 = p * chunk, r1 = min(n, r0 + chunk);
  float s = 0.f;
  if (k < m)
    for (int i = r0 + g; i < r1; i += 4) s += A[(int64_t)i * m + k] * (w ? w[i] : 1.f);
  red[g][c] = s;
  __syncthreads();
  if (g == 0 && k < m) part[(int64_t)p * m + k] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}

// one wavefront per column: lane l adds chunks l, l + 64, ...; the lanes
// are then combined by a fixed shuffle tree
__global__ __launch_bounds__(256) void colsum_final_kernel(const float *__restrict__ part, int np, int m,
                                                           float *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= m) return;   // whole wavefront
  float s = 0.f;
  for (int p = lane; p < np; p += 64) s += part[(int64_t)p * m + k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[k] = s;
}

__global__ void emb_grad_kernel(const int *__restrict__ desc, const float *__restrict__ dX, int n, int din,
                                float *__restrict__ demb) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < (int64_t)n * 64;
       x += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(x / 64), f = (int)(x - (int64_t)i * 64);
    int h = f < 32 ? desc[4 * i] : desc[4 * i + 1];
    atomicAdd(&demb[h * 32 + (f & 31)], dX[(int64_t)i * din + f]);
  }
}

// Adam over all parameter tensors in one launch: tensor q covers elements
// [start[q], start[q+1]) of the concatenation.  Per element:
// m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// p -= lr (m / c1) / (sqrt(v / c2) + eps), with c1, c2 the bias corrections
// 1 - b1^t, 1 - b2^t; the bf16 copy of p is refreshed where there is one.
constexpr int kAdamTensors = 9;
struct AdamSet {
  float *p[kAdamTensors];
  const float *g[kAdamTensors];
  float *m[kAdamTensors];
  float *v[kAdamTensors];
  uint16_t *pb[kAdamTensors];
  int64_t start[kAdamTensors + 1];
};

__global__ void adam_multi_kernel(AdamSet S, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t total = S.start[kAdamTensors];
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
    int q = 0;
#pragma unroll
    for (int t = 1; t < kAdamTensors; ++t) q += (x >= S.start[t]);
    const int64_t i = x - S.start[q];
    const float gi = S.g[q][i];
    const float mi = b1 * S.m[q][i] + (1.f - b1) * gi;
    const float vi = b2 * S.v[q][i] + (1.f - b2) * gi * gi;
    S.m[q][i] = mi;
    S.v[q][i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    const float pi = S.p[q][i] - upd;
    S.p[q][i] = pi;
    if (S.pb[q]) S.pb[q][i] = f2bf(pi);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the allocation-free prediction (qpx_mlp_predict_static): the
// staged descriptors, the features and the three layers' activations of at
// most kStaticRows blocks, in one file-scope device array.  Each GPU the
// library runs on holds its own copy until the process ends; its device
// address is looked up once per device (scratch_address) and only that
// address is handed to kernels and copies.
namespace {
constexpr int kStaticRows = 2048;
constexpr int kDin = 80, kHid = 256;
constexpr size_t al256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t kScratchBytes =
    al256(sizeof(int) * 4 * kStaticRows) + al256(sizeof(float) * 3 * kStaticRows) +
    al256(sizeof(float) * kDin * kStaticRows) + al256(sizeof(uint16_t) * kDin * kStaticRows) +
    3 * al256(sizeof(float) * kHid * kStaticRows) + 3 * al256(sizeof(uint16_t) * kHid * kStaticRows) +
    al256(sizeof(float) * kStaticRows);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// device address of this device's scratch (the current device must be
// `device`); looked up once per device
void *scratch_address(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  QPX_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range");
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    QPX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    QPX_REQUIRE(p != nullptr, "no device address for the prediction scratch");
    g_scratch_dev[device] = p;
  }
  return g_scratch_dev[device];
}
}  // namespace

// ---------------------------------------------------------------------------
struct qpx_mlp {
  int din = 80, hid = 256, n_orb = 0;
  hipStream_t st = nullptr;
  int device = 0;
  // parameters (fp32 master) and bf16 copies; layer l: W (in x out) row-major
  DBuf<float> W1, W2, W3, b1, b2, b3, w4, emb;
  DBuf<uint16_t> W1b, W2b, W3b;
  DBuf<float> gW1, gW2, gW3, gb1, gb2, gb3, gw4, gemb, gb4;
  DBuf<float> ws;   // split-K partial products of the weight gradients (sized once)
  DBuf<float> mW1, mW2, mW3, mb1, mb2, mb3, mw4, memb, vW1, vW2, vW3, vb1, vb2, vb3, vw4, vemb;
  DBuf<float> b4, mb4, vb4;   // the output bias and its Adam moments (one element each, on the device)
  int step = 0;
  // activations for a batch
  int cap = 0;
  DBuf<float> X, Z1, Z2, Z3, y, dy, dZ3, dZ2, dZ1, dX, lossp;
  DBuf<uint16_t> Xb, H1b, H2b, H3b, dZ3b, dZ2b, dZ1b;
  DBuf<int> desc;
  DBuf<float> fdesc, tgt;
  void *scratch = nullptr;   // device address of this device's prediction scratch
  std::string err;
};

namespace {

// C = A B (+ epilogue).  With a workspace, a product whose tile grid
// leaves most CUs idle and whose K is long (the weight gradients, K = batch)
// is cut along K across workgroups and reduced in slice order.
void gemm(hipStream_t st, const uint16_t *A, int64_t ars, int64_t acs, const uint16_t *B, int64_t brs, int64_t bcs,
          float *C, int64_t crs, int M, int N, int K, const float *bias, int relu, uint16_t *Cb, float beta,
          DBuf<float> *ws = nullptr, const float *mask = nullptr) {
  // 16-byte runs need a 16-byte-aligned base and a multiple-of-8 stride
  // across runs
  const auto run_of = [](const uint16_t *P, int64_t rs, int64_t cs) {
    if (((uintptr_t)P & 15) != 0) return 0;
    if (cs == 1 && rs % 8 == 0) return 1;
    if (rs == 1 && cs % 8 == 0) return 2;
    return 0;
  };
  GemmArgs g{A, ars, acs, B, brs, bcs, C, crs, M, N, K, bias, relu, Cb, beta, nullptr, 1,
             run_of(A, ars, acs), run_of(B, bcs, brs), mask};
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
  const int tiles = (int)(grid.x * grid.y);
  constexpr int kGrain = 16;   // split-K sizing unit (slices are then rounded up to TK)
  if (ws && tiles < 128 && K >= 8 * kGrain) {
    g.ksplit = std::max(1, std::min({256 / tiles, K / (4 * kGrain), 32}));
    if (g.ksplit > 1 && ws->n >= (size_t)g.ksplit * M * N) {
      g.part = ws->p;
      grid.z = g.ksplit;
    } else {
      g.ksplit = 1;
    }
  }
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, st, g);
  QPX_HIP(hipGetLastError());
  if (g.ksplit > 1) {
    const int64_t mn = (int64_t)M * N;
    hipLaunchKernelGGL(gemm_splitk_reduce_kernel, dim3((unsigned)std::min<int64_t>((mn + 255) / 256, 4096)), dim3(256),
                       0, st, g);
    QPX_HIP(hipGetLastError());
  }
}

// out[k] = sum_i A[i][k] w[i] (w null: 1) over the n x mc row-major A; the
// wide case uses the workspace for its per-chunk partial sums
void colsum(hipStream_t st, const float *A, const float *w, int n, int mc, float *out, DBuf<float> &ws) {
  if (mc < 32) {
    hipLaunchKernelGGL(colsum_kernel, dim3(colsum_grid(mc)), dim3(256), 0, st, A, w, n, mc, out);
    QPX_HIP(hipGetLastError());
    return;
  }
  const int nb = (mc + 63) / 64;
  int np = std::max(1, std::min(256 / nb, (n + 63) / 64));
  np = (int)std::max<int64_t>(1, std::min<int64_t>(np, (int64_t)(ws.n / (size_t)mc)));
  const int chunk = std::max(1, (n + np - 1) / np);
  np = std::max(1, (n + chunk - 1) / chunk);
  hipLaunchKernelGGL(colsum_part_kernel, dim3(nb, np), dim3(256), 0, st, A, w, n, mc, chunk, ws.p);
  hipLaunchKernelGGL(colsum_final_kernel, dim3((mc + 3) / 4), dim3(256), 0, st, (const float *)ws.p, np, mc, out);
  QPX_HIP(hipGetLastError());
}

void ensure(qpx_mlp *m, int n) {
  if (n <= m->cap) return;
  int c = n;
  m->X.alloc((size_t)c * m->din); m->Xb.alloc((size_t)c * m->din); m->dX.alloc((size_t)c * m->din);
  m->Z1.alloc((size_t)c * m->hid); m->Z2.alloc((size_t)c * m->hid); m->Z3.alloc((size_t)c * m->hid);
  m->H1b.alloc((size_t)c * m->hid); m->H2b.alloc((size_t)c * m->hid); m->H3b.alloc((size_t)c * m->hid);
  m->dZ1.alloc((size_t)c * m->hid); m->dZ2.alloc((size_t)c * m->hid); m->dZ3.alloc((size_t)c * m->hid);
  m->dZ1b.alloc((size_t)c * m->hid); m->dZ2b.alloc((size_t)c * m->hid); m->dZ3b.alloc((size_t)c * m->hid);
  m->y.alloc(c); m->dy.alloc(c); m->lossp.alloc(1025);   // head partials + one loss slot
  m->desc.alloc((size_t)4 * c); m->fdesc.alloc((size_t)3 * c); m->tgt.alloc(c);
  if (!m->ws.p) m->ws.alloc((size_t)32 * m->hid * m->hid);   // the largest split: 32 x hid x hid
  m->cap = c;
}

int grid1(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 16384); }

// the buffers of one forward pass over n rows
struct FwdBufs {
  const int *desc;
  const float *fdesc;
  float *X, *Z1, *Z2, *Z3, *y;
  uint16_t *Xb, *H1b, *H2b, *H3b;
};

void forward_bufs(qpx_mlp *m, int n, const FwdBufs &b) {
  hipStream_t st = m->st;
  const int D = m->din, Hd = m->hid;
  hipLaunchKernelGGL(features_kernel, dim3(grid1((int64_t)n * D)), dim3(256), 0, st, b.desc, b.fdesc, m->emb.p, n, D,
                     b.X, b.Xb);
  // Z = act(X W + b): A = X (n x D, row-major), B = W (D x H, row-major)
  gemm(st, b.Xb, D, 1, m->W1b.p, Hd, 1, b.Z1, Hd, n, Hd, D, m->b1.p, 1, b.H1b, 0.f);
  gemm(st, b.H1b, Hd, 1, m->W2b.p, Hd, 1, b.Z2, Hd, n, Hd, Hd, m->b2.p, 1, b.H2b, 0.f);
  gemm(st, b.H2b, Hd, 1, m->W3b.p, Hd, 1, b.Z3, Hd, n, Hd, Hd, m->b3.p, 1, b.H3b, 0.f);
}

void forward(qpx_mlp *m, int n) {
  forward_bufs(m, n, FwdBufs{m->desc.p, m->fdesc.p, m->X.p, m->Z1.p, m->Z2.p, m->Z3.p, m->y.p, m->Xb.p, m->H1b.p,
                             m->H2b.p, m->H3b.p});
}

// the forward buffers carved out of the prediction scratch at its device
// address base (the layout kScratchBytes is sized for)
FwdBufs static_bufs(void *base) {
  unsigned char *p = static_cast<unsigned char *>(base);
  auto take = [&](size_t bytes) {
    unsigned char *q = p;
    p += al256(bytes);
    return q;
  };
  FwdBufs b{};
  b.desc = reinterpret_cast<const int *>(take(sizeof(int) * 4 * kStaticRows));
  b.fdesc = reinterpret_cast<const float *>(take(sizeof(float) * 3 * kStaticRows));
  b.X = reinterpret_cast<float *>(take(sizeof(float) * kDin * kStaticRows));
  b.Xb = reinterpret_cast<uint16_t *>(take(sizeof(uint16_t) * kDin * kStaticRows));
  b.Z1 = reinterpret_cast<float *>(take(sizeof(float) * kHid * kStaticRows));
  b.Z2 = reinterpret_cast<float *>(take(sizeof(float) * kHid * kStaticRows));
  b.Z3 = reinterpret_cast<float *>(take(sizeof(float) * kHid * kStaticRows));
  b.H1b = reinterpret_cast<uint16_t *>(take(sizeof(uint16_t) * kHid * kStaticRows));
  b.H2b = reinterpret_cast<uint16_t *>(take(sizeof(uint16_t) * kHid * kStaticRows));
  b.H3b = reinterpret_cast<uint16_t *>(take(sizeof(uint16_t) * kHid * kStaticRows));
  b.y = reinterpret_cast<float *>(take(sizeof(float) * kStaticRows));
  return b;
}

// One Adam step on the batch in place, enqueued on the network's stream
// without waiting for it; the batch's mean squared error goes to *loss_out
// (device memory).
void train_step_async(qpx_mlp *m, int n, float lr, float *loss_out) {
  hipStream_t st = m->st;
  const int D = m->din, Hd = m->hid;
  forward(m, n);
  const int nb = head_blocks(n);
  hipLaunchKernelGGL(head_kernel, dim3(nb), dim3(256), 0, st, m->Z3.p, m->w4.p, m->b4.p, n, Hd, m->tgt.p, m->y.p,
                     m->dy.p, m->lossp.p);
  // head grads
  colsum(st, m->Z3.p, m->dy.p, n, Hd, m->gw4.p, m->ws);
  colsum(st, m->dy.p, nullptr, n, 1, m->gb4.p, m->ws);
  hipLaunchKernelGGL(head_back_kernel, dim3(grid1((int64_t)n * Hd)), dim3(256), 0, st, m->Z3.p, m->w4.p, m->dy.p,
                     n, Hd, m->dZ3.p, m->dZ3b.p);
  // layer 3: dW3 = H2^T dZ3 (H x H, K = n); db3 = colsum dZ3; dH2 = dZ3 W3^T
  gemm(st, m->H2b.p, 1, Hd, m->dZ3b.p, Hd, 1, m->gW3.p, Hd, Hd, Hd, n, nullptr, 0, nullptr, 0.f, &m->ws);
  colsum(st, m->dZ3.p, nullptr, n, Hd, m->gb3.p, m->ws);
  // dZ2 = (dZ3 W3^T) masked by Z2 > 0, with its bf16 copy, in the epilogue
  gemm(st, m->dZ3b.p, Hd, 1, m->W3b.p, 1, Hd, m->dZ2.p, Hd, n, Hd, Hd, nullptr, 0, m->dZ2b.p, 0.f, nullptr, m->Z2.p);
  // layer 2
  gemm(st, m->H1b.p, 1, Hd, m->dZ2b.p, Hd, 1, m->gW2.p, Hd, Hd, Hd, n, nullptr, 0, nullptr, 0.f, &m->ws);
  colsum(st, m->dZ2.p, nullptr, n, Hd, m->gb2.p, m->ws);
  gemm(st, m->dZ2b.p, Hd, 1, m->W2b.p, 1, Hd, m->dZ1.p, Hd, n, Hd, Hd, nullptr, 0, m->dZ1b.p, 0.f, nullptr, m->Z1.p);
  // layer 1: dW1 = X^T dZ1 (D x H); dX = dZ1 W1^T (n x D)
  gemm(st, m->Xb.p, 1, D, m->dZ1b.p, Hd, 1, m->gW1.p, Hd, D, Hd, n, nullptr, 0, nullptr, 0.f, &m->ws);
  colsum(st, m->dZ1.p, nullptr, n, Hd, m->gb1.p, m->ws);
  gemm(st, m->dZ1b.p, Hd, 1, m->W1b.p, 1, Hd, m->dX.p, D, n, D, Hd, nullptr, 0, nullptr, 0.f);
  QPX_HIP(hipMemsetAsync(m->gemb.p, 0, sizeof(float) * m->emb.n, st));
  hipLaunchKernelGGL(emb_grad_kernel, dim3(grid1((int64_t)n * 64)), dim3(256), 0, st, m->desc.p, m->dX.p, n, D,
                     m->gemb.p);
  QPX_HIP(hipGetLastError());
  if (lr > 0.f) {
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    AdamSet S{};
    int nq = 0;
    auto ad = [&](DBuf<float> &p, DBuf<float> &g, DBuf<float> &mm, DBuf<float> &vv, uint16_t *pb) {
      S.p[nq] = p.p;
      S.g[nq] = g.p;
      S.m[nq] = mm.p;
      S.v[nq] = vv.p;
      S.pb[nq] = pb;
      S.start[nq + 1] = S.start[nq] + (int64_t)p.n;
      ++nq;
    };
    ad(m->W1, m->gW1, m->mW1, m->vW1, m->W1b.p);
    ad(m->W2, m->gW2, m->mW2, m->vW2, m->W2b.p);
    ad(m->W3, m->gW3, m->mW3, m->vW3, m->W3b.p);
    ad(m->b1, m->gb1, m->mb1, m->vb1, nullptr);
    ad(m->b2, m->gb2, m->mb2, m->vb2, nullptr);
    ad(m->b3, m->gb3, m->mb3, m->vb3, nullptr);
    ad(m->w4, m->gw4, m->mw4, m->vw4, nullptr);
    ad(m->emb, m->gemb, m->memb, m->vemb, nullptr);
    ad(m->b4, m->gb4, m->mb4, m->vb4, nullptr);
    QPX_REQUIRE(nq == kAdamTensors, "Adam tensor count");
    hipLaunchKernelGGL(adam_multi_kernel, dim3(grid1(S.start[kAdamTensors])), dim3(256), 0, st, S, lr, b1, b2, eps,
                       c1, c2);
    QPX_HIP(hipGetLastError());
  }
  // the step's mean squared error, left on the device
  hipLaunchKernelGGL(loss_reduce_kernel, dim3(1), dim3(256), 0, st, m->lossp.p, nb, n, loss_out);
  QPX_HIP(hipGetLastError());
}

// one step with its loss returned (synchronises the stream)
float train_step(qpx_mlp *m, int n, float lr) {
  train_step_async(m, n, lr, m->lossp.p + 1024);
  float l = 0.f;
  QPX_HIP(hipMemcpyAsync(&l, m->lossp.p + 1024, sizeof(float), hipMemcpyDeviceToHost, m->st));
  QPX_HIP(hipStreamSynchronize(m->st));
  return l;
}

// batch rows of the device-resident training set, in permutation order
__global__ void gather_batch_kernel(const int *__restrict__ perm, int nb, const int *__restrict__ desc_all,
                                    const float *__restrict__ fdesc_all, const float *__restrict__ tgt_all,
                                    int *__restrict__ desc, float *__restrict__ fdesc, float *__restrict__ tgt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int64_t r = perm[i];
#pragma unroll
  for (int q = 0; q < 4; ++q) desc[4 * (int64_t)i + q] = desc_all[4 * r + q];
#pragma unroll
  for (int q = 0; q < 3; ++q) fdesc[3 * (int64_t)i + q] = fdesc_all[3 * r + q];
  tgt[i] = tgt_all[r];
}

template <typename F>
int mguard(qpx_mlp *m, F &&f) {
  if (!m) return 1;
  try {
    f();
    return 0;
  } catch (const qpx::Error &e) {
    m->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    m->err = e.what();
    return 3;
  }
}

// The descriptors index the embedding table and the one-hot slots: a row
// outside {h1, h2 in [0, n_orb), spin class in 0..2, degree in 0..4} would
// make the feature and embedding-gradient kernels read or write out of
// bounds, so it is rejected on the host before anything is uploaded.
void check_desc(const qpx_mlp *m, int64_t n, const int *desc) {
  QPX_REQUIRE(desc, "null descriptor array");
  for (int64_t i = 0; i < n; ++i) {
    const int *d = desc + 4 * i;
    if (d[0] < 0 || d[0] >= m->n_orb || d[1] < 0 || d[1] >= m->n_orb || d[2] < 0 || d[2] > 2 || d[3] < 0 ||
        d[3] > 4) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "descriptor row %lld = (%d, %d, %d, %d) out of range (n_orb = %d)",
                    (long long)i, d[0], d[1], d[2], d[3], m->n_orb);
      QPX_REQUIRE(false, buf);
    }
  }
}

uint64_t sm64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI (no counterpart in the reference)
// ---------------------------------------------------------------------------
extern "C" int qpx_mlp_create(int device, int n_orb, uint64_t seed, qpx_mlp **out) {
  if (!out) return 1;
  qpx_mlp *m = new qpx_mlp();
  int rc = mguard(m, [&] {
    QPX_REQUIRE(n_orb > 0 && n_orb <= 256, "n_orb must be in 1..256");
    QPX_HIP(hipSetDevice(device));
    m->device = device;
    m->n_orb = n_orb;
    QPX_REQUIRE(m->din == kDin && m->hid == kHid, "network shape differs from the prediction scratch layout");
    m->scratch = scratch_address(device);
    QPX_HIP(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
    const int D = m->din, Hd = m->hid;
    uint64_t s = seed;
    auto init = [&](DBuf<float> &p, size_t n, float scale) {
      std::vector<float> h(n);
      for (auto &x : h) {
        double u1 = 1.0 - (double)(sm64(s) >> 11) / 9007199254740992.0, u2 = (double)(sm64(s) >> 11) / 9007199254740992.0;
        x = (float)(scale * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
      }
      p.upload(h, m->st);
    };
    init(m->W1, (size_t)D * Hd, std::sqrt(2.0f / 75.f));
    init(m->W2, (size_t)Hd * Hd, std::sqrt(2.0f / Hd));
    init(m->W3, (size_t)Hd * Hd, std::sqrt(2.0f / Hd));
    init(m->w4, Hd, std::sqrt(1.0f / Hd));
    init(m->emb, (size_t)n_orb * 32, 0.3f);
    std::vector<float> z(Hd, 0.f);
    m->b1.upload(z, m->st);
    m->b2.upload(z, m->st);
    m->b3.upload(z, m->st);
    m->W1b.alloc((size_t)D * Hd); m->W2b.alloc((size_t)Hd * Hd); m->W3b.alloc((size_t)Hd * Hd);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(64), dim3(256), 0, m->st, m->W1.p, m->W1b.p, (int64_t)D * Hd);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(64), dim3(256), 0, m->st, m->W2.p, m->W2b.p, (int64_t)Hd * Hd);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(64), dim3(256), 0, m->st, m->W3.p, m->W3b.p, (int64_t)Hd * Hd);
    auto zero = [&](DBuf<float> &p, size_t n) {
      p.alloc(n);
      QPX_HIP(hipMemsetAsync(p.p, 0, sizeof(float) * n, m->st));
    };
    zero(m->gW1, (size_t)D * Hd); zero(m->gW2, (size_t)Hd * Hd); zero(m->gW3, (size_t)Hd * Hd);
    zero(m->gb1, Hd); zero(m->gb2, Hd); zero(m->gb3, Hd); zero(m->gw4, Hd); zero(m->gb4, 1);
    zero(m->b4, 1); zero(m->mb4, 1); zero(m->vb4, 1);
    zero(m->gemb, (size_t)n_orb * 32);
    zero(m->mW1, (size_t)D * Hd); zero(m->mW2, (size_t)Hd * Hd); zero(m->mW3, (size_t)Hd * Hd);
    zero(m->vW1, (size_t)D * Hd); zero(m->vW2, (size_t)Hd * Hd); zero(m->vW3, (size_t)Hd * Hd);
    zero(m->mb1, Hd); zero(m->mb2, Hd); zero(m->mb3, Hd); zero(m->mw4, Hd);
    zero(m->vb1, Hd); zero(m->vb2, Hd); zero(m->vb3, Hd); zero(m->vw4, Hd);
    zero(m->memb, (size_t)n_orb * 32); zero(m->vemb, (size_t)n_orb * 32);
    QPX_HIP(hipGetLastError());
    QPX_HIP(hipStreamSynchronize(m->st));
  });
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return 0;
}

extern "C" int qpx_mlp_destroy(qpx_mlp *m) {
  if (!m) return 0;
  (void)hipSetDevice(m->device);
  if (m->st) {
    (void)hipStreamSynchronize(m->st);
    (void)hipStreamDestroy(m->st);
  }
  delete m;
  return 0;
}

extern "C" int qpx_mlp_last_error(qpx_mlp *m, char *buf, int len) {
  if (!m || !buf || len <= 0) return 1;
  std::strncpy(buf, m->err.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return 0;
}

// One Adam step on a batch.  desc (n x 4 int32: h1, h2 0-based, spin class,
// degree), fdesc (n x 3 float: eps1, eps2, log|c|), target (n) = label.
// lr = 0 computes loss and gradients only.  *loss = mean squared error.
extern "C" int qpx_mlp_train_step(qpx_mlp *m, int n, const int *desc, const float *fdesc, const float *target,
                                  float lr, float *loss) {
  return mguard(m, [&] {
    QPX_HIP(hipSetDevice(m->device));
    QPX_REQUIRE(n > 0, "empty batch");
    QPX_REQUIRE(fdesc && target, "null argument");
    check_desc(m, n, desc);
    ensure(m, n);
    QPX_HIP(hipMemcpyAsync(m->desc.p, desc, sizeof(int) * 4 * n, hipMemcpyHostToDevice, m->st));
    QPX_HIP(hipMemcpyAsync(m->fdesc.p, fdesc, sizeof(float) * 3 * n, hipMemcpyHostToDevice, m->st));
    QPX_HIP(hipMemcpyAsync(m->tgt.p, target, sizeof(float) * n, hipMemcpyHostToDevice, m->st));
    float l = train_step(m, n, lr);
    if (loss) *loss = l;
  });
}

extern "C" int qpx_mlp_train_epochs(qpx_mlp *m, int n, const int *desc, const float *fdesc, const float *target,
                                    float lr, int epochs, int batch, uint64_t seed, float *epoch_loss) {
  return mguard(m, [&] {
    QPX_HIP(hipSetDevice(m->device));
    QPX_REQUIRE(n > 0 && batch >= 1 && batch <= n && epochs >= 1, "need 1 <= batch <= n and epochs >= 1");
    QPX_REQUIRE(desc && fdesc && target && epoch_loss, "null argument");
    check_desc(m, n, desc);
    ensure(m, batch);
    DBuf<int> desc_all, perm_d;
    DBuf<float> fdesc_all, tgt_all;
    desc_all.upload(desc, (size_t)4 * n, m->st);
    fdesc_all.upload(fdesc, (size_t)3 * n, m->st);
    tgt_all.upload(target, (size_t)n, m->st);
    perm_d.alloc(n);
    const int steps = n / batch;
    DBuf<float> step_loss;   // one loss per step of an epoch, read back once per epoch
    step_loss.alloc(steps);
    std::vector<float> sl(steps);
    std::vector<int> perm(n);
    uint64_t s = seed;
    for (int ep = 0; ep < epochs; ++ep) {
      for (int i = 0; i < n; ++i) perm[i] = i;
      for (int i = n - 1; i > 0; --i) {   // Fisher-Yates
        const int j = (int)(sm64(s) % (uint64_t)(i + 1));
        std::swap(perm[i], perm[j]);
      }
      QPX_HIP(hipMemcpyAsync(perm_d.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, m->st));
      for (int q = 0; q < steps; ++q) {
        hipLaunchKernelGGL(gather_batch_kernel, dim3((batch + 255) / 256), dim3(256), 0, m->st,
                           perm_d.p + (size_t)q * batch, batch, desc_all.p, fdesc_all.p, tgt_all.p, m->desc.p,
                           m->fdesc.p, m->tgt.p);
        QPX_HIP(hipGetLastError());
        train_step_async(m, batch, lr, step_loss.p + q);
      }
      QPX_HIP(hipMemcpyAsync(sl.data(), step_loss.p, sizeof(float) * steps, hipMemcpyDeviceToHost, m->st));
      QPX_HIP(hipStreamSynchronize(m->st));
      double lsum = 0.0;
      for (float x : sl) lsum += x;
      epoch_loss[ep] = (float)(lsum / steps);
    }
  });
}

// Predictions for n blocks (processed in chunks of up to 65536).
extern "C" int qpx_mlp_predict(qpx_mlp *m, int n, const int *desc, const float *fdesc, float *out) {
  return mguard(m, [&] {
    QPX_HIP(hipSetDevice(m->device));
    QPX_REQUIRE(n >= 0 && (n == 0 || (fdesc && out)), "bad prediction arguments");
    check_desc(m, n, desc);
    const int chunk = 65536;
    for (int i0 = 0; i0 < n; i0 += chunk) {
      int c = std::min(chunk, n - i0);
      ensure(m, c);
      QPX_HIP(hipMemcpyAsync(m->desc.p, desc + 4 * (size_t)i0, sizeof(int) * 4 * c, hipMemcpyHostToDevice, m->st));
      QPX_HIP(hipMemcpyAsync(m->fdesc.p, fdesc + 3 * (size_t)i0, sizeof(float) * 3 * c, hipMemcpyHostToDevice, m->st));
      forward(m, c);
      const int nb = head_blocks(c);
      hipLaunchKernelGGL(head_kernel, dim3(nb), dim3(256), 0, m->st, m->Z3.p, m->w4.p, m->b4.p, c, m->hid,
                         (const float *)nullptr, m->y.p, (float *)nullptr, (float *)nullptr);
      QPX_HIP(hipGetLastError());
      QPX_HIP(hipMemcpyAsync(out + i0, m->y.p, sizeof(float) * c, hipMemcpyDeviceToHost, m->st));
    }
    QPX_HIP(hipStreamSynchronize(m->st));
  });
}

// Prediction without any allocation: the descriptors are staged, and the
// activations kept, in this device's file-scope scratch (kStaticRows rows at
// most).  Every call on a GPU shares that scratch: calls that overlap on one
// GPU (from several threads or networks) race.
extern "C" int qpx_mlp_predict_static(qpx_mlp *m, int n, const int *desc, const float *fdesc, float *out) {
  return mguard(m, [&] {
    QPX_REQUIRE(n >= 1 && n <= kStaticRows, "n must be in 1..2048 (the rows the static scratch holds)");
    QPX_REQUIRE(fdesc && out, "null argument");
    check_desc(m, n, desc);
    QPX_HIP(hipSetDevice(m->device));
    QPX_REQUIRE(m->scratch, "no prediction scratch on this device");
    const FwdBufs b = static_bufs(m->scratch);
    QPX_HIP(hipMemcpyAsync(const_cast<int *>(b.desc), desc, sizeof(int) * 4 * n, hipMemcpyHostToDevice, m->st));
    QPX_HIP(hipMemcpyAsync(const_cast<float *>(b.fdesc), fdesc, sizeof(float) * 3 * n, hipMemcpyHostToDevice, m->st));
    forward_bufs(m, n, b);
    hipLaunchKernelGGL(head_kernel, dim3(head_blocks(n)), dim3(256), 0, m->st, b.Z3, m->w4.p, m->b4.p, n, m->hid,
                       (const float *)nullptr, b.y, (float *)nullptr, (float *)nullptr);
    QPX_HIP(hipGetLastError());
    QPX_HIP(hipMemcpyAsync(out, b.y, sizeof(float) * n, hipMemcpyDeviceToHost, m->st));
    QPX_HIP(hipStreamSynchronize(m->st));
  });
}

// Parameter exchange (fp32), for checkpoints and the gradient checks:
// which = 0 W1 (80x256), 1 W2, 2 W3 (256x256), 3 b1, 4 b2, 5 b3, 6 w4 (256),
// 7 emb (n_orb x 32), 8 b4 (1); the matching gradients with grad = 1.
extern "C" int qpx_mlp_param(qpx_mlp *m, int which, int grad, int set, float *data, int64_t n) {
  return mguard(m, [&] {
    QPX_HIP(hipSetDevice(m->device));
    DBuf<float> *P[8] = {&m->W1, &m->W2, &m->W3, &m->b1, &m->b2, &m->b3, &m->w4, &m->emb};
    DBuf<float> *G[8] = {&m->gW1, &m->gW2, &m->gW3, &m->gb1, &m->gb2, &m->gb3, &m->gw4, &m->gemb};
    if (which == 8) {
      QPX_REQUIRE(n == 1, "b4 has one element");
      if (grad) {
        QPX_HIP(hipMemcpyAsync(data, m->gb4.p, sizeof(float), hipMemcpyDeviceToHost, m->st));
        QPX_HIP(hipStreamSynchronize(m->st));
      } else if (set) {
        QPX_HIP(hipMemcpyAsync(m->b4.p, data, sizeof(float), hipMemcpyHostToDevice, m->st));
        QPX_HIP(hipStreamSynchronize(m->st));
      } else {
        QPX_HIP(hipMemcpyAsync(data, m->b4.p, sizeof(float), hipMemcpyDeviceToHost, m->st));
        QPX_HIP(hipStreamSynchronize(m->st));
      }
      return;
    }
    QPX_REQUIRE(which >= 0 && which < 8, "unknown parameter");
    DBuf<float> &b = grad ? *G[which] : *P[which];
    QPX_REQUIRE((size_t)n == b.n, "parameter size mismatch");
    if (set) {
      QPX_REQUIRE(!grad, "gradients are read-only");
      QPX_HIP(hipMemcpyAsync(b.p, data, sizeof(float) * n, hipMemcpyHostToDevice, m->st));
      uint16_t *bb = which == 0 ? m->W1b.p : which == 1 ? m->W2b.p : which == 2 ? m->W3b.p : nullptr;
      if (bb) hipLaunchKernelGGL(to_bf16_kernel, dim3(64), dim3(256), 0, m->st, b.p, bb, (int64_t)n);
    } else {
      QPX_HIP(hipMemcpyAsync(data, b.p, sizeof(float) * n, hipMemcpyDeviceToHost, m->st));
    }
    QPX_HIP(hipStreamSynchronize(m->st));
  });
}

// Plain bf16 GEMM C = A B (row-major A n x k, B k x m, fp32 inputs rounded
// to bf16), for the lane-layout and accuracy checks of the MFMA tiles.
// Test entry: C = A B through the network's GEMM in every operand layout it
// is called with.  ta = 0: A given (M x K) row-major; ta = 1: A given as its
// transpose, (K x M) row-major.  tb = 0: B (K x N) row-major; tb = 1: (N x K).
// split = 1 passes a workspace, so the K split runs where it would for a
// weight gradient.
extern "C" int qpx_debug_mlp_gemm(qpx_mlp *m, int M, int N, int K, const float *A, const float *B, int ta, int tb,
                                  int split, float *C) {
  if (M < 1 || N < 1 || K < 1 || !A || !B || !C) return 1;
  return mguard(m, [&] {
    QPX_HIP(hipSetDevice(m->device));
    DBuf<float> a, b, c, ws;
    DBuf<uint16_t> ab, bb;
    a.upload(A, (size_t)M * K, m->st);
    b.upload(B, (size_t)K * N, m->st);
    ab.alloc((size_t)M * K);
    bb.alloc((size_t)K * N);
    c.alloc((size_t)M * N);
    if (split) ws.alloc((size_t)32 * M * N);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(64), dim3(256), 0, m->st, a.p, ab.p, (int64_t)M * K);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(64), dim3(256), 0, m->st, b.p, bb.p, (int64_t)K * N);
    const int64_t ars = ta ? 1 : K, acs = ta ? M : 1;
    const int64_t brs = tb ? 1 : N, bcs = tb ? K : 1;
    gemm(m->st, ab.p, ars, acs, bb.p, brs, bcs, c.p, N, M, N, K, nullptr, 0, nullptr, 0.f, split ? &ws : nullptr);
    QPX_HIP(hipMemcpyAsync(C, c.p, sizeof(float) * M * N, hipMemcpyDeviceToHost, m->st));
    QPX_HIP(hipStreamSynchronize(m->st));
  });
}

extern "C" int qpx_mlp_gemm_check(qpx_mlp *m, int M, int N, int K, const float *A, const float *B, float *C) {
  return mguard(m, [&] {
    QPX_HIP(hipSetDevice(m->device));
    DBuf<float> a, b, c;
    DBuf<uint16_t> ab, bb;
    a.upload(A, (size_t)M * K, m->st);
    b.upload(B, (size_t)K * N, m->st);
    ab.alloc((size_t)M * K);
    bb.alloc((size_t)K * N);
    c.alloc((size_t)M * N);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(64), dim3(256), 0, m->st, a.p, ab.p, (int64_t)M * K);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(64), dim3(256), 0, m->st, b.p, bb.p, (int64_t)K * N);
    gemm(m->st, ab.p, K, 1, bb.p, N, 1, c.p, N, M, N, K, nullptr, 0, nullptr, 0.f);
    QPX_HIP(hipMemcpyAsync(C, c.p, sizeof(float) * M * N, hipMemcpyDeviceToHost, m->st));
    QPX_HIP(hipStreamSynchronize(m->st));
  });
}
