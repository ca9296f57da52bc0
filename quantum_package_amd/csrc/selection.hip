// selection.hip -- CIPSI selection with Epstein-Nesbet PT2 on the GPU.
//
// Reference: plugins/Full_CI_ZMQ/selection.irp.f (select_connected :50-79,
// select_singles_and_doubles :258-578, fill_buffer_double :582-644,
// splash_pq :647-735, get_d2/d1/d0 :738-1085, past_d1/d2 :1088-1126,
// spot_isinwf :1130-1171), selection_buffer.irp.f:2-138 and
// zmq_selection.irp.f:1-164.
//
// Work decomposition (one GPU replaces the ZeroMQ task farm):
//   * per batch of generators, a near-list kernel scans psi once per
//     generator and keeps, in index order, the selectors within 4
//     excitations and the non-selectors within 2 (the reference's
//     preinteresting / prefullinteresting lists);
//   * one workgroup owns one (generator, h1, h2) block -- the reference's
//     maskInd unit.  The block's coupling matrix mat(N_states, M, M) lives
//     in LDS (in an HBM slot per workgroup when it leaves no room for a
//     queue); with the rows split into two bands, cell (i,j) is owned by
//     thread 256 * band(i) + (37 (i-1) + (j-1)) mod 256, so every row and
//     every column of cells maps to distinct threads;
//   * the block's selectors are filtered in parallel into an LDS queue in
//     near-list order, with their determinants.  At a flush, the <= 6
//     couplings of each four-particle selector (get_d2) are computed into
//     its queue entry, one entry per thread; the replay then turns them
//     into per-owner record lists (count, scan, scatter, sort);
//     each thread then walks the get_d1 / get_d0 entries in queue order and
//     applies its earlier records before each one, so every cell receives
//     its terms in selector order, the same order as the reference
//     (bit-identical FP64 sums);
//   * bans (spot_isinwf, past_d1/d2, symmetrisation) are order-independent
//     sets and are built in parallel with LDS atomics before any coupling is
//     accumulated;
//   * candidates with min-state e <= tau go to a global pool; tau tightens
//     between batches to the n_keep-th best key (the reference's buffer
//     threshold `mini`).  Blocks that hit a full pool are re-run after the
//     pool is compacted.  Per-block PT2 is written to its own slot and summed
//     in block order on the host.
// Floating-point expressions keep the reference's operation order and this
// file is compiled without FMA contraction.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <utility>
#include <vector>
#include "qpx_internal.h"
#include "qpx_device.h"

using namespace qpx;

namespace {

constexpr int SB = 512;             // threads per block workgroup
constexpr int OW = 256;             // owners per band of rows
constexpr int NBAND = SB / OW;      // bands of rows of the coupling matrix
constexpr int MAXNS = 4;

// ---------------------------------------------------------------- helpers
template <int NI>
struct Det {
  uint64_t w[2 * NI];
};

__device__ inline bool bit(const uint64_t *s, int o) {  // o 1-based
  return (s[(o - 1) >> 6] >> ((o - 1) & 63)) & 1;
}

// Determinant-shaped arrays (2 * NI words, spin 1 then spin 2) that live in
// registers are only ever indexed with compile-time word numbers: a single
// run-time index into such an array places all of it in scratch memory.
// These helpers pick the word with selects instead.
template <int NI>
__device__ __forceinline__ uint64_t spin_word(const uint64_t *a, int s, int w) {   // w a constant
  uint64_t x0 = a[w], x1 = a[NI + w];
  // the empty asm keeps the selection on the two values: folded into one
  // load from a selected address, it would be a run-time index again
  asm("" : "+v"(x0), "+v"(x1));
  return s == 1 ? x0 : x1;
}
// the word of spin s (1-based) holding orbital o (1-based)
template <int NI>
__device__ __forceinline__ uint64_t orb_word(const uint64_t *a, int s, int o) {
  const int wo = (o - 1) >> 6;
  uint64_t x = spin_word<NI>(a, s, 0);
#pragma unroll
  for (int w = 1; w < NI; ++w)
    if (wo == w) x = spin_word<NI>(a, s, w);
  return x;
}
template <int NI>
__device__ __forceinline__ void set_orb(uint64_t *a, int s, int o) {
  const int idx = (s - 1) * NI + ((o - 1) >> 6);
  const uint64_t b = 1ULL << ((o - 1) & 63);
#pragma unroll
  for (int w = 0; w < 2 * NI; ++w) a[w] |= b & (0ULL - (uint64_t)(w == idx));   // branch-free: no merged store
}
template <int NI>
__device__ __forceinline__ void clear_orb(uint64_t *a, int s, int o) {
  const int idx = (s - 1) * NI + ((o - 1) >> 6);
  const uint64_t b = 1ULL << ((o - 1) & 63);
#pragma unroll
  for (int w = 0; w < 2 * NI; ++w) a[w] &= ~(b & (0ULL - (uint64_t)(w == idx)));
}

// parity of the occupied orbitals with index <= o in spin s of det
// (get_mask_phase)
template <int NI>
__device__ inline int pm(const uint64_t *det, int s, int o) {
  int c = 0;
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    const int lo = 64 * w;
    const uint64_t x = spin_word<NI>(det, s, w);
    if (o - 1 >= lo + 63) c += __popcll(x);
    else if (o - 1 >= lo) c += __popcll(x & ((2ULL << ((o - 1) - lo)) - 1));
  }
  return c & 1;
}

template <int NI>
__device__ inline double phase_bi(const uint64_t *sel, int s1, int s2, int h1, int p1, int h2, int p2) {
  int np = pm<NI>(sel, s1, h1) + pm<NI>(sel, s1, p1) + pm<NI>(sel, s2, h2) + pm<NI>(sel, s2, p2);
  if (p1 < h1) np++;
  if (p2 < h2) np++;
  if (s1 == s2 && max(h1, p1) > min(h2, p2)) np++;
  return (np & 1) ? -1.0 : 1.0;
}

// Prefix-parity masks of a selector (the reference's phasemask,
// get_mask_phase, selection.irp.f:24-47): bit o-1 of spin s's words is the
// parity of the occupied orbitals <= o, so pm(s, o) is one bit test.
template <int NI>
__device__ inline void prefix_parity(const uint64_t *sel, uint64_t *pmk) {
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    uint64_t carry = 0;   // all ones when the lower words hold an odd count
#pragma unroll
    for (int w = 0; w < NI; ++w) {
      const uint64_t x = sel[s * NI + w];
      uint64_t y = x;
      y ^= y << 1;
      y ^= y << 2;
      y ^= y << 4;
      y ^= y << 8;
      y ^= y << 16;
      y ^= y << 32;
      pmk[s * NI + w] = y ^ carry;
      carry ^= (uint64_t)0 - (uint64_t)(__popcll(x) & 1);
    }
  }
}

template <int NI>
__device__ inline int pm_bit(const uint64_t *pmk, int s, int o) { return (int)((orb_word<NI>(pmk, s, o) >> ((o - 1) & 63)) & 1); }

// phase_bi from the prefix-parity masks: 1 when the phase is -1
template <int NI>
__device__ inline uint32_t odd_pm(const uint64_t *pmk, int s1, int s2, int h1, int p1, int h2, int p2) {
  int np = pm_bit<NI>(pmk, s1, h1) + pm_bit<NI>(pmk, s1, p1) + pm_bit<NI>(pmk, s2, h2) + pm_bit<NI>(pmk, s2, p2);
  if (p1 < h1) np++;
  if (p2 < h2) np++;
  if (s1 == s2 && max(h1, p1) > min(h2, p2)) np++;
  return (uint32_t)(np & 1);
}

struct Tabs {
  const double *V, *hcore, *J, *JA, *fock_cs;
  int M, npair;
  double ref_energy;
  // M^4 table, G4[((a*M + b)*M + y)*M + x] = (ax|by) (0-based), or null
  // above its orbital limit: the lines of cells get_d1 / get_d0 fill read
  // it along their varying orbital, so a wavefront's lanes read neighbouring
  // addresses (the values are those of V)
  const double *G4;
};

// Offset of <ij|kl> = (ik|jl) (1-based orbitals) in the table the line
// kernels read: in G4 with argument `vary` (0..3 = i, j, k, l) in the
// contiguous slot, by the 8-fold symmetry (ik|jl) = (ki|jl) = (jl|ik) =
// (lj|ik); in V (pair matrix) when there is no G4.
// 32-bit offsets: npair^2 < 2^31 for M <= 256, and the M^4 table stops at
// kG4MaxOrb (160^4 < 2^30); every factor stays below 2^24
__device__ inline uint32_t pair32(int p, int q) {
  const int lo = p < q ? p : q, hi = p < q ? q : p;
  return (uint32_t)(hi * (hi + 1) / 2 + lo);
}
__device__ inline uint32_t line_off(const Tabs &T, int i, int j, int k, int l, int vary) {
  if (!T.G4) return pair32(i - 1, k - 1) * (uint32_t)T.npair + pair32(j - 1, l - 1);
  int a, x, b, y;   // (a x | b y)
  switch (vary) {
    case 0: a = k; x = i; b = j; y = l; break;
    case 1: a = l; x = j; b = i; y = k; break;
    case 2: a = i; x = k; b = j; y = l; break;
    default: a = j; x = l; b = i; y = k; break;
  }
  const uint32_t M = (uint32_t)T.M;
  return (((uint32_t)(a - 1) * M + (uint32_t)(b - 1)) * M + (uint32_t)(y - 1)) * M + (uint32_t)(x - 1);
}

__device__ inline double I4(const Tabs &T, int i, int j, int k, int l) {  // <ij|kl>, 1-based
  return T.V[pair_index(i - 1, k - 1) * T.npair + pair_index(j - 1, l - 1)];
}

// holes/particles of spin s between determinants a (initial) and b (final),
// 1-based, ascending (the first two of each)
struct SpinExc {
  int nh, np, h0, h1, p0, p1;
};
template <int NI>
__device__ inline SpinExc spin_exc(const uint64_t *a, const uint64_t *b, int s) {
  SpinExc e{0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    const uint64_t aw = spin_word<NI>(a, s, w), bw = spin_word<NI>(b, s, w);
    const uint64_t x = aw ^ bw;
    uint64_t pp = x & bw, hh = x & aw;
    while (pp) {
      const int o = 64 * w + __ffsll((unsigned long long)pp);
      pp &= pp - 1;
      if (e.np == 0) e.p0 = o;
      else if (e.np == 1) e.p1 = o;
      e.np++;
    }
    while (hh) {
      const int o = 64 * w + __ffsll((unsigned long long)hh);
      hh &= hh - 1;
      if (e.nh == 0) e.h0 = o;
      else if (e.nh == 1) e.h1 = o;
      e.nh++;
    }
  }
  return e;
}

// occupied orbitals of spin s of det strictly between h and p (1-based)
template <int NI>
__device__ inline int count_between1(const uint64_t *det, int s, int h, int p) {
  int lo = min(h, p), hi = max(h, p);
  int c = 0;
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    int b0 = lo - 64 * w, e = hi - 1 - 64 * w;  // 0-based bit positions lo..hi-2 <=> orbitals lo+1..hi-1
    if (b0 < 0) b0 = 0;
    if (e > 64) e = 64;
    if (b0 < e) {
      uint64_t m = ((e == 64) ? ~0ULL : ((1ULL << e) - 1)) & ~((1ULL << b0) - 1);
      c += __popcll(spin_word<NI>(det, s, w) & m);
    }
  }
  return c;
}

template <int NI>
__device__ inline int spin_nperm(const uint64_t *det, int s, const SpinExc &e) {
  int n = count_between1<NI>(det, s, e.h0, e.p0);
  if (e.nh == 2) {
    n += count_between1<NI>(det, s, e.h1, e.p1);
    int A = min(e.h0, e.p0), B = max(e.h0, e.p0), Cc = min(e.h1, e.p1), D = max(e.h1, e.p1);
    if (A < Cc && Cc < B && B < D) n++;
  }
  return n;
}

// diag_H_mat_elem (slater_rules.irp.f:1869-2022): incremental from the HF
// reference with a/ac operators, in the reference's order
template <int NI>
__device__ double diag_H_incremental(const Tabs &T, const uint64_t *det, int na_e, int nb_e) {
  // every array below is indexed with constants only (the spin loop and the
  // word loops are unrolled), so it stays in registers
  const int M = T.M;
  uint64_t ref[2 * NI];
#pragma unroll
  for (int w = 0; w < NI; ++w) {
    const int na = na_e - 64 * w, nb = nb_e - 64 * w;
    ref[w] = na <= 0 ? 0ULL : (na >= 64 ? ~0ULL : ((1ULL << na) - 1));
    ref[NI + w] = nb <= 0 ? 0ULL : (nb >= 64 ? ~0ULL : ((1ULL << nb) - 1));
  }
  uint64_t part[2 * NI], hole[2 * NI];
  int nexc0 = 0, nexc1 = 0;
#pragma unroll
  for (int w = 0; w < 2 * NI; ++w) {
    uint64_t x = det[w] ^ ref[w];
    part[w] = x & det[w];
    hole[w] = x & ref[w];
    if (w < NI) nexc0 += __popcll(hole[w]);
    else nexc1 += __popcll(hole[w]);
  }
  double e = T.ref_energy;
  if (nexc0 + nexc1 == 0) return e;
  uint64_t key[2 * NI];
#pragma unroll
  for (int w = 0; w < 2 * NI; ++w) key[w] = ref[w];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    int na = s == 0 ? na_e : nb_e;
    const int nb = s == 0 ? nb_e : na_e;
    const int nexc = s == 0 ? nexc0 : nexc1;
    // i-th particle / hole of this spin, ascending
    uint64_t pw[NI], hw[NI];
#pragma unroll
    for (int w = 0; w < NI; ++w) { pw[w] = part[s * NI + w]; hw[w] = hole[s * NI + w]; }
    for (int i = 0; i < nexc; ++i) {
      int po = 0, ho = 0;
#pragma unroll
      for (int w = 0; w < NI; ++w)
        if (!po && pw[w]) { po = 64 * w + __ffsll((unsigned long long)pw[w]); pw[w] &= pw[w] - 1; }
#pragma unroll
      for (int w = 0; w < NI; ++w)
        if (!ho && hw[w]) { ho = 64 * w + __ffsll((unsigned long long)hw[w]); hw[w] &= hw[w] - 1; }
      // ac_operator(po): loops over the current occupations before adding po
      {
        e += T.hcore[(po - 1) * (M + 1)];
        int cnt = 0;
#pragma unroll
        for (int w = 0; w < NI; ++w) {
          uint64_t y = key[s * NI + w];
          while (y && cnt < na) {
            int o = 64 * w + __ffsll((unsigned long long)y);
            y &= y - 1;
            e += T.JA[(o - 1) + M * (po - 1)];
            ++cnt;
          }
        }
        cnt = 0;
#pragma unroll
        for (int w = 0; w < NI; ++w) {
          uint64_t y = key[(1 - s) * NI + w];
          while (y && cnt < nb) {
            int o = 64 * w + __ffsll((unsigned long long)y);
            y &= y - 1;
            e += T.J[(o - 1) + M * (po - 1)];
            ++cnt;
          }
        }
        set_orb<NI>(key, s + 1, po);
        na += 1;
      }
      // a_operator(ho): removes ho, then loops over the remaining occupations
      {
        clear_orb<NI>(key, s + 1, ho);
        na -= 1;
        e -= T.hcore[(ho - 1) * (M + 1)];
        int cnt = 0;
#pragma unroll
        for (int w = 0; w < NI; ++w) {
          uint64_t y = key[s * NI + w];
          while (y && cnt < na) {
            int o = 64 * w + __ffsll((unsigned long long)y);
            y &= y - 1;
            e -= T.JA[(o - 1) + M * (ho - 1)];
            ++cnt;
          }
        }
        cnt = 0;
#pragma unroll
        for (int w = 0; w < NI; ++w) {
          uint64_t y = key[(1 - s) * NI + w];
          while (y && cnt < nb) {
            int o = 64 * w + __ffsll((unsigned long long)y);
            y &= y - 1;
            e -= T.J[(o - 1) + M * (ho - 1)];
            ++cnt;
          }
        }
      }
    }
  }
  return e;
}

// i_H_j (slater_rules.irp.f:478-571) with the Fock-based single excitation
// (mono_excitations.irp.f:97-155); na_e/nb_e electron numbers
template <int NI>
__device__ double i_H_j(const Tabs &T, const uint64_t *ki, const uint64_t *kj, int na_e, int nb_e) {
  int degree = 0;
#pragma unroll
  for (int w = 0; w < 2 * NI; ++w) degree += __popcll(ki[w] ^ kj[w]);
  degree >>= 1;
  if (degree > 2) return 0.0;
  const SpinExc ea = spin_exc<NI>(ki, kj, 1), eb = spin_exc<NI>(ki, kj, 2);
  if (degree == 2) {
    int nperm = 0;
    if (ea.nh) nperm += spin_nperm<NI>(ki, 1, ea);
    if (ea.nh != 2 && eb.nh) nperm += spin_nperm<NI>(ki, 2, eb);
    double phase = (nperm & 1) ? -1.0 : 1.0;
    if (ea.nh == 1) {
      return phase * I4(T, ea.h0, eb.h0, ea.p0, eb.p0);
    }
    const bool sa = ea.nh == 2;
    const int h0 = sa ? ea.h0 : eb.h0, h1 = sa ? ea.h1 : eb.h1, p0 = sa ? ea.p0 : eb.p0, p1 = sa ? ea.p1 : eb.p1;
    return phase * (I4(T, h0, h1, p0, p1) - I4(T, h0, h1, p1, p0));
  }
  if (degree == 1) {
    const int s = (ea.nh == 1) ? 1 : 2;
    const SpinExc e = (s == 1) ? ea : eb;
    double phase = (spin_nperm<NI>(ki, s, e) & 1) ? -1.0 : 1.0;
    // get_mono_excitation_from_fock(ki, kj, h:=particle, p:=hole, spin)
    const int h = e.p0, p = e.h0;
    const int M = T.M;
    const int64_t Php = pair_index(h - 1, p - 1);
    double hij = T.fock_cs[(h - 1) + M * (p - 1)];
    // reference closed shell: orbitals 1..nb_e in both spins
    uint64_t refcs[NI];
#pragma unroll
    for (int w = 0; w < NI; ++w) {
      const int n = nb_e - 64 * w;
      refcs[w] = n <= 0 ? 0ULL : (n >= 64 ? ~0ULL : ((1ULL << n) - 1));
    }
    auto coul = [&](int i) { return T.V[Php * T.npair + pair_index(i - 1, i - 1)]; };
    auto exch = [&](int i) { return T.V[pair_index(h - 1, i - 1) * T.npair + pair_index(i - 1, p - 1)]; };
    // hij -/+= f(o) over the set bits o of y in ascending order; the
    // integrals of four orbitals at a time are loaded together, the sum
    // keeps its order
    auto acc = [&](uint64_t y, int w, bool sub, auto f) {
      while (y) {
        int o[4];
        double v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          o[t] = 0;
          if (y) {
            o[t] = 64 * w + __ffsll((unsigned long long)y);
            y &= y - 1;
          }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) v[t] = o[t] ? f(o[t]) : 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
          if (o[t]) hij = sub ? hij - v[t] : hij + v[t];
      }
    };
#pragma unroll
    for (int ss = 0; ss < 2; ++ss)
#pragma unroll
      for (int w = 0; w < NI; ++w) acc((ki[ss * NI + w] ^ refcs[w]) & refcs[w], w, true, coul);
#pragma unroll
    for (int w = 0; w < NI; ++w) acc((spin_word<NI>(ki, s, w) ^ refcs[w]) & refcs[w], w, false, exch);
#pragma unroll
    for (int ss = 0; ss < 2; ++ss)
#pragma unroll
      for (int w = 0; w < NI; ++w) acc((ki[ss * NI + w] ^ refcs[w]) & ki[ss * NI + w], w, false, coul);
#pragma unroll
    for (int w = 0; w < NI; ++w) {
      const uint64_t kw = spin_word<NI>(ki, s, w);
      acc((kw ^ refcs[w]) & kw, w, true, exch);
    }
    return hij * phase;
  }
  return diag_H_incremental<NI>(T, ki, na_e, nb_e);
}

// ---------------------------------------------------------------- near lists
struct NearArgs {
  const uint64_t *psi;
  int64_t n_det;
  int n_sel;
  const int *gens;  // generator indices (1-based) of this batch
  int n_gens;
  const int64_t *off;  // fill offsets (per generator)
  int *cnt;            // counts (per generator)
  int *list;           // output det indices (1-based)
  int *past;           // per generator: kept entries with index < the generator
  uint64_t *sig;       // output hole signatures (hole_sig), parallel to list
};

// The holes of d relative to the generator (orbitals the generator occupies
// and d does not; at most 4 for a near-list entry), spin 1 first, ascending:
// 16-bit fields (s - 1) << 8 | (o - 1), unused fields 0xffff.  A block
// (h1, h2) then tests an entry without its determinant: the entry has
// e + [h1 not a hole] + [h2 not a hole] particles outside the block's mask,
// e its number of holes.
template <int NI>
__device__ inline uint64_t hole_sig(const uint64_t *gen, const uint64_t *d) {
  uint64_t sig = ~0ULL;
  int n = 0;
#pragma unroll
  for (int s = 1; s <= 2; ++s)
#pragma unroll
    for (int w = 0; w < NI; ++w) {
      uint64_t hh = spin_word<NI>(gen, s, w) & ~spin_word<NI>(d, s, w);
      while (hh) {
        const int o = 64 * w + __ffsll((unsigned long long)hh) - 1;
        hh &= hh - 1;
        if (n < 4) sig = (sig & ~(0xffffULL << (16 * n))) | ((uint64_t)(((s - 1) << 8) | o) << (16 * n));
        ++n;
      }
    }
  return sig;
}
__device__ inline uint32_t hole_code(int s, int o) { return (uint32_t)(((s - 1) << 8) | (o - 1)); }
// particles outside the mask of block (c1, c2) of an entry with signature sig
__device__ inline int sig_nt(uint64_t sig, uint32_t c1, uint32_t c2) {
  int e = 0;
  bool m1 = false, m2 = false;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const uint32_t f = (uint32_t)(sig >> (16 * t)) & 0xffffu;
    if (f != 0xffffu) {
      ++e;
      m1 = m1 || f == c1;
      m2 = m2 || f == c2;
    }
  }
  return e + (m1 ? 0 : 1) + (m2 ? 0 : 1);
}
// every hole of the entry is h1 or h2: it contains the block's mask
__device__ inline bool sig_in_mask(uint64_t sig, uint32_t c1, uint32_t c2) {
  bool ok = true;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const uint32_t f = (uint32_t)(sig >> (16 * t)) & 0xffffu;
    ok = ok && (f == 0xffffu || f == c1 || f == c2);
  }
  return ok;
}

template <int NI>
__global__ __launch_bounds__(256) void near_kernel(NearArgs A, int fill) {
  __shared__ int wcount[4], wbefore[4];
  __shared__ int base, before_gen;
  const int g = A.gens[blockIdx.x];
  const uint64_t *gen = A.psi + (int64_t)(g - 1) * 2 * NI;
  uint64_t neg[2 * NI];
#pragma unroll
  for (int w = 0; w < 2 * NI; ++w) neg[w] = ~gen[w];
  if (threadIdx.x == 0) {
    base = 0;
    before_gen = 0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int64_t out0 = fill ? A.off[blockIdx.x] : 0;
  for (int64_t t0 = 0; t0 < A.n_det; t0 += 256) {
    int64_t i = t0 + threadIdx.x;
    bool keep = false;
    if (i < A.n_det) {
      const uint64_t *d = A.psi + i * 2 * NI;
      int nt = 0;
#pragma unroll
      for (int w = 0; w < 2 * NI; ++w) nt += __popcll(neg[w] & d[w]);
      keep = (i + 1 <= A.n_sel) ? (nt <= 4) : (nt <= 2);
    }
    uint64_t m = __ballot(keep);
    const uint64_t mb = __ballot(keep && i + 1 < g);   // kept entries before the generator
    int my = __popcll(m & ((1ULL << lane) - 1));
    if (lane == 0) {
      wcount[wv] = __popcll(m);
      wbefore[wv] = __popcll(mb);
    }
    __syncthreads();
    int pre = 0, tot = 0;
    for (int x = 0; x < 4; ++x) {
      if (x < wv) pre += wcount[x];
      tot += wcount[x];
    }
    if (fill && keep) {
      A.list[out0 + base + pre + my] = (int)(i + 1);
      A.sig[out0 + base + pre + my] = hole_sig<NI>(gen, A.psi + i * 2 * NI);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      base += tot;
      for (int x = 0; x < 4; ++x) before_gen += wbefore[x];
    }
    __syncthreads();
  }
  if (!fill && threadIdx.x == 0) A.cnt[blockIdx.x] = base;
  // the kept entries before the generator form a prefix of the list
  if (!fill && threadIdx.x == 0) A.past[blockIdx.x] = before_gen;
}

// ---------------------------------------------------------------- blocks
struct BlockDesc {
  int gen;      // 1-based generator index
  int gslot;    // index of the generator inside the batch (near-list slot)
  int8_t s1, s2, unbanA, unbanB;
  int16_t h1, h2;
  int mask_ind;
};

struct SelArgs {
  const uint64_t *psi;
  const double *coef;  // (n_det, NS) column-major
  int64_t n_det;
  int n_sel, NS, na, nb;
  double E0[MAXNS];
  const BlockDesc *blocks;
  int n_blocks;
  const int64_t *near_off;
  const int *near_cnt;
  const int *near_list;
  // pool
  double *pool_key;
  uint64_t *pool_det;
  int *pool_blk;
  unsigned long long *pool_count;
  int64_t pool_cap;
  double tau;
  int blk_base;   // global id of blocks[0]
  int qcap;       // entries of the LDS selector queue
  // outputs per block
  double *blk_pt2;     // n_blocks * NS
  int *blk_status;     // 0 done, 1 skipped (owned by an earlier generator), 2 overflow
  int64_t *blk_ncand;
  int64_t *blk_ngath;
  int64_t *blk_ticks;  // duration of the workgroup in 100 MHz wall-clock ticks (cost model of the proposal)
  double *mat_g;       // coupling matrices in HBM (one NS*M*M slot per workgroup), when they do not fit LDS
  const double *fockg; // fock_diag_tmp of each generator of the batch, 2*(M+1) per near-list slot
  const int *near_past;   // per near-list slot: entries with index < the generator (a prefix of the list)
  const uint64_t *near_sig;   // hole signature of each near-list entry
};

// fock_diag_tmp (Fock_diag.irp.f:1-93) of each generator of a batch: the
// diagonal Fock energies of both spins (slot M+1 holds E_gen), as
// F_[(s-1) + 2*(i-1)], one workgroup per generator.  Every block of the
// generator copies it.
template <int NI>
__global__ __launch_bounds__(256) void fock_gen_kernel(const uint64_t *__restrict__ psi, const int *__restrict__ gens,
                                                      Tabs T, double *__restrict__ fockg) {
  const int M = T.M, tid = threadIdx.x;
  const uint64_t *gen = psi + (int64_t)(gens[blockIdx.x] - 1) * 2 * NI;
  double *F_ = fockg + (size_t)blockIdx.x * 2 * (M + 1);
    const uint64_t *ga = gen, *gb = gen + NI;
#define F(s, i) F_[((s) - 1) + 2 * ((i) - 1)]
#define JA(a, b) T.JA[((a) - 1) + M * ((b) - 1)]
#define JJ(a, b) T.J[((a) - 1) + M * ((b) - 1)]
    for (int x = tid; x < 2 * M; x += blockDim.x) {
      int s = x / M + 1, i = x % M + 1;
      const uint64_t *same = s == 1 ? ga : gb, *oth = s == 1 ? gb : ga;
      double f = 0.0;
      if (bit(same, i)) {
        f += T.hcore[(i - 1) * (M + 1)];
        for (int w = 0; w < NI; ++w) {
          uint64_t y = same[w];
          while (y) { int j = 64 * w + __ffsll((unsigned long long)y); y &= y - 1; if (j != i) f += JA(i, j); }
        }
        for (int w = 0; w < NI; ++w) {
          uint64_t y = oth[w];
          while (y) { int j = 64 * w + __ffsll((unsigned long long)y); y &= y - 1; f += JJ(i, j); }
        }
      }
      if (f == 0.0) {
        f += T.hcore[(i - 1) * (M + 1)];
        for (int w = 0; w < NI; ++w) {
          uint64_t y = same[w];
          while (y) { int j = 64 * w + __ffsll((unsigned long long)y); y &= y - 1; f += JA(i, j); }
        }
        for (int w = 0; w < NI; ++w) {
          uint64_t y = oth[w];
          while (y) { int j = 64 * w + __ffsll((unsigned long long)y); y &= y - 1; f += JJ(i, j); }
        }
      }
      F(s, i) = f;
    }
    if (tid == 0) {
      double E0 = 0.0;
      for (int w = 0; w < NI; ++w) {
        uint64_t y = ga[w];
        while (y) {
          int i = 64 * w + __ffsll((unsigned long long)y);
          y &= y - 1;
          E0 += T.hcore[(i - 1) * (M + 1)];
          for (int w2 = 0; w2 < NI; ++w2) {
            uint64_t z = ga[w2];
            while (z) { int j = 64 * w2 + __ffsll((unsigned long long)z); z &= z - 1; if (j != i) E0 += 0.5 * JA(i, j); }
          }
          for (int w2 = 0; w2 < NI; ++w2) {
            uint64_t z = gb[w2];
            while (z) { int j = 64 * w2 + __ffsll((unsigned long long)z); z &= z - 1; E0 += JJ(i, j); }
          }
        }
      }
      for (int w = 0; w < NI; ++w) {
        uint64_t y = gb[w];
        while (y) {
          int i = 64 * w + __ffsll((unsigned long long)y);
          y &= y - 1;
          E0 += T.hcore[(i - 1) * (M + 1)];
          for (int w2 = 0; w2 < NI; ++w2) {
            uint64_t z = gb[w2];
            while (z) { int j = 64 * w2 + __ffsll((unsigned long long)z); z &= z - 1; if (j != i) E0 += 0.5 * JA(i, j); }
          }
        }
      }
      F(1, M + 1) = E0;
      F(2, M + 1) = E0;
    }
#undef F
#undef JA
#undef JJ
  }

template <int NI>
__device__ inline void build_mask(const uint64_t *gen, int s1, int h1, int s2, int h2, uint64_t *mask) {
#pragma unroll
  for (int w = 0; w < 2 * NI; ++w) mask[w] = gen[w];
  clear_orb<NI>(mask, s1, h1);
  clear_orb<NI>(mask, s2, h2);
}

// LDS state of one block
struct BlockLds {
  uint32_t *ban1, *ban2;  // M*M bits each, cell (i,j) -> bit (i-1)+M*(j-1)
  uint32_t *borb;         // 2 * 8 words: bannedOrb per spin (1-based orbitals)
  double *fock;           // 2*(M+1): fock_diag_tmp(s, i) at [(s-1) + 2*(i-1)]
  double *mat;            // NS*M*M
  int *flag;
};

__device__ inline bool ban_get(const uint32_t *b, int M, int i, int j) {
  int x = (i - 1) + M * (j - 1);
  return (b[x >> 5] >> (x & 31)) & 1;
}
__device__ inline void ban_set(uint32_t *b, int M, int i, int j) {
  int x = (i - 1) + M * (j - 1);
  atomicOr(&b[x >> 5], 1u << (x & 31));
}
__device__ inline bool borb_get(const uint32_t *b, int o, int s) {
  return (b[(s - 1) * 8 + ((o - 1) >> 5)] >> ((o - 1) & 31)) & 1;
}
__device__ inline void borb_set(uint32_t *b, int o, int s) {
  atomicOr(&b[(s - 1) * 8 + ((o - 1) >> 5)], 1u << ((o - 1) & 31));
}

// Ownership of the coupling-matrix cells: the M rows are split into NBAND
// bands; cell (i, j) of band b belongs to thread
//   OW * b + ((kOwnA * (i - 1) + (j - 1)) mod OW),
// kOwnA odd.  Every row and every column then has distinct owners (M <= OW,
// and kOwnA is invertible mod OW), so a thread meets at most one cell of any
// line; the odd multiplier spreads a dense patch of cells (the low virtual
// pairs most selectors land on) over the whole band instead of the few
// threads of a diagonal stripe.
constexpr int kOwnA = 37, kOwnAInv = 173;   // 37 * 173 = 1 mod 256
static_assert((kOwnA * kOwnAInv) % OW == 1, "kOwnAInv is the inverse of kOwnA mod OW");
static_assert(NBAND == 2, "band_of assumes two bands");
__device__ inline int band_of(int i, int M) { return (2 * (i - 1) >= M) ? 1 : 0; }   // floor(2 (i-1) / M)
__device__ inline int owner_of(int i, int j, int M) {
  return ((kOwnA * (i - 1) + (j - 1)) & (OW - 1)) + OW * band_of(i, M);
}
__device__ inline bool owns(int i, int j, int M) { return owner_of(i, j, M) == (int)threadIdx.x; }
// the unique j (1..M) such that thread owns (i, j); 0 if none
__device__ inline int owned_j(int i, int M) {
  if (band_of(i, M) != (int)threadIdx.x / OW) return 0;
  const int j = (((int)threadIdx.x & (OW - 1)) - kOwnA * (i - 1)) & (OW - 1);
  return (j < M) ? j + 1 : 0;
}
// the unique i (1..M) such that thread owns (i, j); 0 if none
__device__ inline int owned_i(int j, int M) {
  const int i0 = (kOwnAInv * (((int)threadIdx.x & (OW - 1)) - (j - 1))) & (OW - 1);
  if (i0 >= M) return 0;
  return (band_of(i0 + 1, M) == (int)threadIdx.x / OW) ? i0 + 1 : 0;
}
// rows [lo, hi] of this thread's band
__device__ inline void band_rows(int M, int &lo, int &hi) {
  const int b = (int)threadIdx.x / OW;
  lo = (b * M + 1) / 2 + 1;        // smallest i with band_of(i) == b
  hi = ((b + 1) * M + 1) / 2;      // largest such i
}

template <int NI>
struct Ctx {
  const Tabs &T;
  const SelArgs &A;
  BlockLds &L;
  const uint64_t *mask;
  int M, NS;
  int64_t gath;
  __device__ Ctx(const Tabs &t, const SelArgs &a, BlockLds &l, const uint64_t *m)
      : T(t), A(a), L(l), mask(m), M(t.M), NS(a.NS), gath(0) {}
  __device__ inline void add(int i, int j, const double *coefs, double hij) {
    double *c = L.mat + (int64_t)NS * ((i - 1) + M * (j - 1));
    for (int st = 0; st < NS; ++st) c[st] += coefs[st] * hij;
  }
  __device__ inline double I(int i, int j, int k, int l) {
    gath++;
    return I4(T, i, j, k, l);
  }
};

// Couplings of one selector with four particles outside the block's mask
// (get_d2, selection.irp.f:738-853): at most 6 cells, each met once,
// written as (cell, value) pairs in the reference's order straight into the
// selector's LDS queue entry.
struct D2List {
  uint32_t *cellw;   // 3 words, two cells (i - 1) | (j - 1) << 8 per word
  double *hij;
  int n;
  __device__ inline void push(uint32_t c, double v) {   // c = (i - 1) | (j - 1) << 8
    if (n & 1) cellw[n >> 1] |= c << 16;
    else cellw[n >> 1] = c;
    hij[n] = v;
    ++n;
  }
};

// Up to 6 cells of one selector, collected before their integrals are
// loaded together: cell (i - 1) | (j - 1) << 8, table offsets of the value's
// one or two integrals (offsets fit 32 bits: M <= 256 for V, the M^4 table
// stops at kG4MaxOrb), and flags: bit k, value va - vb (else va); bit 8 + k,
// negative phase; bit 16 + k, slot used.  The value is (va - vb) * sign or
// va * sign with sign = +-1, formed as an exact negation.
struct CellSet {
  uint32_t cl[6], ia[6], ib[6], fl;
  __device__ inline void clear() {
#pragma unroll
    for (int k = 0; k < 6; ++k) cl[k] = ia[k] = ib[k] = 0u;
    fl = 0u;
  }
  __device__ inline bool used(int k) const { return (fl >> (16 + k)) & 1u; }
  __device__ inline void put2(int k, int i, int j, uint32_t a, uint32_t b, uint32_t odd) {
    cl[k] = (uint32_t)(i - 1) | ((uint32_t)(j - 1) << 8);
    ia[k] = a;
    ib[k] = b;
    fl |= (1u << k) | (odd << (8 + k)) | (1u << (16 + k));
  }
  __device__ inline void put1(int k, int i, int j, uint32_t a, uint32_t odd) {
    cl[k] = (uint32_t)(i - 1) | ((uint32_t)(j - 1) << 8);
    ia[k] = a;
    fl |= (odd << (8 + k)) | (1u << (16 + k));
  }
  // loads the integrals of the used cells and forms their values
  __device__ inline void values(const double *TT, double *v, int64_t &gath) const {
    double va[6], vb[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      va[k] = used(k) ? TT[ia[k]] : 0.0;
      vb[k] = ((fl >> k) & 1u) ? TT[ib[k]] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      double x = ((fl >> k) & 1u) ? va[k] - vb[k] : va[k];
      if ((fl >> (8 + k)) & 1u) x = -x;
      v[k] = x;
      if (used(k)) gath += ((fl >> k) & 1u) ? 2 : 1;
    }
  }
};
__device__ inline int cell_i(uint32_t c) { return (int)(c & 0xffu) + 1; }
__device__ inline int cell_j(uint32_t c) { return (int)((c >> 8) & 0xffu) + 1; }

// The particles (orbitals of a selector outside the block's mask, the first
// 4 kept) and holes (mask orbitals it lacks, the first 2 kept) of each spin,
// ascending, 1-based.  One orbital per byte (0-based): a run-time spin or
// position selects with shifts rather than an array index.
struct SelL {
  uint32_t pa, pb, ha, hb;
  int npa, npb, nha, nhb;
  __device__ int P(int n, int s) const { return (int)(((s == 1 ? pa : pb) >> (8 * (n - 1))) & 0xffu) + 1; }
  __device__ int H(int n, int s) const { return (int)(((s == 1 ? ha : hb) >> (8 * (n - 1))) & 0xffu) + 1; }
  __device__ int np(int s) const { return s == 1 ? npa : npb; }
};

template <int NI>
__device__ inline SelL sel_lists(const uint64_t *mask, const uint64_t *d) {
  SelL L{0u, 0u, 0u, 0u, 0, 0, 0, 0};
#pragma unroll
  for (int s = 1; s <= 2; ++s) {
    uint32_t pw = 0u, hw = 0u;
    int np = 0, nh = 0;
#pragma unroll
    for (int w = 0; w < NI; ++w) {
      uint64_t mob = ~spin_word<NI>(mask, s, w) & spin_word<NI>(d, s, w);
      uint64_t per = spin_word<NI>(mask, s, w) & ~spin_word<NI>(d, s, w);
      while (mob) {
        const int o = 64 * w + __ffsll((unsigned long long)mob) - 1;
        mob &= mob - 1;
        if (np < 4) pw |= (uint32_t)o << (8 * np);
        np++;
      }
      while (per) {
        const int o = 64 * w + __ffsll((unsigned long long)per) - 1;
        per &= per - 1;
        if (nh < 2) hw |= (uint32_t)o << (8 * nh);
        nh++;
      }
    }
    if (s == 1) {
      L.pa = pw; L.ha = hw; L.npa = np; L.nha = nh;
    } else {
      L.pb = pw; L.hb = hw; L.npb = np; L.nhb = nh;
    }
  }
  return L;
}

template <int NI>
__device__ __forceinline__ void get_d2(Ctx<NI> &C, const uint64_t *sel, const SelL &SL, int sp, D2List &out) {
  uint64_t pmk[2 * NI];
  prefix_parity<NI>(sel, pmk);
  // P(k, s) = k-th particle of spin s; H similarly.  The (at most 6)
  // cells are collected first, their integrals loaded together, then the
  // values pushed in the reference's order.
#define P(n, s) SL.P((n), (s))
#define H(n, s) SL.H((n), (s))
  int tip = SL.npa * SL.npb;
  int ma = sp;
  if (SL.npa > SL.npb) ma = 1;
  if (SL.npa < SL.npb) ma = 2;
  int mi = (ma % 2) + 1;
  const uint32_t *B1 = C.L.ban1;
  const int M = C.M;
  CellSet cs;
  cs.clear();
  auto idx4 = [&](int i, int j, int k, int l) -> uint32_t {   // <ij|kl>, 1-based
    return pair32(i - 1, k - 1) * (uint32_t)C.T.npair + pair32(j - 1, l - 1);
  };
  if (sp == 3) {
    const uint32_t *BB = (ma == 2) ? C.L.ban2 : C.L.ban1;
    if (tip == 3) {
      const int puti = P(1, mi);
#pragma unroll
      for (int i = 1; i <= 3; ++i) {
        const int putj = P(i, ma);
        if (ban_get(BB, M, putj, puti)) continue;
        const int i1 = (i == 1) ? 2 : 1, i2 = (i == 3) ? 2 : 3;
        const int p1 = P(i1, ma), p2 = P(i2, ma), h1 = H(1, ma), h2 = H(2, ma);
        const int k = i - 1;
        cs.put2(k, ma == 1 ? putj : puti, ma == 1 ? puti : putj, idx4(p1, p2, h1, h2), idx4(p2, p1, h1, h2),
                odd_pm<NI>(pmk, ma, ma, h1, p1, h2, p2));
      }
    } else {
      const int h1 = H(1, 1), h2 = H(1, 2);
#pragma unroll
      for (int j = 1; j <= 2; ++j) {
        const int putj = P(j, 2), p2 = P(3 - j, 2);
#pragma unroll
        for (int i = 1; i <= 2; ++i) {
          const int puti = P(i, 1);
          if (ban_get(BB, M, puti, putj)) continue;
          const int p1 = P(3 - i, 1);
          const int k = 2 * (j - 1) + (i - 1);
          cs.put1(k, puti, putj, idx4(p1, p2, h1, h2), odd_pm<NI>(pmk, 1, 2, h1, p1, h2, p2));
        }
      }
    }
  } else {
    if (tip == 0) {
      const int h1 = H(1, ma), h2 = H(2, ma);
      int k = 0;
#pragma unroll
      for (int i = 1; i <= 3; ++i) {
        const int puti = P(i, ma);
#pragma unroll
        for (int j = i + 1; j <= 4; ++j) {
          const int putj = P(j, ma);
          const int kk = k++;
          if (ban_get(B1, M, puti, putj)) continue;
          int c[2], n = 0;
#pragma unroll
          for (int q = 1; q <= 4; ++q)
            if (q != i && q != j) c[n++] = q;
          const int p1 = P(c[0], ma), p2 = P(c[1], ma);
          cs.put2(kk, puti, putj, idx4(p1, p2, h1, h2), idx4(p2, p1, h1, h2), odd_pm<NI>(pmk, ma, ma, h1, p1, h2, p2));
        }
      }
    } else if (tip == 3) {
      const int h1 = H(1, mi), h2 = H(1, ma), p1 = P(1, mi);
#pragma unroll
      for (int i = 1; i <= 3; ++i) {
        const int i1 = (i == 1) ? 2 : 1, i2 = (i == 3) ? 2 : 3;
        const int puti = P(i1, ma), putj = P(i2, ma);
        if (ban_get(B1, M, puti, putj)) continue;
        const int p2 = P(i, ma);
        const int k = i - 1;
        cs.put1(k, min(puti, putj), max(puti, putj), idx4(p1, p2, h1, h2), odd_pm<NI>(pmk, mi, ma, h1, p1, h2, p2));
      }
    } else {
      const int puti = P(1, sp), putj = P(2, sp);
      if (!ban_get(B1, M, puti, putj)) {
        const int p1 = P(1, mi), p2 = P(2, mi), h1 = H(1, mi), h2 = H(2, mi);
        cs.put2(0, puti, putj, idx4(p1, p2, h1, h2), idx4(p2, p1, h1, h2), odd_pm<NI>(pmk, mi, mi, h1, p1, h2, p2));
      }
    }
  }
  double v[6];
  cs.values(C.T.V, v, C.gath);
#pragma unroll
  for (int k = 0; k < 6; ++k)
    if (cs.used(k)) out.push(cs.cl[k], v[k]);
}


template <int NI>
__device__ __forceinline__ void get_d1(Ctx<NI> &C, const uint64_t *sel, const SelL &SL, int sp, const double *coefs,
                                       int na_e, int nb_e, bool singles) {
  uint64_t pmk[2 * NI];
  prefix_parity<NI>(sel, pmk);
  const int M = C.M;
  const uint32_t *BO = C.L.borb;
  const uint32_t *B1 = C.L.ban1;
  // lbanned = bannedOrb + the selector's particles
  auto lbanned = [&](int o, int s) -> bool {
    if (borb_get(BO, o, s)) return true;
    const int n = SL.np(s);
#pragma unroll
    for (int k = 1; k <= 4; ++k)
      if (k <= n && P(k, s) == o) return true;
    return false;
  };
  int ma = 1;
  if (SL.npb >= 2) ma = 2;
  int mi = (ma == 1) ? 2 : 1;
  // The cells this thread owns on the selector's lines (at most 6) are
  // collected first, their integrals loaded together, then added.
  CellSet cs;
  cs.clear();
  // <ij|kl> (1-based) with argument `vary` the orbital that changes along
  // the line (the cells' owners are consecutive threads)
  auto idx4 = [&](int i, int j, int k, int l, int vary) -> uint32_t { return line_off(C.T, i, j, k, l, vary); };
  const double *TT = C.T.G4 ? C.T.G4 : C.T.V;
  auto row_of = [&](int j) -> int { return owned_i(j, M); };   // i with owns(i, j), 0 if none
  // phases: odd_pm's terms of the orbitals fixed for the whole selector are
  // computed once; a cell adds those of its varying orbital
  auto pb = [&](int s, int o) -> uint32_t { return (uint32_t)pm_bit<NI>(pmk, s, o); };
  // same-spin double (h, hfix) -> (p1, p2) with h the cell's varying orbital
  auto same_odd = [&](uint32_t fixed, int h, int hfix, int p1, int p2) -> uint32_t {
    const int h1 = min(h, hfix), h2 = max(h, hfix);
    return fixed ^ pb(ma, h) ^ (uint32_t)(p1 < h1) ^ (uint32_t)(p2 < h2) ^ (uint32_t)(max(h1, p1) > min(h2, p2));
  };
  if (sp == 3) {
    const uint32_t *BB = (ma == 2) ? C.L.ban2 : C.L.ban1;
    const int puti = P(1, mi), hfix = H(1, ma), p1 = P(1, ma), p2 = P(2, ma);
    const uint32_t f_ma = pb(ma, hfix) ^ pb(ma, p1) ^ pb(ma, p2);
    if (!borb_get(BO, puti, mi)) {
      const int putj = (ma == 1) ? row_of(puti) : owned_j(puti, M);
      if (putj && putj != hfix && !lbanned(putj, ma) && !ban_get(BB, M, putj, puti)) {
        const int ci = (ma == 1) ? putj : puti, cj = (ma == 1) ? puti : putj;
        const uint32_t odd = same_odd(f_ma, putj, hfix, p1, p2);
        if (putj < hfix)
          cs.put2(0, ci, cj, idx4(p1, p2, putj, hfix, 2), idx4(p2, p1, putj, hfix, 2), odd);
        else
          cs.put2(0, ci, cj, idx4(p1, p2, hfix, putj, 3), idx4(p2, p1, hfix, putj, 3), odd);
      }
    }
    const int pfix = P(1, mi);
    const uint32_t f_mi = pb(ma, hfix) ^ pb(mi, pfix);
    // lines with fixed ma-orbital p1 and p2, varying mi-orbital
#pragma unroll
    for (int which = 0; which < 2; ++which) {
      const int pf = which == 0 ? p1 : p2;      // fixed (ma) orbital of the cell
      const int other = which == 0 ? p2 : p1;   // the ma particle in the integral
      const int pu = (mi == 1) ? row_of(pf) : owned_j(pf, M);
      if (!pu || lbanned(pu, mi)) continue;
      if (ban_get(BB, M, pf, pu)) continue;
      // odd_pm(ma, mi, hfix, other, pu, pfix), two spins
      const uint32_t odd = f_mi ^ pb(ma, other) ^ (uint32_t)(other < hfix) ^ pb(mi, pu) ^ (uint32_t)(pfix < pu);
      cs.put1(1 + which, (mi == 1) ? pu : pf, (mi == 1) ? pf : pu, idx4(other, pfix, hfix, pu, 3), odd);
    }
  } else if (SL.np(ma) == 3) {
    const int hfix = H(1, ma);
    const uint32_t f_all = pb(ma, hfix) ^ pb(ma, P(1, ma)) ^ pb(ma, P(2, ma)) ^ pb(ma, P(3, ma));
#pragma unroll
    for (int i = 1; i <= 3; ++i) {
      const int puti = P(i, ma);
      const int i1 = (i == 1) ? 2 : 1, i2 = (i == 3) ? 2 : 3;
      const int p1 = P(i1, ma), p2 = P(i2, ma);
      const uint32_t f_i = f_all ^ pb(ma, puti);   // hfix, p1 and p2
      // cells {x, puti}: (x, puti) for x < puti, (puti, x) for x >= puti
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int k = 2 * (i - 1) + half;
        int x;
        if (half == 0) {
          x = row_of(puti);
          if (!x || x >= puti) continue;
        } else {
          x = owned_j(puti, M);
          if (!x || x < puti) continue;
        }
        const int putj = x;
        if (putj == hfix) continue;
        if (lbanned(putj, ma) || ban_get(B1, M, puti, putj)) continue;
        const int ci = half == 0 ? x : puti, cj = half == 0 ? puti : x;
        const uint32_t odd = same_odd(f_i, putj, hfix, p1, p2);
        if (putj < hfix)
          cs.put2(k, ci, cj, idx4(p1, p2, putj, hfix, 2), idx4(p2, p1, putj, hfix, 2), odd);
        else
          cs.put2(k, ci, cj, idx4(p1, p2, hfix, putj, 3), idx4(p2, p1, hfix, putj, 3), odd);
      }
    }
  } else {
    const int hfix = H(1, mi), pfix = P(1, mi), p1 = P(1, ma), p2 = P(2, ma);
    const uint32_t f_mi = pb(mi, hfix) ^ pb(mi, pfix) ^ (uint32_t)(pfix < hfix);
#pragma unroll
    for (int which = 0; which < 2; ++which) {
      const int pc = which == 0 ? p2 : p1;   // line {x, pc}
      const int other = which == 0 ? p1 : p2;
      const uint32_t f_w = f_mi ^ pb(ma, other);
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int k = 2 * which + half;
        int x;
        if (half == 0) {
          x = row_of(pc);
          if (!x || x >= pc) continue;
        } else {
          x = owned_j(pc, M);
          if (!x || x < pc) continue;
        }
        const int puti = x;
        if (lbanned(puti, ma)) continue;
        if (ban_get(B1, M, puti, pc)) continue;
        // odd_pm(mi, ma, hfix, pfix, puti, other), two spins
        cs.put1(k, half == 0 ? x : pc, half == 0 ? pc : x, idx4(pfix, other, hfix, puti, 3),
                f_w ^ pb(ma, puti) ^ (uint32_t)(other < puti));
      }
    }
  }
  {
    double v[6];
    cs.values(TT, v, C.gath);
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (cs.used(k)) C.add(cell_i(cs.cl[k]), cell_j(cs.cl[k]), coefs, v[k]);
  }
  // singles folded into this block (normally applied as records of the
  // selector's queue entry instead)
  if (!singles) return;
  int s1, s2;
  if (sp == 3) { s1 = 1; s2 = 2; } else { s1 = sp; s2 = sp; }
  for (int i1 = 1; i1 <= SL.np(s1); ++i1) {
    int ib = (s1 == s2) ? i1 + 1 : 1;
    for (int i2 = ib; i2 <= SL.np(s2); ++i2) {
      int p1 = P(i1, s1), p2 = P(i2, s2);
      if (!owns(p1, p2, M)) continue;
      if (borb_get(BO, p1, s1) || borb_get(BO, p2, s2) || ban_get(B1, M, p1, p2)) continue;
      uint64_t det[2 * NI];
#pragma unroll
      for (int w = 0; w < 2 * NI; ++w) det[w] = C.mask[w];
      set_orb<NI>(det, s1, p1);
      set_orb<NI>(det, s2, p2);
      double hij = i_H_j<NI>(C.T, sel, det, na_e, nb_e);
      C.gath += 4;
      C.add(p1, p2, coefs, hij);
    }
  }
}

template <int NI>
__device__ __forceinline__ void get_d0(Ctx<NI> &C, const uint64_t *sel, const SelL &SL, int sp, const double *coefs,
                                       int na_e, int nb_e) {
  // one cell per orbital row and thread.  Cells on the rows and columns of
  // the selector's own particles need the full matrix element (i_H_j) and
  // are done separately; for the others the integrals of 8 (N_int = 1) or 4
  // rows are loaded
  // together before they are used.
  constexpr int RB8 = (NI == 1) ? 8 : 4;   // rows per batch of loads
  const int M = C.M;
  const uint32_t *BO = C.L.borb;
  const uint32_t *B1 = C.L.ban1;
  auto full = [&](int i, int j, int si, int sj) __attribute__((always_inline)) {   // cell (i, j) via i_H_j
    uint64_t det[2 * NI];
#pragma unroll
    for (int w = 0; w < 2 * NI; ++w) det[w] = C.mask[w];
    set_orb<NI>(det, si, i);
    set_orb<NI>(det, sj, j);
    C.gath += 4;
    C.add(i, j, coefs, i_H_j<NI>(C.T, sel, det, na_e, nb_e));
  };
  auto col_owner_row = [&](int j) -> int { return owned_i(j, M); };   // i with owns(i, j), 0 if none
  // the cells of a row belong to consecutive threads: their integrals are
  // read along the column orbital
  const double *TT = C.T.G4 ? C.T.G4 : C.T.V;
  int r_lo, r_hi;
  band_rows(M, r_lo, r_hi);
  if (sp == 3) {
    const int h1 = P(1, 1), h2 = P(1, 2);
    for (int r0 = r_lo; r0 <= r_hi; r0 += RB8) {
      int cj[RB8];
      bool ok[RB8];
      double v[RB8];
#pragma unroll
      for (int b = 0; b < RB8; ++b) {
        const int p1 = r0 + b;
        const int p2 = (p1 <= r_hi) ? owned_j(p1, M) : 0;
        const bool o = p2 && p1 != h1 && p2 != h2 &&
                       !(borb_get(BO, p1, 1) || borb_get(BO, p2, 2) || ban_get(B1, M, p1, p2));
        cj[b] = p2;
        ok[b] = o;
        v[b] = o ? TT[line_off(C.T, p1, p2, h1, h2, 1)] : 0.0;
      }
#pragma unroll
      for (int b = 0; b < RB8; ++b) {
        if (!ok[b]) continue;
        const int p1 = r0 + b, p2 = cj[b];
        C.gath++;
        C.add(p1, p2, coefs, v[b] * phase_bi<NI>(sel, 1, 2, h1, p1, h2, p2));
      }
    }
    // row h1 and column h2
    {
      const int p2 = owned_j(h1, M);
      if (p2 && !(borb_get(BO, h1, 1) || borb_get(BO, p2, 2) || ban_get(B1, M, h1, p2))) full(h1, p2, 1, 2);
      const int p1 = col_owner_row(h2);
      if (p1 && p1 != h1 && !(borb_get(BO, p1, 1) || borb_get(BO, h2, 2) || ban_get(B1, M, p1, h2)))
        full(p1, h2, 1, 2);
    }
  } else {
    const int p1 = P(1, sp), p2 = P(2, sp);
    for (int r0 = r_lo; r0 <= r_hi; r0 += RB8) {
      int cj[RB8];
      bool ok[RB8];
      double va[RB8], vb[RB8];
#pragma unroll
      for (int b = 0; b < RB8; ++b) {
        const int puti = r0 + b;
        const int putj = (puti <= r_hi) ? owned_j(puti, M) : 0;
        const bool o = putj && putj > puti && puti != p1 && putj != p2 && puti != p2 && putj != p1 &&
                       !(borb_get(BO, puti, sp) || borb_get(BO, putj, sp) || ban_get(B1, M, puti, putj));
        cj[b] = putj;
        ok[b] = o;
        va[b] = o ? TT[line_off(C.T, p1, p2, puti, putj, 3)] : 0.0;
        vb[b] = o ? TT[line_off(C.T, p2, p1, puti, putj, 3)] : 0.0;
      }
#pragma unroll
      for (int b = 0; b < RB8; ++b) {
        if (!ok[b]) continue;
        const int puti = r0 + b, putj = cj[b];
        C.gath += 2;
        C.add(puti, putj, coefs, (va[b] - vb[b]) * phase_bi<NI>(sel, sp, sp, puti, p1, putj, p2));
      }
    }
    // cells (puti < putj) on the rows and columns of p1 and p2, each once
    auto try_cell = [&](int i, int j) {
      if (!(i >= 1 && j > i && j <= M)) return;
      if (borb_get(BO, i, sp) || borb_get(BO, j, sp) || ban_get(B1, M, i, j)) return;
      full(i, j, sp, sp);
    };
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int o = a == 0 ? p1 : p2;
      // row o: cell (o, j) with j > o
      const int j = owned_j(o, M);
      if (j) try_cell(o, j);
      // column o: cell (i, o) with i < o, unless row i is p1 or p2 (done above)
      const int i = col_owner_row(o);
      if (i && i != p1 && i != p2) try_cell(i, o);
    }
  }
#undef P
#undef H
}


// diag_H_mat_elem_fock (slater_rules.irp.f:1800-1867)
__device__ inline double diag_fock(const Tabs &T, const double *f, int sp, int h1, int p1, int h2, int p2) {
  const int M = T.M;
#define F(s, i) f[((s) - 1) + 2 * ((i) - 1)]
#define JA(a, b) T.JA[((a) - 1) + M * ((b) - 1)]
#define JJ(a, b) T.J[((a) - 1) + M * ((b) - 1)]
  double E0 = F(1, M + 1);
  // degree and ordering as get_double_excitation(gen, det) would return
  if (sp == 3) {
    bool da = (p1 != h1), db = (p2 != h2);
    if (da && db) {
      return E0 - F(1, h1) + (F(1, p1) - JA(h1, p1)) - (F(2, h2) - JJ(h1, h2) + JJ(p1, h2)) +
             (F(2, p2) - JJ(h1, p2) + JJ(p1, p2) - JA(h2, p2));
    }
    if (da) return E0 - F(1, h1) + (F(1, p1) - JA(h1, p1));
    if (db) return E0 - F(2, h2) + (F(2, p2) - JA(h2, p2));
    return E0;
  }
  // same spin: holes {h1 < h2}, particles {p1 < p2}
  int hs[2] = {h1, h2}, ps[2] = {p1, p2};
  int hr[2], pr[2], n = 0;
  bool usedp[2] = {false, false}, usedh[2] = {false, false};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      if (!usedh[a] && !usedp[b] && hs[a] == ps[b]) { usedh[a] = true; usedp[b] = true; }
  int k = 0;
  for (int a = 0; a < 2; ++a) if (!usedh[a]) hr[k++] = hs[a];
  n = k;
  k = 0;
  for (int b = 0; b < 2; ++b) if (!usedp[b]) pr[k++] = ps[b];
  const int s = sp;
  if (n == 2) {
    return E0 - F(s, hr[0]) + (F(s, pr[0]) - JA(hr[0], pr[0])) - (F(s, hr[1]) - JA(hr[0], hr[1]) + JA(pr[0], hr[1])) +
           (F(s, pr[1]) - JA(hr[0], pr[1]) + JA(pr[0], pr[1]) - JA(hr[1], pr[1]));
  }
  if (n == 1) return E0 - F(s, hr[0]) + (F(s, pr[0]) - JA(hr[0], pr[0]));
  return E0;
#undef F
#undef JA
#undef JJ
}

// The LDS queue of a block's surviving selectors, as arrays (qcap entries):
// per entry its index, kind (0: four particles outside the mask; 1: three;
// 2: two or fewer) with the number and first slot of its records, six
// value slots (hij) with their cells ((i - 1) | (j - 1) << 8, two per word)
// and its coefficients.  A four-particle entry's slots hold its couplings;
// a three-particle entry's hold its determinant (N_int <= 3) and then its
// folded single excitations; another entry's hold its determinant.  The
// replay of a flush keeps, per owner thread, the list of (entry, slot)
// records that fall on its cells (rec, cursor), and the positions of the
// three- and fewer-particle entries (k12).
struct QueueLds {
  double *hij;      // 6 * qcap
  double *coef;     // NS * qcap
  int *idx;         // qcap
  int *kind_n;      // qcap
  uint32_t *cellw;  // 3 * qcap
  int *cursor;      // SB
  uint16_t *rec;    // 6 * qcap
  uint16_t *k12;    // qcap
  __host__ __device__ static size_t bytes(int qcap, int NS) {
    return (size_t)qcap * (8 * (6 + NS) + 4 * 5 + 2 * 7) + 4 * (size_t)SB;
  }
  __device__ void carve(unsigned char *base, int qcap, int NS) {
    hij = reinterpret_cast<double *>(base);
    coef = hij + 6 * (size_t)qcap;
    idx = reinterpret_cast<int *>(coef + (size_t)NS * qcap);
    kind_n = idx + qcap;
    cellw = reinterpret_cast<uint32_t *>(kind_n + qcap);
    cursor = reinterpret_cast<int *>(cellw + 3 * (size_t)qcap);
    rec = reinterpret_cast<uint16_t *>(cursor + SB);
    k12 = rec + 6 * (size_t)qcap;
  }
};

// a queue entry of three or fewer particles keeps its determinant in the
// six coupling slots (48 bytes) when it fits there (N_int <= 3)
template <int NI>
constexpr bool kQueueHoldsDet = 2 * NI * sizeof(uint64_t) <= 6 * sizeof(double);


// GM = false: the coupling matrix lives in LDS.  GM = true (N_states * M^2
// doubles do not fit next to the queue): it lives in HBM, one slot of
// A.mat_g per workgroup of the launch; each cell is still read and written
// by its owner thread only.
template <int NI, bool GM>
__global__ __launch_bounds__(SB) void select_block_kernel(SelArgs A, Tabs T) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int M = T.M, NS = A.NS;
  const int nbits = (M * M + 31) / 32;
  BlockLds L;
  if constexpr (GM) {
    L.mat = A.mat_g + (size_t)blockIdx.x * NS * M * M;
    L.fock = reinterpret_cast<double *>(lds_raw);
  } else {
    L.mat = reinterpret_cast<double *>(lds_raw);
    L.fock = L.mat + (size_t)NS * M * M;
  }
  L.ban1 = reinterpret_cast<uint32_t *>(L.fock + 2 * (M + 1));
  L.ban2 = L.ban1 + nbits;
  L.borb = L.ban2 + nbits;
  L.flag = reinterpret_cast<int *>(L.borb + 16);
  int *wsum = L.flag + 4;  // per-wavefront counts of the queue compaction
  // reduction scratch (SB * MAXNS doubles), shared with the selector queue
  double *red = reinterpret_cast<double *>(wsum + 8);
  red = reinterpret_cast<double *>((reinterpret_cast<uintptr_t>(red) + 7) & ~uintptr_t(7));

  const uint64_t t_start = wall_clock64();
  const BlockDesc B = A.blocks[blockIdx.x];
  const int tid = threadIdx.x;
  const uint64_t *gen = A.psi + (int64_t)(B.gen - 1) * 2 * NI;
  const int sp = (B.s1 != B.s2) ? 3 : B.s1;
  uint64_t mask[2 * NI];
  build_mask<NI>(gen, B.s1, B.h1, B.s2, B.h2, mask);

  // ---- init ----
  for (int x = tid; x < NS * M * M; x += SB) L.mat[x] = 0.0;
  for (int x = tid; x < 2 * nbits; x += SB) L.ban1[x] = 0u;  // ban1 and ban2 are contiguous
  if (tid < 16) {
    // bannedOrb: every orbital occupied in the generator (holes of the
    // particle mask), then the mono unbans
    int s = tid / 8 + 1, w = tid & 7;
    uint32_t v = 0;
    for (int bb = 0; bb < 32; ++bb) {
      int o = w * 32 + bb + 1;
      bool banned = (o > M) || bit(gen + (s - 1) * NI, o);
      if (banned) v |= 1u << bb;
    }
    L.borb[tid] = v;
  }
  if (tid == 0) L.flag[0] = 0;
  // fock_diag_tmp of the generator (computed once per generator by fock_gen_kernel)
  for (int x = tid; x < 2 * (M + 1); x += SB) L.fock[x] = A.fockg[(size_t)B.gslot * 2 * (M + 1) + x];
  __syncthreads();
  if (tid == 0) {
    if (B.unbanA) L.borb[(B.s1 - 1) * 8 + ((B.h1 - 1) >> 5)] &= ~(1u << ((B.h1 - 1) & 31));
    if (B.unbanB) L.borb[(B.s2 - 1) * 8 + ((B.h2 - 1) >> 5)] &= ~(1u << ((B.h2 - 1) & 31));
  }
  __syncthreads();

  const int *near = A.near_list + A.near_off[B.gslot];
  const uint64_t *nsig = A.near_sig + A.near_off[B.gslot];
  const int n_near = A.near_cnt[B.gslot];
  const uint32_t c1 = hole_code(B.s1, B.h1), c2 = hole_code(B.s2, B.h2);

  // ---- spot_isinwf ----
  for (int x = tid; x < n_near; x += SB) {
    if (!sig_in_mask(nsig[x], c1, c2)) continue;   // the entry does not contain the mask
    const int idx = near[x];
    const uint64_t *d = A.psi + (int64_t)(idx - 1) * 2 * NI;
    if (idx < B.gen) {
      atomicOr(&L.flag[0], 1);
      continue;
    }
    int l0 = 0, l1 = 0, n = 0;
#pragma unroll
    for (int w = 0; w < 2 * NI; ++w) {
      uint64_t y = d[w] & ~mask[w];
      while (y && n < 2) {
        const int o = 64 * (w % NI) + __ffsll((unsigned long long)y);
        if (n == 0) l0 = o;
        else l1 = o;
        ++n;
        y &= y - 1;
      }
    }
    if (n == 2) ban_set(L.ban1, M, l0, l1);
  }
  __syncthreads();
  if (L.flag[0]) {
    if (tid == 0) {
      A.blk_status[blockIdx.x] = 1;
      for (int st = 0; st < NS; ++st) A.blk_pt2[(int64_t)blockIdx.x * NS + st] = 0.0;
      A.blk_ncand[blockIdx.x] = 0;
      A.blk_ngath[blockIdx.x] = 0;
      A.blk_ticks[blockIdx.x] = (int64_t)(wall_clock64() - t_start);
    }
    return;
  }
  // ---- past_d1 / past_d2 (selectors before the generator) ----
  // the near list is in index order: those selectors are its first entries
  const int n_past = A.near_past[B.gslot];
  for (int x = tid; x < n_past; x += SB) {
    const int idx = near[x];
    if (idx >= B.gen || idx > A.n_sel) continue;
    const int nt = sig_nt(nsig[x], c1, c2);
    if (nt < 3 || nt > 4) continue;
    const uint64_t *d = A.psi + (int64_t)(idx - 1) * 2 * NI;
    const SelL SL = sel_lists<NI>(mask, d);
    if (nt == 4) {
      if (sp == 3) {
        for (int i = 1; i <= SL.npa; ++i)
          for (int j = 1; j <= SL.npb; ++j) ban_set(L.ban1, M, SL.P(i, 1), SL.P(j, 2));
      } else {
        const int n = SL.np(sp);
        for (int i = 1; i <= n; ++i)
          for (int j = 1; j < i; ++j) {
            ban_set(L.ban1, M, SL.P(j, sp), SL.P(i, sp));
            ban_set(L.ban1, M, SL.P(i, sp), SL.P(j, sp));
          }
      }
    } else {
#pragma unroll
      for (int s = 1; s <= 2; ++s)
        for (int i = 1; i <= SL.np(s); ++i) borb_set(L.borb, SL.P(i, s), s);
    }
  }
  __syncthreads();
  // ---- symmetrisation when the generator itself is met (splash_pq:692-706) ----
  if (sp == 3) {
    for (int x = tid; x < M * M; x += SB) {
      int j = x % M + 1, k = x / M + 1;
      if (ban_get(L.ban1, M, k, j)) ban_set(L.ban2, M, j, k);
    }
  } else {
    for (int x = tid; x < M * M; x += SB) {
      int k = x % M + 1, l = x / M + 1;
      if (l > k && ban_get(L.ban1, M, k, l)) ban_set(L.ban1, M, l, k);
    }
  }
  __syncthreads();
  // lower-triangle entries that the reference overwrites with a false upper
  // entry: banned(l,k,1) = banned(k,l,1) also clears bits set only below the
  // diagonal by past_d2 (never happens: past_d2 sets both orientations)

  // ---- splash_pq: couplings in selector order ----
  // Phase A: FR * SB near-list entries at a time are filtered in parallel and the
  // survivors appended, in near-list order, to an LDS queue with their
  // determinants.  Phase B (flush): the couplings of the four-particle
  // entries are computed, one entry per thread (get_d2); then the queue in
  // order: a four-particle entry adds its values in the cells' owner
  // threads, the others run get_d1 / get_d0 with every thread.  Each cell
  // thus receives its terms in selector order, as in the reference.
  Ctx<NI> C(T, A, L, mask);
  QueueLds Q;
  Q.carve(reinterpret_cast<unsigned char *>(red), A.qcap, NS);
  const int lane = tid & 63, wv = tid >> 6;
  // Replay of the queue (entries 0..nq-1, in near-list order).  Each cell
  // of the coupling matrix is touched only by its owner thread, so each
  // thread can apply its own terms in queue order without synchronisation:
  //   1. every thread counts the records of its entries per owner and the
  //      positions of the three- and fewer-particle entries are compacted in
  //      order (k12);
  //   2. an exclusive scan of the counts gives each owner a range of rec,
  //      and the records are scattered into it;
  //   3. each thread sorts its records by (entry, slot) -- the values are
  //      distinct, so the order is fixed -- and walks the k12 list: before
  //      entry k it adds its records of earlier entries, then runs get_d1 /
  //      get_d0 of entry k on its cells; entry k's own records (its folded
  //      singles, on cells its lines never touch) and those of later entries
  //      follow.
  auto cell_of = [&](int e, int c) -> uint32_t {
    return (Q.cellw[3 * e + (c >> 1)] >> (16 * (c & 1))) & 0xffffu;
  };
  auto flush = [&](int nq) {
    // couplings of the four-particle entries (get_d2), one entry per thread:
    // the queue is dense, so the lanes are busy where the filter's survivors
    // were scattered over its chunk
    // The same pass computes the single excitations a three-particle
    // selector folds into the block (the end of get_d1) as records of its
    // entry, in the slots after its determinant; when they do not fit there
    // the replay computes them (flag bit 12).  kind_n = kind | records << 4
    // | first record slot << 8.
    // Four-particle entries go to the lower half of the workgroup, the others
    // to the upper half, so that the singles' integral loads overlap those
    // of get_d2 instead of following them in the same lanes.
    const int dhalf = tid / (SB / 2), dt = tid - dhalf * (SB / 2);
    for (int e = dt; e < nq; e += SB / 2) {
      // the other half may already have rewritten this entry's kind_n; the
      // kind nibble stays
      const int kind = Q.kind_n[e] & 0xf;
      if ((kind == 0) != (dhalf == 0)) continue;
      const int idx = Q.idx[e];
      for (int st = 0; st < NS; ++st) Q.coef[(size_t)NS * e + st] = A.coef[(idx - 1) + A.n_det * st];
      C.gath += NS;
      uint64_t d[2 * NI];
      const uint64_t *dp = A.psi + (int64_t)(idx - 1) * 2 * NI;
#pragma unroll
      for (int w = 0; w < 2 * NI; ++w) d[w] = dp[w];
      if (kQueueHoldsDet<NI> && kind != 0) {
        // the determinant of a three- or fewer-particle selector goes where
        // a four-particle selector's couplings go, for the replay to read
        uint64_t *qd = reinterpret_cast<uint64_t *>(Q.hij + 6 * e);
#pragma unroll
        for (int w = 0; w < 2 * NI; ++w) qd[w] = d[w];
      }
      if (kind == 2) continue;
      const SelL SL = sel_lists<NI>(mask, d);
      if (kind == 0) {
        D2List dl{Q.cellw + 3 * e, Q.hij + 6 * e, 0};
        get_d2<NI>(C, d, SL, sp, dl);
        Q.kind_n[e] = dl.n << 4;
      } else {
        constexpr int c0 = kQueueHoldsDet<NI> ? 2 * NI : 0;
        const int s1 = (sp == 3) ? 1 : sp, s2 = (sp == 3) ? 2 : sp;
        auto single_ok = [&](int p1, int p2) {
          return !(borb_get(L.borb, p1, s1) || borb_get(L.borb, p2, s2) || ban_get(L.ban1, M, p1, p2));
        };
        int ns = 0;
        for (int i1 = 1; i1 <= SL.np(s1); ++i1)
          for (int i2 = (s1 == s2) ? i1 + 1 : 1; i2 <= SL.np(s2); ++i2) ns += single_ok(SL.P(i1, s1), SL.P(i2, s2));
        if (c0 + ns > 6) {
          Q.kind_n[e] = 1 | (1 << 12);
        } else {
          D2List dl{Q.cellw + 3 * e, Q.hij + 6 * e, c0};
          for (int i1 = 1; i1 <= SL.np(s1); ++i1)
            for (int i2 = (s1 == s2) ? i1 + 1 : 1; i2 <= SL.np(s2); ++i2) {
              const int p1 = SL.P(i1, s1), p2 = SL.P(i2, s2);
              if (!single_ok(p1, p2)) continue;
              uint64_t det[2 * NI];
#pragma unroll
              for (int w = 0; w < 2 * NI; ++w) det[w] = mask[w];
              set_orb<NI>(det, s1, p1);
              set_orb<NI>(det, s2, p2);
              C.gath += 4;
              dl.push((uint32_t)(p1 - 1) | ((uint32_t)(p2 - 1) << 8), i_H_j<NI>(T, d, det, A.na, A.nb));
            }
          Q.kind_n[e] = 1 | (ns << 4) | (c0 << 8);
        }
      }
    }
    Q.cursor[tid] = 0;
    __syncthreads();
    int n12 = 0;
    for (int e0 = 0; e0 < nq; e0 += SB) {
      const int e = e0 + tid;
      bool is12 = false;
      if (e < nq) {
        const int kn = Q.kind_n[e];
        const int c0 = (kn >> 8) & 7, c1 = c0 + ((kn >> 4) & 0xf);
        for (int c = c0; c < c1; ++c) {
          const uint32_t cell = cell_of(e, c);
          atomicAdd(&Q.cursor[owner_of(cell_i(cell), cell_j(cell), M)], 1);
        }
        is12 = (kn & 0xf) != 0;
      }
      const unsigned long long bal = __ballot(is12);
      const int before = __popcll(bal & ((1ULL << lane) - 1ULL));
      if (lane == 0) wsum[wv] = __popcll(bal);
      __syncthreads();
      int off = 0, tot = 0;
      for (int w = 0; w < SB / 64; ++w) {
        if (w < wv) off += wsum[w];
        tot += wsum[w];
      }
      if (is12) Q.k12[n12 + off + before] = (uint16_t)e;
      n12 += tot;
      __syncthreads();
    }
    // exclusive scan of the per-owner counts
    const int my_cnt = Q.cursor[tid];
    int incl = my_cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int my_start = incl - my_cnt;
    for (int w = 0; w < wv; ++w) my_start += wsum[w];
    Q.cursor[tid] = my_start;
    __syncthreads();
    for (int e = tid; e < nq; e += SB) {
      const int kn = Q.kind_n[e];
      const int c0 = (kn >> 8) & 7, c1 = c0 + ((kn >> 4) & 0xf);
      for (int c = c0; c < c1; ++c) {
        const uint32_t cell = cell_of(e, c);
        const int slot = atomicAdd(&Q.cursor[owner_of(cell_i(cell), cell_j(cell), M)], 1);
        Q.rec[slot] = (uint16_t)((e << 3) | c);
      }
    }
    __syncthreads();
    uint16_t *mine = Q.rec + my_start;
    for (int a = 1; a < my_cnt; ++a) {   // insertion sort (a few records per thread)
      const uint16_t x = mine[a];
      int b = a - 1;
      while (b >= 0 && mine[b] > x) {
        mine[b + 1] = mine[b];
        --b;
      }
      mine[b + 1] = x;
    }
    int r = 0;
    auto apply_until = [&](int k) {
      while (r < my_cnt) {
        const uint16_t x = mine[r];
        const int e = x >> 3;
        if (e >= k) break;
        const int c = x & 7;
        const uint32_t cell = cell_of(e, c);
        C.add(cell_i(cell), cell_j(cell), Q.coef + (size_t)NS * e, Q.hij[6 * e + c]);
        ++r;
      }
    };
    for (int q = 0; q < n12; ++q) {
      const int k = Q.k12[q];
      apply_until(k);
      const int kn = Q.kind_n[k];
      const int kind = kn & 0xf;
      uint64_t d[2 * NI];
      const uint64_t *dp = kQueueHoldsDet<NI> ? reinterpret_cast<const uint64_t *>(Q.hij + 6 * k)
                                              : A.psi + (int64_t)(Q.idx[k] - 1) * 2 * NI;
#pragma unroll
      for (int w = 0; w < 2 * NI; ++w) d[w] = dp[w];
      const SelL SL = sel_lists<NI>(mask, d);
      if (kind == 1) get_d1<NI>(C, d, SL, sp, Q.coef + (size_t)NS * k, A.na, A.nb, (kn >> 12) & 1);
      else get_d0<NI>(C, d, SL, sp, Q.coef + (size_t)NS * k, A.na, A.nb);
    }
    apply_until(nq);
  };
  int qn = 0;
  // Each round, every thread tests FR consecutive near-list entries from
  // their hole signatures alone (no determinant is loaded); the survivors
  // keep their near-list order through a workgroup scan of the per-thread
  // counts.
  constexpr int FR = 4;
  // the selectors before the generator (the first n_past entries of the
  // near list) never pass the filter: the scan starts after them
  // one extra round past the end of the list flushes what is left
  for (int x0 = n_past;; x0 += FR * SB) {
    const bool more = x0 < n_near;
    int idx[FR];
    uint32_t passm = 0, kinds = 0;   // bit j: entry j survives; 2 bits per entry: its kind
    int cnt = 0, pos0 = 0;
    if (more) {
      const int xt = x0 + FR * tid;
#pragma unroll
      for (int j = 0; j < FR; ++j) {
        idx[j] = 0;
        uint64_t sg = 0;
        if (xt + j < n_near) {
          idx[j] = near[xt + j];
          sg = nsig[xt + j];
        }
        if (idx[j] >= B.gen && idx[j] <= A.n_sel) {   // idx == 0 past the end of the list
          const int nt = sig_nt(sg, c1, c2);
          if (nt <= 4) {
            passm |= 1u << j;
            kinds |= (uint32_t)(nt == 4 ? 0 : (nt == 3 ? 1 : 2)) << (2 * j);
          }
        }
      }
      // order-preserving compaction: exclusive scan of the survivor counts
      const int c = __popc(passm);
      int incl = c;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
      }
      if (lane == 63) wsum[wv] = incl;
      __syncthreads();
      int off = 0;
      for (int w = 0; w < SB / 64; ++w) {
        if (w < wv) off += wsum[w];
        cnt += wsum[w];
      }
      pos0 = off + incl - c;
    }
    int done = 0;
    while (true) {
      const int space = A.qcap - qn;
      int pos = pos0;
#pragma unroll
      for (int j = 0; j < FR; ++j) {
        if (!((passm >> j) & 1u)) continue;
        if (pos >= done && pos < done + space) {
          // the survivor fits in the queue: its index and kind (its
          // coefficients, determinant and couplings are filled at the flush)
          const int e = qn + pos - done;
          Q.idx[e] = idx[j];
          Q.kind_n[e] = (int)((kinds >> (2 * j)) & 3u);
        }
        ++pos;
      }
      const int adv = min(cnt - done, space);
      qn += adv;
      done += adv;
      if (done == cnt && more) break;
      // queue full, or the end of the list: flush
      __syncthreads();
      flush(qn);
      __syncthreads();
      qn = 0;
      if (done == cnt) break;
    }
    if (!more) break;
    __syncthreads();
  }

  // ---- fill_buffer_double ----
  const int s1 = (sp == 3) ? 1 : sp, s2 = (sp == 3) ? 2 : sp;
  double pt2[MAXNS];   // indexed with constants only (kept in registers)
#pragma unroll
  for (int st = 0; st < MAXNS; ++st) pt2[st] = 0.0;
  int64_t ncand = 0;
  bool overflow = false;
  int r_lo, r_hi;
  band_rows(M, r_lo, r_hi);
  for (int p1 = r_lo; p1 <= r_hi; ++p1) {
    int p2 = owned_j(p1, M);
    if (!p2) continue;
    if (sp != 3 && p2 <= p1) continue;
    if (borb_get(L.borb, p1, s1) || borb_get(L.borb, p2, s2)) continue;
    if (ban_get(L.ban1, M, p1, p2)) continue;
    const double *m = L.mat + (int64_t)NS * ((p1 - 1) + M * (p2 - 1));
    double sa = 0.0;
    for (int st = 0; st < NS; ++st) sa += fabs(m[st]);
    if (sa == 0.0) continue;
    double Hii = diag_fock(T, L.fock, sp, B.h1, p1, B.h2, p2);
    double min_e = 0.0;
#pragma unroll
    for (int st = 0; st < MAXNS; ++st) {
      if (st >= NS) break;
      double dE = A.E0[st] - Hii;
      double val = m[st] + m[st];
      double tmp = sqrt(dE * dE + val * val);
      if (dE < 0.0) tmp = -tmp;
      double e = 0.5 * (tmp - dE);
      pt2[st] += e;
      min_e = min(e, min_e);
    }
    ++ncand;
    if (min_e <= A.tau) {
      unsigned long long slot = atomicAdd(A.pool_count, 1ULL);
      if ((int64_t)slot < A.pool_cap) {
        uint64_t det[2 * NI];
#pragma unroll
        for (int w = 0; w < 2 * NI; ++w) det[w] = mask[w];
        set_orb<NI>(det, s1, p1);
        set_orb<NI>(det, s2, p2);
        A.pool_key[slot] = min_e;
#pragma unroll
        for (int w = 0; w < 2 * NI; ++w) A.pool_det[slot * 2 * NI + w] = det[w];
        A.pool_blk[slot] = A.blk_base + blockIdx.x;
      } else {
        overflow = true;
      }
    }
  }
  // ---- block reductions (fixed tree order) ----
  __syncthreads();
  if (tid == 0) L.flag[1] = 0;
  __syncthreads();
  if (overflow) atomicOr(&L.flag[1], 1);
#pragma unroll
  for (int st = 0; st < MAXNS; ++st)
    if (st < NS) red[st * SB + tid] = pt2[st];
  __syncthreads();
  for (int off = SB / 2; off > 0; off >>= 1) {
    if (tid < off)
      for (int st = 0; st < NS; ++st) red[st * SB + tid] += red[st * SB + tid + off];
    __syncthreads();
  }
  if (tid == 0)
    for (int st = 0; st < NS; ++st) A.blk_pt2[(int64_t)blockIdx.x * NS + st] = red[st * SB];
  __syncthreads();
  // counters: candidates and gathers
  int64_t *ired = reinterpret_cast<int64_t *>(red);
  ired[tid] = ncand;
  __syncthreads();
  for (int off = SB / 2; off > 0; off >>= 1) {
    if (tid < off) ired[tid] += ired[tid + off];
    __syncthreads();
  }
  if (tid == 0) A.blk_ncand[blockIdx.x] = ired[0];
  __syncthreads();
  ired[tid] = C.gath;
  __syncthreads();
  for (int off = SB / 2; off > 0; off >>= 1) {
    if (tid < off) ired[tid] += ired[tid + off];
    __syncthreads();
  }
  if (tid == 0) {
    A.blk_ngath[blockIdx.x] = ired[0];
    A.blk_status[blockIdx.x] = L.flag[1] ? 2 : 0;
    A.blk_ticks[blockIdx.x] = (int64_t)(wall_clock64() - t_start);
  }
}

// drop entries from given blocks and entries above tau; returns via count
template <int NI>
__global__ void pool_compact_kernel(const double *__restrict__ key, const uint64_t *__restrict__ det,
                                    const int *__restrict__ blk, int64_t n, double tau,
                                    const uint8_t *__restrict__ drop_blk, double *__restrict__ key2,
                                    uint64_t *__restrict__ det2, int *__restrict__ blk2,
                                    unsigned long long *__restrict__ count2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (key[i] > tau) continue;
    if (drop_blk && drop_blk[blk[i]]) continue;
    unsigned long long s = atomicAdd(count2, 1ULL);
    key2[s] = key[i];
    for (int w = 0; w < 2 * NI; ++w) det2[s * 2 * NI + w] = det[i * 2 * NI + w];
    blk2[s] = blk[i];
  }
}

}  // namespace

// --------------------------------------------------------------------------
// host driver
// --------------------------------------------------------------------------
namespace {

struct Pool {
  DBuf<double> key, key2;
  DBuf<uint64_t> det, det2;
  DBuf<int> blk, blk2;
  DBuf<unsigned long long> count;
  int64_t cap = 0;
};

// A list of the reference's selection tasks: task t enumerates the blocks of
// generator gen[t] (1-based) whose maskInd satisfies the subset filter
// subset[t] (0 = every block), as run_selection_slave does for one
// "i_generator subset" task (plugins/Full_CI_ZMQ/run_selection_slave.irp.f:46-106).
struct TaskList {
  std::vector<int> gen, subset;
};

template <int NI>
void run_selection(qpx_ctx *c, int NS, const double *coef, int n_gen, int n_sel, const double *E0, int64_t n_keep,
                   int subset, double *pt2_out, uint64_t *out_dets, double *out_vals, int64_t *n_out,
                   const uint8_t *keep_mask, double *block_pt2_out, int64_t *block_info_out,
                   const TaskList *tasks = nullptr, double *task_pt2_out = nullptr) {
  Wfn &W = c->W;
  Ham &H = c->H;
  hipStream_t st = c->stream;
  const int64_t n_det = W.n_det;
  const int M = H.M;
  // LDS budget
  const int nbits = (M * M + 31) / 32;
  // mat | fock | ban1, ban2, bannedOrb | flags, wavefront counts | reduction
  // scratch shared with the selector queue
  const size_t mat_bytes = sizeof(double) * (size_t)NS * M * M;
  const size_t red_bytes = sizeof(double) * SB * std::max(NS, 1);
  const size_t rest_bytes = sizeof(double) * 2 * (M + 1) + sizeof(uint32_t) * (2 * nbits + 16) + sizeof(int) * 12 + 8;
  // the coupling matrix goes to HBM when it leaves no room for a queue of
  // 64 entries in LDS (the reference keeps mat(N_states, M, M) on the heap
  // for any size, selection.irp.f:409)
  const bool gm = mat_bytes + rest_bytes + std::max(red_bytes, QueueLds::bytes(64, NS)) > (size_t)c->lds_per_block;
  const size_t lds_base = rest_bytes + (gm ? 0 : mat_bytes);
  QPX_REQUIRE(lds_base + red_bytes <= (size_t)c->lds_per_block, "selection: mo_num too large for the LDS layout");
  int qcap = 0;
  if ((size_t)c->lds_per_block > lds_base + QueueLds::bytes(0, NS))
    qcap = (int)std::min<size_t>(1024, ((size_t)c->lds_per_block - lds_base - QueueLds::bytes(0, NS)) /
                                           (QueueLds::bytes(1, NS) - QueueLds::bytes(0, NS)));
  QPX_REQUIRE(qcap >= 16, "selection: no LDS left for the selector queue");
  const size_t lds = lds_base + std::max(red_bytes, QueueLds::bytes(qcap, NS));
  auto kfun = gm ? select_block_kernel<NI, true> : select_block_kernel<NI, false>;
  QPX_HIP(hipFuncSetAttribute((const void *)kfun, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  // HBM coupling matrices: one slot per workgroup of a launch, launches of
  // at most mat_slots workgroups (about 1 GiB of slots)
  DBuf<double> mat_g;
  int64_t mat_slots = 0;
  if (gm) {
    mat_slots = std::max<int64_t>(256, std::min<int64_t>(4096, ((int64_t)1 << 30) / (int64_t)mat_bytes));
    mat_g.alloc((size_t)mat_slots * NS * M * M);
  }

  DBuf<double> coef_d;
  coef_d.upload(coef, (size_t)n_det * NS, st);
  // the M^4 table (shared with the sparse sigma path), built on first use
  const bool g4 = ensure_g4(c);
  Tabs T{H.V.p, H.hcore.p, H.Jd.p, H.JAd.p, H.fock_cs.p, M, H.npair, H.ref_energy, g4 ? H.G4.p : nullptr};

  // hole lists = occupied orbitals of each generator (full generator masks)
  std::vector<uint64_t> psi_h((size_t)n_det * 2 * NI);
  QPX_HIP(hipMemcpyAsync(psi_h.data(), W.psi.p, sizeof(uint64_t) * psi_h.size(), hipMemcpyDeviceToHost, st));
  QPX_HIP(hipStreamSynchronize(st));

  const int fragment_count = H.na * H.na;
  // all blocks in generator order
  std::vector<BlockDesc> all;
  std::vector<int> gen_first;  // first block of each generator
  std::vector<int> blk_task;   // task of each block (task lists only; -1 = none)
  // task lists: the tasks of each generator, as (first, count) into task_of
  std::vector<int> gt_first, gt_count, task_of;
  if (tasks) {
    gt_first.assign((size_t)n_gen + 2, 0);
    gt_count.assign((size_t)n_gen + 2, 0);
    for (int g : tasks->gen) gt_count[g]++;
    for (int g = 1; g <= n_gen; ++g) gt_first[g + 1] = gt_first[g] + gt_count[g];
    task_of.assign(tasks->gen.size(), 0);
    std::vector<int> fill(gt_first);
    for (size_t t = 0; t < tasks->gen.size(); ++t) task_of[fill[tasks->gen[t]]++] = (int)t;
  }
  for (int g = 1; g <= n_gen; ++g) {
    gen_first.push_back((int)all.size());
    if (tasks && !gt_count[g]) continue;   // only the listed generators get blocks
    const uint64_t *gd = &psi_h[(size_t)(g - 1) * 2 * NI];
    std::vector<int> hl[2];
    for (int s = 0; s < 2; ++s)
      for (int w = 0; w < NI; ++w) {
        uint64_t y = gd[s * NI + w];
        while (y) { hl[s].push_back(64 * w + __builtin_ctzll(y) + 1); y &= y - 1; }
      }
    int maskInd = -1;
    bool monoB = true;
    for (int s1 = 1; s1 <= 2; ++s1)
      for (int i1 = (int)hl[s1 - 1].size(); i1 >= 1; --i1) {
        int h1 = hl[s1 - 1][i1 - 1];
        for (int s2 = s1; s2 <= 2; ++s2) {
          int ib = (s1 == s2) ? i1 + 1 : 1;
          bool monoA = true;
          for (int i2 = (int)hl[s2 - 1].size(); i2 >= ib; --i2) {
            int h2 = hl[s2 - 1][i2 - 1];
            BlockDesc b{};
            b.gen = g;
            b.s1 = (int8_t)s1; b.s2 = (int8_t)s2;
            b.h1 = (int16_t)h1; b.h2 = (int16_t)h2;
            b.unbanA = (s1 != s2 && monoB) ? 1 : 0;
            b.unbanB = 0;
            if (s1 != s2 && monoA) { b.unbanB = 1; monoA = false; }
            ++maskInd;
            b.mask_ind = maskInd;
            bool keep = (subset == 0 || (maskInd % fragment_count) == subset - 1);
            if (keep_mask) keep = keep && keep_mask[all.size()] != 0;
            if (tasks) {
              int owner = -1;
              for (int q = 0; q < gt_count[g]; ++q) {
                const int t = task_of[gt_first[g] + q];
                const int sb = tasks->subset[t];
                if (sb == 0 || (maskInd % fragment_count) == sb - 1) owner = t;
              }
              keep = keep && owner >= 0;
              blk_task.push_back(owner);
            }
            b.gslot = keep ? 1 : 0;  // temporary: keep flag
            all.push_back(b);
          }
          if (s1 != s2) monoB = false;
        }
      }
  }
  gen_first.push_back((int)all.size());
  const int64_t n_blocks_total = (int64_t)all.size();

  Pool P;
  P.cap = std::max<int64_t>(8 * n_keep + (1 << 20), 1 << 22);
  P.key.alloc(P.cap); P.key2.alloc(P.cap);
  P.det.alloc((size_t)P.cap * 2 * NI); P.det2.alloc((size_t)P.cap * 2 * NI);
  P.blk.alloc(P.cap); P.blk2.alloc(P.cap);
  P.count.alloc(2);
  QPX_HIP(hipMemsetAsync(P.count.p, 0, sizeof(unsigned long long) * 2, st));
  int64_t pool_n = 0;
  // the pool keeps candidates with min-state e <= tau; nothing is kept when
  // no determinant is asked for
  double tau = n_keep > 0 ? 0.0 : -INFINITY;

  std::vector<double> blk_pt2_all((size_t)n_blocks_total * NS, 0.0);
  std::vector<int> blk_done(n_blocks_total, 0);
  std::vector<int64_t> blk_ticks_all(n_blocks_total, 0);
  int64_t ncand_total = 0, ngath_total = 0, nproc = 0;

  DBuf<int> gens_d, near_cnt_d, near_list_d, near_past_d;
  DBuf<uint64_t> near_sig_d;
  DBuf<double> fockg_d;
  DBuf<int64_t> near_off_d;
  DBuf<BlockDesc> blocks_d;
  DBuf<double> bpt2_d;
  DBuf<int> bstat_d;
  DBuf<int64_t> bncand_d, bngath_d, bticks_d;
  DBuf<uint8_t> drop_d;
  std::vector<uint8_t> drop(n_blocks_total, 0);

  auto compact = [&](bool with_drop) {
    // tighten tau to the n_keep-th best key (ties kept), drop entries above it
    // and entries of blocks that will be re-run
    std::vector<double> keys(pool_n);
    QPX_HIP(hipMemcpyAsync(keys.data(), P.key.p, sizeof(double) * pool_n, hipMemcpyDeviceToHost, st));
    QPX_HIP(hipStreamSynchronize(st));
    std::vector<int> blkh;
    if (with_drop) {
      blkh.resize(pool_n);
      QPX_HIP(hipMemcpyAsync(blkh.data(), P.blk.p, sizeof(int) * pool_n, hipMemcpyDeviceToHost, st));
      QPX_HIP(hipStreamSynchronize(st));
      std::vector<double> k2;
      k2.reserve(pool_n);
      for (int64_t i = 0; i < pool_n; ++i)
        if (!drop[blkh[i]]) k2.push_back(keys[i]);
      keys.swap(k2);
    }
    if ((int64_t)keys.size() >= n_keep && n_keep > 0) {
      std::nth_element(keys.begin(), keys.begin() + (n_keep - 1), keys.end());
      tau = std::min(tau, keys[n_keep - 1]);
    }
    if (with_drop) drop_d.upload(drop, st);
    QPX_HIP(hipMemsetAsync(P.count.p + 1, 0, sizeof(unsigned long long), st));
    int grid = (int)std::min<int64_t>((pool_n + 255) / 256, 65536);
    if (pool_n > 0)
      hipLaunchKernelGGL(pool_compact_kernel<NI>, dim3(grid), dim3(256), 0, st, P.key.p, P.det.p, P.blk.p, pool_n,
                         tau, with_drop ? drop_d.p : (const uint8_t *)nullptr, P.key2.p, P.det2.p, P.blk2.p,
                         P.count.p + 1);
    QPX_HIP(hipGetLastError());
    unsigned long long n2 = 0;
    QPX_HIP(hipMemcpyAsync(&n2, P.count.p + 1, sizeof(n2), hipMemcpyDeviceToHost, st));
    QPX_HIP(hipStreamSynchronize(st));
    std::swap(P.key.p, P.key2.p);
    std::swap(P.det.p, P.det2.p);
    std::swap(P.blk.p, P.blk2.p);
    pool_n = (int64_t)n2;
    unsigned long long pn = (unsigned long long)pool_n;
    QPX_HIP(hipMemcpyAsync(P.count.p, &pn, sizeof(pn), hipMemcpyHostToDevice, st));
    QPX_HIP(hipStreamSynchronize(st));
  };

  // generators with at least one block to run (all of them unless a keep
  // mask or a subset leaves some out): near lists are built for these only
  std::vector<int> active;
  active.reserve(n_gen);
  for (int g = 1; g <= n_gen; ++g)
    for (int b = gen_first[g - 1]; b < gen_first[g]; ++b)
      if (all[b].gslot) {
        active.push_back(g);
        break;
      }
  // batches of generators: bounded near-list size and block count
  size_t a0 = 0;
  int batch_gens = 1;  // the first generator alone establishes tau
  while (a0 < active.size()) {
    const size_t a1 = std::min(active.size(), a0 + (size_t)batch_gens);
    const int ng = (int)(a1 - a0);
    std::vector<int> gens(active.begin() + a0, active.begin() + a1);
    gens_d.upload(gens, st);
    near_cnt_d.alloc(ng);
    near_off_d.alloc(ng);
    near_past_d.alloc(ng);
    NearArgs NA{W.psi.p, n_det, n_sel, gens_d.p, ng, near_off_d.p, near_cnt_d.p, nullptr, near_past_d.p, nullptr};
    hipLaunchKernelGGL(near_kernel<NI>, dim3(ng), dim3(256), 0, st, NA, 0);
    QPX_HIP(hipGetLastError());
    std::vector<int> cnt(ng);
    QPX_HIP(hipMemcpyAsync(cnt.data(), near_cnt_d.p, sizeof(int) * ng, hipMemcpyDeviceToHost, st));
    QPX_HIP(hipStreamSynchronize(st));
    std::vector<int64_t> off(ng);
    int64_t tot = 0;
    for (int i = 0; i < ng; ++i) { off[i] = tot; tot += cnt[i]; }
    near_off_d.upload(off, st);
    near_list_d.alloc(std::max<int64_t>(tot, 1));
    near_sig_d.alloc(std::max<int64_t>(tot, 1));
    NA.list = near_list_d.p;
    NA.sig = near_sig_d.p;
    hipLaunchKernelGGL(near_kernel<NI>, dim3(ng), dim3(256), 0, st, NA, 1);
    QPX_HIP(hipGetLastError());
    fockg_d.alloc((size_t)ng * 2 * (M + 1));
    hipLaunchKernelGGL(fock_gen_kernel<NI>, dim3(ng), dim3(256), 0, st, W.psi.p, gens_d.p, T, fockg_d.p);
    QPX_HIP(hipGetLastError());
    // blocks of this batch
    std::vector<BlockDesc> bl;
    std::vector<int64_t> bl_gid;
    for (int q = 0; q < ng; ++q) {
      const int g = gens[q];
      for (int b = gen_first[g - 1]; b < gen_first[g]; ++b) {
        if (!all[b].gslot) continue;  // filtered out by subset / keep mask
        BlockDesc d = all[b];
        d.gslot = q;
        bl.push_back(d);
        bl_gid.push_back(b);
      }
    }
    std::vector<int64_t> pending(bl.size());
    std::iota(pending.begin(), pending.end(), 0);
    while (!pending.empty()) {
      std::vector<BlockDesc> run(pending.size());
      for (size_t i = 0; i < pending.size(); ++i) run[i] = bl[pending[i]];
      const int nb = (int)run.size();
      blocks_d.upload(run, st);
      bpt2_d.alloc((size_t)nb * NS);
      bstat_d.alloc(nb);
      bncand_d.alloc(nb);
      bngath_d.alloc(nb);
      bticks_d.alloc(nb);
      // global ids for the pool: the run index maps back through pending
      SelArgs SA{};
      SA.psi = W.psi.p; SA.coef = coef_d.p; SA.n_det = n_det; SA.n_sel = n_sel; SA.NS = NS;
      SA.na = H.na; SA.nb = H.nb;
      for (int s = 0; s < NS; ++s) SA.E0[s] = E0[s];
      SA.blocks = blocks_d.p; SA.n_blocks = nb;
      SA.near_off = near_off_d.p; SA.near_cnt = near_cnt_d.p; SA.near_list = near_list_d.p;
      SA.near_past = near_past_d.p; SA.fockg = fockg_d.p; SA.near_sig = near_sig_d.p;
      SA.pool_key = P.key.p; SA.pool_det = P.det.p; SA.pool_blk = P.blk.p; SA.pool_count = P.count.p;
      SA.pool_cap = P.cap; SA.tau = tau;
      SA.blk_pt2 = bpt2_d.p; SA.blk_status = bstat_d.p; SA.blk_ncand = bncand_d.p; SA.blk_ngath = bngath_d.p;
      SA.blk_ticks = bticks_d.p;
      // the pool records run-local block ids; they are translated to global
      // ids right after the launch
      SA.blk_base = 0;
      SA.qcap = qcap;
      // remember the mapping run index -> global block id for drops
      std::vector<int64_t> run_gid(nb);
      for (int i = 0; i < nb; ++i) run_gid[i] = bl_gid[pending[i]];
      int64_t pool_before = pool_n;
      if (!gm) {
        hipLaunchKernelGGL(kfun, dim3(nb), dim3(SB), lds, st, SA, T);
        QPX_HIP(hipGetLastError());
      } else {
        // launches of at most mat_slots workgroups, each on its own slot;
        // the pool's block ids stay run-local through blk_base
        SA.mat_g = mat_g.p;
        for (int off = 0; off < nb; off += (int)mat_slots) {
          const int nl = (int)std::min<int64_t>(mat_slots, nb - off);
          SelArgs S2 = SA;
          S2.blocks = SA.blocks + off;
          S2.n_blocks = nl;
          S2.blk_pt2 = SA.blk_pt2 + (size_t)off * NS;
          S2.blk_status = SA.blk_status + off;
          S2.blk_ncand = SA.blk_ncand + off;
          S2.blk_ngath = SA.blk_ngath + off;
          S2.blk_ticks = SA.blk_ticks + off;
          S2.blk_base = off;
          hipLaunchKernelGGL(kfun, dim3(nl), dim3(SB), lds, st, S2, T);
          QPX_HIP(hipGetLastError());
        }
      }
      std::vector<double> bp((size_t)nb * NS);
      std::vector<int> bs(nb);
      std::vector<int64_t> bc(nb), bg(nb), bt(nb);
      QPX_HIP(hipMemcpyAsync(bp.data(), bpt2_d.p, sizeof(double) * bp.size(), hipMemcpyDeviceToHost, st));
      QPX_HIP(hipMemcpyAsync(bs.data(), bstat_d.p, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
      QPX_HIP(hipMemcpyAsync(bc.data(), bncand_d.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, st));
      QPX_HIP(hipMemcpyAsync(bg.data(), bngath_d.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, st));
      QPX_HIP(hipMemcpyAsync(bt.data(), bticks_d.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, st));
      unsigned long long pc = 0;
      QPX_HIP(hipMemcpyAsync(&pc, P.count.p, sizeof(pc), hipMemcpyDeviceToHost, st));
      QPX_HIP(hipStreamSynchronize(st));
      pool_n = std::min<int64_t>((int64_t)pc, P.cap);
      // translate run-local block ids stored in the pool to global ids
      if (pool_n > pool_before) {
        std::vector<int> pb(pool_n - pool_before);
        QPX_HIP(hipMemcpyAsync(pb.data(), P.blk.p + pool_before, sizeof(int) * pb.size(), hipMemcpyDeviceToHost, st));
        QPX_HIP(hipStreamSynchronize(st));
        for (auto &x : pb) x = (int)run_gid[x];
        QPX_HIP(hipMemcpyAsync(P.blk.p + pool_before, pb.data(), sizeof(int) * pb.size(), hipMemcpyHostToDevice, st));
      }
      std::vector<int64_t> again;
      bool any_over = false;
      for (int i = 0; i < nb; ++i) {
        int64_t gid = run_gid[i];
        blk_ticks_all[gid] += bt[i];   // a re-run after a full pool adds its time
        if (bs[i] == 2) {
          any_over = true;
          drop[gid] = 1;
          again.push_back(pending[i]);
          continue;
        }
        for (int s = 0; s < NS; ++s) blk_pt2_all[(size_t)gid * NS + s] = bp[(size_t)i * NS + s];
        blk_done[gid] = (bs[i] == 0) ? 1 : 2;
        ncand_total += bc[i];
        ngath_total += bg[i];
        if (bs[i] == 0) ++nproc;
      }
      if (any_over) {
        unsigned long long pn = (unsigned long long)pool_n;
        QPX_HIP(hipMemcpyAsync(P.count.p, &pn, sizeof(pn), hipMemcpyHostToDevice, st));
        compact(true);
        for (auto gi : again) drop[bl_gid[gi]] = 0;
        QPX_REQUIRE(pool_n < P.cap / 2 || n_keep >= P.cap / 4, "selection pool cannot be compacted");
      } else if (pool_n > P.cap / 2) {
        compact(false);
      }
      pending.swap(again);
    }
    if (pool_n > n_keep) compact(false);
    // grow the batch: aim at ~ 32k blocks per launch
    const int glast = gens.back();
    int per_gen = (int)std::max<int64_t>(1, (gen_first[glast] - gen_first[glast - 1]));
    batch_gens = std::max(1, std::min(4096, 32768 / per_gen));
    a0 = a1;
  }
  // final: sort the pool by (key, determinant words) and keep n_keep
  std::vector<double> keys(pool_n);
  std::vector<uint64_t> dets((size_t)pool_n * 2 * NI);
  QPX_HIP(hipMemcpyAsync(keys.data(), P.key.p, sizeof(double) * pool_n, hipMemcpyDeviceToHost, st));
  QPX_HIP(hipMemcpyAsync(dets.data(), P.det.p, sizeof(uint64_t) * dets.size(), hipMemcpyDeviceToHost, st));
  QPX_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> ord(pool_n);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
    if (keys[a] != keys[b]) return keys[a] < keys[b];
    const uint64_t *da = &dets[(size_t)a * 2 * NI], *db = &dets[(size_t)b * 2 * NI];
    for (int w = 0; w < 2 * NI; ++w)
      if (da[w] != db[w]) return da[w] < db[w];
    return a < b;
  });
  int64_t nk = std::min<int64_t>(n_keep, pool_n);
  for (int64_t i = 0; i < nk; ++i) {
    out_vals[i] = keys[ord[i]];
    std::memcpy(out_dets + (size_t)i * 2 * NI, &dets[(size_t)ord[i] * 2 * NI], sizeof(uint64_t) * 2 * NI);
  }
  *n_out = nk;
  for (int s = 0; s < NS; ++s) {
    double acc = 0.0;
    for (int64_t b = 0; b < n_blocks_total; ++b) acc += blk_pt2_all[(size_t)b * NS + s];
    pt2_out[s] = acc;
  }
  if (block_pt2_out) std::memcpy(block_pt2_out, blk_pt2_all.data(), sizeof(double) * blk_pt2_all.size());
  if (tasks && task_pt2_out) {
    // per task, its blocks in block order (select_connected's sum)
    for (size_t t = 0; t < tasks->gen.size(); ++t) {
      const int g = tasks->gen[t];
      for (int st = 0; st < NS; ++st) {
        double acc = 0.0;
        for (int b = gen_first[g - 1]; b < gen_first[g]; ++b)
          if (blk_task[b] == (int)t) acc += blk_pt2_all[(size_t)b * NS + st];
        task_pt2_out[t * NS + st] = acc;
      }
    }
  }
  c->sel_ticks.swap(blk_ticks_all);
  if (block_info_out)
    for (int64_t b = 0; b < n_blocks_total; ++b) {
      const BlockDesc &d = all[b];
      int64_t *r = block_info_out + 6 * b;
      r[0] = d.gen; r[1] = d.s1; r[2] = d.h1; r[3] = d.s2; r[4] = d.h2; r[5] = blk_done[b];
    }
  c->sel_stats[0] = ncand_total;
  c->sel_stats[1] = nproc;
  c->sel_stats[2] = n_blocks_total;
  c->sel_stats[3] = ngath_total;
}

}  // namespace

static int select_impl(qpx_ctx *c, int n_states, const double *coef, int n_gen, int n_sel, const double *E0,
                       int64_t n_keep, int subset, double *pt2, uint64_t *out_dets, double *out_vals, int64_t *n_out,
                       const uint8_t *keep_mask, double *block_pt2, int64_t *block_info) {
  return qpx_guard(c, [&] {
    QPX_HIP(hipSetDevice(c->device));
    QPX_REQUIRE(c->W.valid, "no wavefunction set");
    QPX_REQUIRE(n_states >= 1 && n_states <= MAXNS, "n_states must be in 1..4");
    QPX_REQUIRE(n_gen >= 1 && n_gen <= n_sel && n_sel <= c->W.n_det, "need 1 <= n_gen <= n_sel <= n_det");
    QPX_REQUIRE(n_keep >= 0 && coef && E0 && pt2 && n_out, "bad selection arguments");
    QPX_REQUIRE(subset >= 0, "subset must be >= 0");
    switch (c->W.nint) {
      case 1: run_selection<1>(c, n_states, coef, n_gen, n_sel, E0, n_keep, subset, pt2, out_dets, out_vals, n_out, keep_mask, block_pt2, block_info); break;
      case 2: run_selection<2>(c, n_states, coef, n_gen, n_sel, E0, n_keep, subset, pt2, out_dets, out_vals, n_out, keep_mask, block_pt2, block_info); break;
      case 3: run_selection<3>(c, n_states, coef, n_gen, n_sel, E0, n_keep, subset, pt2, out_dets, out_vals, n_out, keep_mask, block_pt2, block_info); break;
      default: run_selection<4>(c, n_states, coef, n_gen, n_sel, E0, n_keep, subset, pt2, out_dets, out_vals, n_out, keep_mask, block_pt2, block_info); break;
    }
  });
}

extern "C" int qpx_select(qpx_ctx *c, int n_states, const double *coef, int n_gen, int n_sel, const double *E0,
                          int64_t n_keep, int subset, double *pt2, uint64_t *out_dets, double *out_vals,
                          int64_t *n_out) {
  return select_impl(c, n_states, coef, n_gen, n_sel, E0, n_keep, subset, pt2, out_dets, out_vals, n_out, nullptr,
                     nullptr, nullptr);
}

extern "C" int qpx_select_stats(qpx_ctx *c, int64_t *stats4) {
  return qpx_guard(c, [&] {
    for (int i = 0; i < 4; ++i) stats4[i] = c->sel_stats[i];
  });
}

// Block-level variant: keep_mask (one byte per block in the reference's
// enumeration order, all generators; NULL = all) restricts the enumerated
// blocks; block_pt2 (n_blocks * n_states) and block_info (n_blocks * 6:
// gen, s1, h1, s2, h2, status) receive per-block results (status 0 = not
// run, 1 = evaluated, 2 = owned by an earlier generator).
extern "C" int qpx_select_blocks(qpx_ctx *c, int n_states, const double *coef, int n_gen, int n_sel,
                                 const double *E0, int64_t n_keep, const uint8_t *keep_mask, double *pt2,
                                 uint64_t *out_dets, double *out_vals, int64_t *n_out, double *block_pt2,
                                 int64_t *block_info) {
  return select_impl(c, n_states, coef, n_gen, n_sel, E0, n_keep, 0, pt2, out_dets, out_vals, n_out, keep_mask,
                     block_pt2, block_info);
}

// benchmark hook: per-block workgroup durations of the last selection call
// (100 MHz wall-clock ticks, in block enumeration order; 0 for blocks not run)
extern "C" int qpx_debug_select_block_ticks(qpx_ctx *c, int64_t *ticks, int64_t n) {
  return qpx_guard(c, [&] {
    QPX_REQUIRE(ticks, "null output");
    QPX_REQUIRE(n == (int64_t)c->sel_ticks.size(), "n must equal the block count of the last selection call");
    std::copy(c->sel_ticks.begin(), c->sel_ticks.end(), ticks);
  });
}

// Validated task list: generators in 1..n_sel, subsets in 0..fragment_count,
// no generator listed twice with the same subset or with subset 0 and
// another subset.  Returns the largest generator index.
static int check_tasks(qpx_ctx *c, int n_sel, int n_tasks, const int *gens, const int *subsets, TaskList &T) {
  QPX_REQUIRE(n_tasks >= 1 && gens, "need at least one task");
  QPX_REQUIRE(n_sel >= 1 && n_sel <= c->W.n_det, "need 1 <= n_sel <= n_det");
  const int fc = c->H.na * c->H.na;
  T.gen.assign(gens, gens + n_tasks);
  T.subset.assign(n_tasks, 0);
  if (subsets) T.subset.assign(subsets, subsets + n_tasks);
  std::vector<std::pair<int, int>> key(n_tasks);
  for (int t = 0; t < n_tasks; ++t) {
    QPX_REQUIRE(T.gen[t] >= 1 && T.gen[t] <= n_sel, "generator indices must be in 1..n_sel");
    QPX_REQUIRE(T.subset[t] >= 0 && T.subset[t] <= fc, "subset must be in 0..fragment_count (n_alpha^2)");
    key[t] = {T.gen[t], T.subset[t]};
  }
  std::sort(key.begin(), key.end());
  for (int t = 1; t < n_tasks; ++t)
    QPX_REQUIRE(key[t].first != key[t - 1].first || (key[t].second != key[t - 1].second && key[t - 1].second != 0),
                "a generator is listed twice with the same subset, or with subset 0 and another subset");
  return key.back().first;
}

template <typename F>
static void dispatch_nint(int nint, F &&f) {
  switch (nint) {
    case 1: f(std::integral_constant<int, 1>()); break;
    case 2: f(std::integral_constant<int, 2>()); break;
    case 3: f(std::integral_constant<int, 3>()); break;
    default: f(std::integral_constant<int, 4>()); break;
  }
}

// EN-PT2 of each listed generator (select_connected's sum over its blocks)
extern "C" int qpx_select_generators(qpx_ctx *c, int n_states, const double *coef, int n_sel, const double *E0,
                                     int n_list, const int *gens, double *gen_pt2) {
  return qpx_guard(c, [&] {
    QPX_HIP(hipSetDevice(c->device));
    QPX_REQUIRE(c->W.valid, "no wavefunction set");
    QPX_REQUIRE(n_states >= 1 && n_states <= MAXNS, "n_states must be in 1..4");
    QPX_REQUIRE(n_list >= 1 && gens && gen_pt2 && coef && E0, "bad arguments");
    TaskList T;
    const int n_gen = check_tasks(c, n_sel, n_list, gens, nullptr, T);
    double pt2[MAXNS];
    uint64_t dummy_det[2 * NINT_MAX];
    double dummy_val;
    int64_t n_out = 0;
    dispatch_nint(c->W.nint, [&](auto ni) {
      run_selection<decltype(ni)::value>(c, n_states, coef, n_gen, n_sel, E0, 1, 0, pt2, dummy_det, &dummy_val,
                                         &n_out, nullptr, nullptr, nullptr, &T, gen_pt2);
    });
  });
}

// The reference's selection task farm over an explicit list of (generator,
// subset) tasks, with one merged buffer (zmq_selection.irp.f:54-100,
// run_selection_slave.irp.f:46-106)
extern "C" int qpx_select_tasks(qpx_ctx *c, int n_states, const double *coef, int n_sel, const double *E0,
                                int64_t n_keep, int n_tasks, const int *task_gen, const int *task_subset,
                                double *pt2, uint64_t *out_dets, double *out_vals, int64_t *n_out,
                                double *task_pt2) {
  return qpx_guard(c, [&] {
    QPX_HIP(hipSetDevice(c->device));
    QPX_REQUIRE(c->W.valid, "no wavefunction set");
    QPX_REQUIRE(n_states >= 1 && n_states <= MAXNS, "n_states must be in 1..4");
    QPX_REQUIRE(coef && E0 && pt2 && n_out && n_keep >= 0, "bad selection arguments");
    QPX_REQUIRE(n_keep == 0 || (out_dets && out_vals), "null output buffers");
    TaskList T;
    const int n_gen = check_tasks(c, n_sel, n_tasks, task_gen, task_subset, T);
    dispatch_nint(c->W.nint, [&](auto ni) {
      run_selection<decltype(ni)::value>(c, n_states, coef, n_gen, n_sel, E0, n_keep, 0, pt2, out_dets, out_vals,
                                         n_out, nullptr, nullptr, nullptr, &T, task_pt2);
    });
  });
}

// number of blocks of the first n_gen generators (for sizing block arrays)
extern "C" int qpx_select_count_blocks(qpx_ctx *c, int n_gen, int64_t *n_blocks) {
  return qpx_guard(c, [&] {
    QPX_REQUIRE(c->W.valid, "no wavefunction set");
    int na = c->H.na, nb = c->H.nb;
    int64_t per = (int64_t)na * (na - 1) / 2 + (int64_t)nb * (nb - 1) / 2 + (int64_t)na * nb;
    *n_blocks = per * n_gen;
  });
}
