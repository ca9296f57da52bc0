// psi_update.hip -- host-side wavefunction update after a selection step.
//
// make_s2_eigenfunction (src/Determinants/occ_pattern.irp.f:251-309): every
// determinant that shares an occupation pattern (doubly occupied orbitals,
// singly occupied orbitals) with a determinant of psi and is not in psi is
// added, so that psi spans whole spin multiplets.  The reference collects
// them per pattern from OpenMP threads (unordered); here the order is fixed:
// patterns are grouped by (number of open shells, alpha electrons among
// them), ascending; within a group the alpha subsets of the open shells come
// in lexicographic order and, for each subset, the patterns in lexicographic
// order of their words.  Host code: a few hundred thousand patterns are
// handled in well under a second, where the array version of this loop took
// seconds.
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <vector>
#include "../../include/qpx.h"

namespace {

constexpr int kMaxWords = 2 * 4;   // (docc, socc) or (alpha, beta), up to 4 words each
using Words = std::array<uint64_t, kMaxWords>;

bool less_words(const Words &a, const Words &b, int n) {
  for (int w = 0; w < n; ++w)
    if (a[w] != b[w]) return a[w] < b[w];
  return false;
}

bool equal_words(const Words &a, const Words &b, int n) {
  for (int w = 0; w < n; ++w)
    if (a[w] != b[w]) return false;
  return true;
}

}  // namespace

extern "C" int qpx_make_s2_eigenfunction(int nint, int64_t ndet, const uint64_t *psi_det, int n_alpha,
                                         int64_t max_out, uint64_t *out_dets, int64_t *n_out) {
  if (!n_out || nint < 1 || nint > 4 || ndet < 0 || (ndet > 0 && !psi_det) || n_alpha < 0) return 1;
  const int nw = 2 * nint;
  // determinants of psi, sorted, for the membership test
  std::vector<Words> dets((size_t)ndet), pats((size_t)ndet);
  for (int64_t k = 0; k < ndet; ++k) {
    Words d{}, p{};
    const uint64_t *x = psi_det + (size_t)k * nw;
    for (int w = 0; w < nint; ++w) {
      d[w] = x[w];
      d[nint + w] = x[nint + w];
      p[w] = x[w] & x[nint + w];          // doubly occupied
      p[nint + w] = x[w] ^ x[nint + w];   // singly occupied
    }
    dets[(size_t)k] = d;
    pats[(size_t)k] = p;
  }
  auto lt = [nw](const Words &a, const Words &b) { return less_words(a, b, nw); };
  auto eq = [nw](const Words &a, const Words &b) { return equal_words(a, b, nw); };
  std::sort(dets.begin(), dets.end(), lt);
  std::sort(pats.begin(), pats.end(), lt);
  pats.erase(std::unique(pats.begin(), pats.end(), eq), pats.end());

  struct Pat {
    int ns, k;
    size_t idx;
  };
  std::vector<Pat> order;
  order.reserve(pats.size());
  for (size_t i = 0; i < pats.size(); ++i) {
    int nd = 0, ns = 0;
    for (int w = 0; w < nint; ++w) {
      nd += __builtin_popcountll(pats[i][w]);
      ns += __builtin_popcountll(pats[i][nint + w]);
    }
    order.push_back({ns, n_alpha - nd, i});
  }
  std::stable_sort(order.begin(), order.end(), [](const Pat &a, const Pat &b) {
    return a.ns != b.ns ? a.ns < b.ns : a.k < b.k;
  });

  std::vector<uint64_t> res;
  std::vector<int> pos;
  for (size_t g0 = 0; g0 < order.size();) {
    size_t g1 = g0;
    while (g1 < order.size() && order[g1].ns == order[g0].ns && order[g1].k == order[g0].k) ++g1;
    const int ns = order[g0].ns, k = order[g0].k;
    if (ns > 0 && k >= 0 && k <= ns) {
      // open-shell orbital positions of every pattern of the group, ascending
      pos.assign((g1 - g0) * (size_t)ns, 0);
      for (size_t q = g0; q < g1; ++q) {
        const Words &p = pats[order[q].idx];
        int c = 0;
        for (int w = 0; w < nint; ++w)
          for (uint64_t y = p[nint + w]; y; y &= y - 1) pos[(q - g0) * ns + c++] = 64 * w + __builtin_ctzll(y);
      }
      // alpha subsets of size k of the ns open shells, lexicographic
      std::vector<int> comb(k);
      for (int j = 0; j < k; ++j) comb[j] = j;
      while (true) {
        for (size_t q = g0; q < g1; ++q) {
          const Words &p = pats[order[q].idx];
          Words d{};
          for (int w = 0; w < nint; ++w) d[w] = d[nint + w] = p[w];
          int cj = 0;
          for (int j = 0; j < ns; ++j) {
            const int o = pos[(q - g0) * ns + j];
            const bool alpha = cj < k && comb[cj] == j;
            if (alpha) ++cj;
            d[(alpha ? 0 : nint) + (o >> 6)] |= 1ULL << (o & 63);
          }
          if (!std::binary_search(dets.begin(), dets.end(), d, lt))
            res.insert(res.end(), d.begin(), d.begin() + nw);
        }
        // next subset
        int j = k - 1;
        while (j >= 0 && comb[j] == ns - k + j) --j;
        if (j < 0) break;
        ++comb[j];
        for (int l = j + 1; l < k; ++l) comb[l] = comb[l - 1] + 1;
      }
    }
    g0 = g1;
  }
  const int64_t n = (int64_t)(res.size() / nw);
  *n_out = n;
  if (!out_dets) return 0;
  if (max_out < n) return 2;
  if (n) std::memcpy(out_dets, res.data(), res.size() * sizeof(uint64_t));
  return 0;
}
