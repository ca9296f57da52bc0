/*
 * selection.c -- CIPSI selection with Epstein-Nesbet PT2 of the oracle
 * (test infrastructure only).
 *
 * Restates plugins/Full_CI_ZMQ/selection.irp.f:
 *   fragment_count                 :3-10
 *   get_mask_phase                 :24-47
 *   select_connected               :50-79
 *   get_phase_bi                   :82-100
 *   select_singles_and_doubles     :258-578
 *   fill_buffer_double             :582-644
 *   splash_pq                      :647-735
 *   get_d2 / get_d1 / get_d0       :738-853 / :856-1023 / :1028-1085
 *   past_d1 / past_d2              :1088-1101 / :1104-1126
 *   spot_isinwf                    :1130-1171
 *   bitstring_to_list_in_selection :1174-1200
 * plugins/Full_CI_ZMQ/selection_buffer.irp.f:2-138 (buffer add/sort),
 * src/Utils/sort.irp.f:241-270 (dsort; here a stable sort keyed on
 * (value, determinant words) so that ties are deterministic).
 * src/Determinants/determinants.irp.f:752-850 (apply_hole(s)/particle(s)).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "qp_oracle.h"

typedef struct {
  const oq_ham *H;
  oq_sel_ctx *S;
  int M, NS, nint;
  /* work arrays */
  char *banned;      /* M*M*2 */
  char *bannedOrb;   /* M*2 */
  double *mat;       /* NS*M*M */
  double *tmp_row, *tmp_row2; /* NS*M */
  char *lbanned;     /* M*2 */
  int *phasemask;    /* 2 * 64*nint */
  int64_t n_gathers, n_cand;
} selw;

#define BAN(w, i, j, sl) (w)->banned[((i) - 1) + (int64_t)(w)->M * ((j) - 1) + (int64_t)(w)->M * (w)->M * ((sl) - 1)]
#define BORB(w, o, s) (w)->bannedOrb[((o) - 1) + (w)->M * ((s) - 1)]
#define LBORB(w, o, s) (w)->lbanned[((o) - 1) + (w)->M * ((s) - 1)]
#define MAT(w, st, i, j) (w)->mat[(st) + (int64_t)(w)->NS * (((i) - 1) + (int64_t)(w)->M * ((j) - 1))]
#define PM(w, s, o) (w)->phasemask[((s) - 1) + 2 * ((o) - 1)]
#define P(n, s) p[(n) + 5 * ((s) - 1)]
#define Hh(n, s) h[(n) + 3 * ((s) - 1)]

static inline double I4(selw *w, int i, int j, int k, int l) {
  w->n_gathers++;
  return oq_get_integral(w->H, i, j, k, l);
}

static int list_sel(const bk_t *s, int nint, int *list) {
  /* bitstring_to_list_in_selection: 1-based orbital numbers */
  int n = 0;
  for (int i = 0; i < nint; ++i) {
    bk_t l = s[i];
    while (l) { list[n++] = __builtin_ctzll(l) + 1 + 64 * i; l &= l - 1; }
  }
  return n;
}

static void get_mask_phase(selw *w, const bk_t *det) {
  for (int s = 1; s <= 2; ++s) {
    int change = 0;
    for (int ni = 0; ni < w->nint; ++ni)
      for (int i = 0; i < 64; ++i) {
        if ((det[(s - 1) * w->nint + ni] >> i) & 1) change = !change;
        PM(w, s, ni * 64 + i + 1) = change ? 1 : 0;
      }
  }
}

static double phase_bi(selw *w, int s1, int s2, int h1, int p1, int h2, int p2) {
  int np = PM(w, s1, h1) + PM(w, s1, p1) + PM(w, s2, h2) + PM(w, s2, p2);
  if (p1 < h1) np++;
  if (p2 < h2) np++;
  if (s1 == s2 && (h1 > p1 ? h1 : p1) > (h2 < p2 ? h2 : p2)) np++;
  return (np & 1) ? -1.0 : 1.0;
}

static int apply_particles(const bk_t *det, int s1, int p1, int s2, int p2, bk_t *res, int nint) {
  memcpy(res, det, sizeof(bk_t) * 2 * nint);
  if (p1 != 0) {
    int ii = (p1 - 1) >> 6, pos = (p1 - 1) & 63;
    if ((det[(s1 - 1) * nint + ii] >> pos) & 1) return 0;
    res[(s1 - 1) * nint + ii] |= ((bk_t)1) << pos;
  }
  int ii = (p2 - 1) >> 6, pos = (p2 - 1) & 63;
  if ((det[(s2 - 1) * nint + ii] >> pos) & 1) return 0;
  res[(s2 - 1) * nint + ii] |= ((bk_t)1) << pos;
  return 1;
}

static int apply_hole(const bk_t *det, int s1, int h1, bk_t *res, int nint) {
  memcpy(res, det, sizeof(bk_t) * 2 * nint);
  int ii = (h1 - 1) >> 6, pos = (h1 - 1) & 63;
  if (!((det[(s1 - 1) * nint + ii] >> pos) & 1)) return 0;
  res[(s1 - 1) * nint + ii] &= ~(((bk_t)1) << pos);
  return 1;
}

static inline void addc(selw *w, int i, int j, const double *coefs, double hij) {
  for (int st = 0; st < w->NS; ++st) MAT(w, st, i, j) += coefs[st] * hij;
}

static const int turn2[3] = {0, 2, 1};
/* turn3(a,i): complementary pair of {1,2,3} \ {i} */
static int turn3(int a, int i) {
  static const int t[3][2] = {{2, 3}, {1, 3}, {1, 2}};
  return t[i - 1][a - 1];
}
/* turn2d(a,i,j): complementary pair of {1,2,3,4} \ {i,j} */
static int turn2d(int a, int i, int j) {
  int c[2], n = 0;
  for (int k = 1; k <= 4; ++k)
    if (k != i && k != j) c[n++] = k;
  return c[a - 1];
}

static void get_d2(selw *w, const bk_t *gen, const int *h, const int *p, int sp, const double *coefs) {
  (void)gen;
  int bant = 1;
  int tip = P(0, 1) * P(0, 2);
  int ma = sp;
  if (P(0, 1) > P(0, 2)) ma = 1;
  if (P(0, 1) < P(0, 2)) ma = 2;
  int mi = (ma % 2) + 1;
  if (sp == 3) {
    if (ma == 2) bant = 2;
    if (tip == 3) {
      int puti = P(1, mi);
      for (int i = 1; i <= 3; ++i) {
        int putj = P(i, ma);
        if (BAN(w, putj, puti, bant)) continue;
        int p1 = P(turn3(1, i), ma), p2 = P(turn3(2, i), ma);
        int h1 = Hh(1, ma), h2 = Hh(2, ma);
        double hij = (I4(w, p1, p2, h1, h2) - I4(w, p2, p1, h1, h2)) * phase_bi(w, ma, ma, h1, p1, h2, p2);
        if (ma == 1) addc(w, putj, puti, coefs, hij);
        else addc(w, puti, putj, coefs, hij);
      }
    } else {
      int h1 = Hh(1, 1), h2 = Hh(1, 2);
      for (int j = 1; j <= 2; ++j) {
        int putj = P(j, 2), p2 = P(turn2[j], 2);
        for (int i = 1; i <= 2; ++i) {
          int puti = P(i, 1);
          if (BAN(w, puti, putj, bant)) continue;
          int p1 = P(turn2[i], 1);
          double hij = I4(w, p1, p2, h1, h2) * phase_bi(w, 1, 2, h1, p1, h2, p2);
          addc(w, puti, putj, coefs, hij);
        }
      }
    }
  } else {
    if (tip == 0) {
      int h1 = Hh(1, ma), h2 = Hh(2, ma);
      for (int i = 1; i <= 3; ++i) {
        int puti = P(i, ma);
        for (int j = i + 1; j <= 4; ++j) {
          int putj = P(j, ma);
          if (BAN(w, puti, putj, 1)) continue;
          int p1 = P(turn2d(1, i, j), ma), p2 = P(turn2d(2, i, j), ma);
          double hij = (I4(w, p1, p2, h1, h2) - I4(w, p2, p1, h1, h2)) * phase_bi(w, ma, ma, h1, p1, h2, p2);
          addc(w, puti, putj, coefs, hij);
        }
      }
    } else if (tip == 3) {
      int h1 = Hh(1, mi), h2 = Hh(1, ma), p1 = P(1, mi);
      for (int i = 1; i <= 3; ++i) {
        int puti = P(turn3(1, i), ma), putj = P(turn3(2, i), ma);
        if (BAN(w, puti, putj, 1)) continue;
        int p2 = P(i, ma);
        double hij = I4(w, p1, p2, h1, h2) * phase_bi(w, mi, ma, h1, p1, h2, p2);
        addc(w, puti < putj ? puti : putj, puti < putj ? putj : puti, coefs, hij);
      }
    } else {
      int puti = P(1, sp), putj = P(2, sp);
      if (!BAN(w, puti, putj, 1)) {
        int p1 = P(1, mi), p2 = P(2, mi), h1 = Hh(1, mi), h2 = Hh(2, mi);
        double hij = (I4(w, p1, p2, h1, h2) - I4(w, p2, p1, h1, h2)) * phase_bi(w, mi, mi, h1, p1, h2, p2);
        addc(w, puti, putj, coefs, hij);
      }
    }
  }
}

static void get_d1(selw *w, const bk_t *gen, const int *h, const int *p, int sp, const double *coefs,
                   const bk_t *mask) {
  const int M = w->M, NS = w->NS;
  memcpy(w->lbanned, w->bannedOrb, 2 * M);
  for (int i = 1; i <= P(0, 1); ++i) LBORB(w, P(i, 1), 1) = 1;
  for (int i = 1; i <= P(0, 2); ++i) LBORB(w, P(i, 2), 2) = 1;
  int ma = 1;
  if (P(0, 2) >= 2) ma = 2;
  int mi = turn2[ma];
  int bant = 1;
  double *tr = w->tmp_row, *tr2 = w->tmp_row2;
#define TR(a, st, o) a[(st) + NS * ((o) - 1)]
  if (sp == 3) {
    if (ma == 2) bant = 2;
    int puti = P(1, mi), hfix = Hh(1, ma), p1 = P(1, ma), p2 = P(2, ma);
    if (!BORB(w, puti, mi)) {
      memset(tr, 0, sizeof(double) * NS * M);
      for (int putj = 1; putj < hfix; ++putj) {
        if (LBORB(w, putj, ma) || BAN(w, putj, puti, bant)) continue;
        double hij = (I4(w, p1, p2, putj, hfix) - I4(w, p2, p1, putj, hfix)) * phase_bi(w, ma, ma, putj, p1, hfix, p2);
        for (int st = 0; st < NS; ++st) TR(tr, st, putj) += hij * coefs[st];
      }
      for (int putj = hfix + 1; putj <= M; ++putj) {
        if (LBORB(w, putj, ma) || BAN(w, putj, puti, bant)) continue;
        double hij = (I4(w, p1, p2, hfix, putj) - I4(w, p2, p1, hfix, putj)) * phase_bi(w, ma, ma, hfix, p1, putj, p2);
        for (int st = 0; st < NS; ++st) TR(tr, st, putj) += hij * coefs[st];
      }
      for (int o = 1; o <= M; ++o)
        for (int st = 0; st < NS; ++st) {
          if (ma == 1) MAT(w, st, o, puti) += TR(tr, st, o);
          else MAT(w, st, puti, o) += TR(tr, st, o);
        }
    }
    int pfix = P(1, mi);
    memset(tr, 0, sizeof(double) * NS * M);
    memset(tr2, 0, sizeof(double) * NS * M);
    for (int pu = 1; pu <= M; ++pu) {
      if (LBORB(w, pu, mi)) continue;
      int putj = p1;
      if (!BAN(w, putj, pu, bant)) {
        double hij = I4(w, p2, pfix, hfix, pu) * phase_bi(w, ma, mi, hfix, p2, pu, pfix);
        for (int st = 0; st < NS; ++st) TR(tr, st, pu) += hij * coefs[st];
      }
      putj = p2;
      if (!BAN(w, putj, pu, bant)) {
        double hij = I4(w, p1, pfix, hfix, pu) * phase_bi(w, ma, mi, hfix, p1, pu, pfix);
        for (int st = 0; st < NS; ++st) TR(tr2, st, pu) += hij * coefs[st];
      }
    }
    for (int o = 1; o <= M; ++o)
      for (int st = 0; st < NS; ++st) {
        if (mi == 1) { MAT(w, st, o, p1) += TR(tr, st, o); MAT(w, st, o, p2) += TR(tr2, st, o); }
        else { MAT(w, st, p1, o) += TR(tr, st, o); MAT(w, st, p2, o) += TR(tr2, st, o); }
      }
  } else {
    if (P(0, ma) == 3) {
      for (int i = 1; i <= 3; ++i) {
        int hfix = Hh(1, ma), puti = P(i, ma), p1 = P(turn3(1, i), ma), p2 = P(turn3(2, i), ma);
        memset(tr, 0, sizeof(double) * NS * M);
        for (int putj = 1; putj < hfix; ++putj) {
          if (LBORB(w, putj, ma) || BAN(w, puti, putj, 1)) continue;
          double hij = (I4(w, p1, p2, putj, hfix) - I4(w, p2, p1, putj, hfix)) * phase_bi(w, ma, ma, putj, p1, hfix, p2);
          for (int st = 0; st < NS; ++st) TR(tr, st, putj) += hij * coefs[st];
        }
        for (int putj = hfix + 1; putj <= M; ++putj) {
          if (LBORB(w, putj, ma) || BAN(w, puti, putj, 1)) continue;
          double hij = (I4(w, p1, p2, hfix, putj) - I4(w, p2, p1, hfix, putj)) * phase_bi(w, ma, ma, hfix, p1, putj, p2);
          for (int st = 0; st < NS; ++st) TR(tr, st, putj) += hij * coefs[st];
        }
        for (int o = 1; o < puti; ++o)
          for (int st = 0; st < NS; ++st) MAT(w, st, o, puti) += TR(tr, st, o);
        for (int o = puti; o <= M; ++o)
          for (int st = 0; st < NS; ++st) MAT(w, st, puti, o) += TR(tr, st, o);
      }
    } else {
      int hfix = Hh(1, mi), pfix = P(1, mi), p1 = P(1, ma), p2 = P(2, ma);
      memset(tr, 0, sizeof(double) * NS * M);
      memset(tr2, 0, sizeof(double) * NS * M);
      for (int puti = 1; puti <= M; ++puti) {
        if (LBORB(w, puti, ma)) continue;
        int putj = p2;
        if (!BAN(w, puti, putj, 1)) {
          double hij = I4(w, pfix, p1, hfix, puti) * phase_bi(w, mi, ma, hfix, pfix, puti, p1);
          for (int st = 0; st < NS; ++st) TR(tr, st, puti) += hij * coefs[st];
        }
        putj = p1;
        if (!BAN(w, puti, putj, 1)) {
          double hij = I4(w, pfix, p2, hfix, puti) * phase_bi(w, mi, ma, hfix, pfix, puti, p2);
          for (int st = 0; st < NS; ++st) TR(tr2, st, puti) += hij * coefs[st];
        }
      }
      for (int o = 1; o < p2; ++o)
        for (int st = 0; st < NS; ++st) MAT(w, st, o, p2) += TR(tr, st, o);
      for (int o = p2; o <= M; ++o)
        for (int st = 0; st < NS; ++st) MAT(w, st, p2, o) += TR(tr, st, o);
      for (int o = 1; o < p1; ++o)
        for (int st = 0; st < NS; ++st) MAT(w, st, o, p1) += TR(tr2, st, o);
      for (int o = p1; o <= M; ++o)
        for (int st = 0; st < NS; ++st) MAT(w, st, p1, o) += TR(tr2, st, o);
    }
  }
#undef TR
  /* singles folded into this block */
  int s1, s2;
  if (sp == 3) { s1 = 1; s2 = 2; } else { s1 = sp; s2 = sp; }
  bk_t det[2 * OQ_NINT_MAX];
  for (int i1 = 1; i1 <= P(0, s1); ++i1) {
    int ib = (s1 == s2) ? i1 + 1 : 1;
    for (int i2 = ib; i2 <= P(0, s2); ++i2) {
      int p1 = P(i1, s1), p2 = P(i2, s2);
      if (BORB(w, p1, s1) || BORB(w, p2, s2) || BAN(w, p1, p2, 1)) continue;
      apply_particles(mask, s1, p1, s2, p2, det, w->nint);
      double hij = oq_i_H_j(w->H, gen, det);
      w->n_gathers += 4;
      addc(w, p1, p2, coefs, hij);
    }
  }
}

static void get_d0(selw *w, const bk_t *gen, const int *h, const int *p, int sp, const double *coefs,
                   const bk_t *mask) {
  (void)h;
  const int M = w->M;
  int bant = 1;
  bk_t det[2 * OQ_NINT_MAX];
  if (sp == 3) {
    int h1 = P(1, 1), h2 = P(1, 2);
    for (int p1 = 1; p1 <= M; ++p1) {
      if (BORB(w, p1, 1)) continue;
      for (int p2 = 1; p2 <= M; ++p2) {
        if (BORB(w, p2, 2)) continue;
        if (BAN(w, p1, p2, bant)) continue;
        double hij;
        if (p1 == h1 || p2 == h2) {
          apply_particles(mask, 1, p1, 2, p2, det, w->nint);
          hij = oq_i_H_j(w->H, gen, det);
          w->n_gathers += 4;
        } else {
          hij = I4(w, p1, p2, h1, h2) * phase_bi(w, 1, 2, h1, p1, h2, p2);
        }
        addc(w, p1, p2, coefs, hij);
      }
    }
  } else {
    int p1 = P(1, sp), p2 = P(2, sp);
    for (int puti = 1; puti <= M; ++puti) {
      if (BORB(w, puti, sp)) continue;
      for (int putj = puti + 1; putj <= M; ++putj) {
        if (BORB(w, putj, sp)) continue;
        if (BAN(w, puti, putj, bant)) continue;
        double hij;
        if (puti == p1 || putj == p2 || puti == p2 || putj == p1) {
          apply_particles(mask, sp, puti, sp, putj, det, w->nint);
          hij = oq_i_H_j(w->H, gen, det);
          w->n_gathers += 4;
        } else {
          hij = (I4(w, p1, p2, puti, putj) - I4(w, p2, p1, puti, putj)) * phase_bi(w, sp, sp, puti, p1, putj, p2);
        }
        addc(w, puti, putj, coefs, hij);
      }
    }
  }
}

static void past_d1(selw *w, const int *p) {
  for (int s = 1; s <= 2; ++s)
    for (int i = 1; i <= P(0, s); ++i) BORB(w, P(i, s), s) = 1;
}

static void past_d2(selw *w, const int *p, int sp) {
  if (sp == 3) {
    for (int i = 1; i <= P(0, 1); ++i)
      for (int j = 1; j <= P(0, 2); ++j) BAN(w, P(i, 1), P(j, 2), 1) = 1;
  } else {
    for (int i = 1; i <= P(0, sp); ++i)
      for (int j = 1; j < i; ++j) {
        BAN(w, P(j, sp), P(i, sp), 1) = 1;
        BAN(w, P(i, sp), P(j, sp), 1) = 1;
      }
  }
}

/* returns fullMatch */
static int spot_isinwf(selw *w, const bk_t *mask, const bk_t *dets, const int *idx, int N, int i_gen) {
  int nint = w->nint;
  for (int i = 0; i < N; ++i) {
    const bk_t *d = dets + (int64_t)i * 2 * nint;
    int ok = 1;
    for (int j = 0; j < 2 * nint; ++j)
      if ((d[j] & mask[j]) != mask[j]) { ok = 0; break; }
    if (!ok) continue;
    if (idx[i] < i_gen) return 1;
    bk_t my[2 * OQ_NINT_MAX];
    for (int j = 0; j < 2 * nint; ++j) my[j] = d[j] & ~mask[j];
    int list[4 * 64 * OQ_NINT_MAX];
    int na = list_sel(my, nint, list);
    list_sel(my + nint, nint, list + na);
    BAN(w, list[0], list[1], 1) = 1;
  }
  return 0;
}

static void splash_pq(selw *w, const bk_t *mask, int sp, const bk_t *dets, const int *idx, int N_sel, int i_gen) {
  const int M = w->M, NS = w->NS, nint = w->nint;
  memset(w->mat, 0, sizeof(double) * NS * M * M);
  bk_t neg[2 * OQ_NINT_MAX], mob[2 * OQ_NINT_MAX], per[2 * OQ_NINT_MAX];
  for (int i = 0; i < 2 * nint; ++i) neg[i] = ~mask[i];
  for (int i = 0; i < N_sel; ++i) {
    const bk_t *d = dets + (int64_t)i * 2 * nint;
    int nt = 0;
    for (int j = 0; j < 2 * nint; ++j) { mob[j] = neg[j] & d[j]; nt += __builtin_popcountll(mob[j]); }
    if (nt > 4) continue;
    if (idx[i] == i_gen) {
      if (sp == 3) {
        for (int j = 1; j <= M; ++j)
          for (int k = 1; k <= M; ++k) BAN(w, j, k, 2) = BAN(w, k, j, 1);
      } else {
        for (int k = 1; k <= M; ++k)
          for (int l = k + 1; l <= M; ++l) BAN(w, l, k, 1) = BAN(w, k, l, 1);
      }
    }
    int p[10], h[6];
    memset(p, 0, sizeof(p));
    memset(h, 0, sizeof(h));
    int tmpl[4 * 64 * OQ_NINT_MAX];
    int n1 = list_sel(mob, nint, tmpl);
    P(0, 1) = n1;
    for (int k = 0; k < n1 && k < 4; ++k) P(k + 1, 1) = tmpl[k];
    int n2 = list_sel(mob + nint, nint, tmpl);
    P(0, 2) = n2;
    for (int k = 0; k < n2 && k < 4; ++k) P(k + 1, 2) = tmpl[k];
    for (int j = 0; j < 2 * nint; ++j) per[j] = mask[j] & ~d[j];
    n1 = list_sel(per, nint, tmpl);
    Hh(0, 1) = n1;
    for (int k = 0; k < n1 && k < 2; ++k) Hh(k + 1, 1) = tmpl[k];
    n2 = list_sel(per + nint, nint, tmpl);
    Hh(0, 2) = n2;
    for (int k = 0; k < n2 && k < 2; ++k) Hh(k + 1, 2) = tmpl[k];
    if (idx[i] >= i_gen) {
      get_mask_phase(w, w->S->psi_det_sorted + (int64_t)(idx[i] - 1) * 2 * nint);
      const double *coefs = w->S->coef_transp + (int64_t)(idx[i] - 1) * NS;
      w->n_gathers += NS;
      if (nt == 4) get_d2(w, d, h, p, sp, coefs);
      else if (nt == 3) get_d1(w, d, h, p, sp, coefs, mask);
      else get_d0(w, d, h, p, sp, coefs, mask);
    } else {
      if (nt == 4) past_d2(w, p, sp);
      if (nt == 3) past_d1(w, p);
    }
  }
}

/* ---- selection buffer ---- */
oq_sel_buffer *oq_sel_buffer_create(int N, int siz, int nint) {
  oq_sel_buffer *b = (oq_sel_buffer *)calloc(1, sizeof(oq_sel_buffer));
  if (siz < 1) siz = 1;
  b->N = N; b->cap = siz; b->nint = nint; b->cur = 0; b->mini = 0.0;
  b->det = (bk_t *)calloc((size_t)siz * 2 * nint, sizeof(bk_t));
  b->val = (double *)calloc(siz, sizeof(double));
  return b;
}
void oq_sel_buffer_destroy(oq_sel_buffer *b) {
  if (!b) return;
  free(b->det); free(b->val); free(b);
}

static int g_sort_nint;
static const double *g_sort_val;
static const bk_t *g_sort_det;
static int cmp_entry(const void *x, const void *y) {
  int a = *(const int *)x, b = *(const int *)y;
  double va = g_sort_val[a], vb = g_sort_val[b];
  if (va != vb) return va < vb ? -1 : 1;
  const bk_t *da = g_sort_det + (int64_t)a * 2 * g_sort_nint, *db = g_sort_det + (int64_t)b * 2 * g_sort_nint;
  for (int k = 0; k < 2 * g_sort_nint; ++k)
    if (da[k] != db[k]) return da[k] < db[k] ? -1 : 1;
  return a < b ? -1 : (a > b);
}

void oq_sel_buffer_sort(oq_sel_buffer *b) {
  if (b->cur == 0) return;
  int nmwen = b->cur < b->N ? b->cur : b->N;
  int *ord = (int *)malloc(sizeof(int) * b->cur);
  for (int i = 0; i < b->cur; ++i) ord[i] = i;
#pragma omp critical(oq_sort)
  {
    g_sort_nint = b->nint; g_sort_val = b->val; g_sort_det = b->det;
    qsort(ord, b->cur, sizeof(int), cmp_entry);
  }
  bk_t *nd = (bk_t *)calloc((size_t)b->cap * 2 * b->nint, sizeof(bk_t));
  double *nv = (double *)calloc(b->cap, sizeof(double));
  for (int i = 0; i < b->cur; ++i) {
    nv[i] = b->val[ord[i]];
    if (i < nmwen) memcpy(nd + (int64_t)i * 2 * b->nint, b->det + (int64_t)ord[i] * 2 * b->nint, sizeof(bk_t) * 2 * b->nint);
  }
  free(b->det); free(b->val); free(ord);
  b->det = nd; b->val = nv;
  /* val(N) after the sort; entries past cur are zero as in the reference */
  double vN = (b->N - 1 < b->cur) ? b->val[b->N - 1] : 0.0;
  b->mini = b->mini < vN ? b->mini : vN;
  b->cur = nmwen;
}

static void buffer_add(oq_sel_buffer *b, const bk_t *det, double val) {
  if (val <= b->mini) {
    memcpy(b->det + (int64_t)b->cur * 2 * b->nint, det, sizeof(bk_t) * 2 * b->nint);
    b->val[b->cur] = val;
    b->cur++;
    if (b->cur == b->cap) oq_sel_buffer_sort(b);
  }
}

static void fill_buffer_double(selw *w, const bk_t *gen, int sp, int h1, int h2, const double *fock,
                               const double *E0, double *pt2, oq_sel_buffer *buf, double *blk_e, int64_t *blk_c) {
  const int M = w->M, NS = w->NS, nint = w->nint;
  int s1, s2;
  if (sp == 3) { s1 = 1; s2 = 2; } else { s1 = sp; s2 = sp; }
  bk_t mask[2 * OQ_NINT_MAX], det[2 * OQ_NINT_MAX], tmp[2 * OQ_NINT_MAX];
  apply_hole(gen, s1, h1, tmp, nint);
  apply_hole(tmp, s2, h2, mask, nint);
  for (int p1 = 1; p1 <= M; ++p1) {
    if (BORB(w, p1, s1)) continue;
    int ib = (sp != 3) ? p1 + 1 : 1;
    for (int p2 = ib; p2 <= M; ++p2) {
      if (BORB(w, p2, s2)) continue;
      if (BAN(w, p1, p2, 1)) continue;
      double sa = 0.0;
      for (int st = 0; st < NS; ++st) sa += fabs(MAT(w, st, p1, p2));
      if (sa == 0.0) continue;
      apply_particles(mask, s1, p1, s2, p2, det, nint);
      double Hii = oq_diag_H_mat_elem_fock(w->H, gen, det, fock);
      double min_e = 0.0;
      for (int st = 0; st < NS; ++st) {
        double dE = E0[st] - Hii;
        double val = MAT(w, st, p1, p2) + MAT(w, st, p1, p2);
        double t = sqrt(dE * dE + val * val);
        if (dE < 0.0) t = -t;
        double e = 0.5 * (t - dE);
        pt2[st] += e;
        if (st == 0) *blk_e += e;
        min_e = e < min_e ? e : min_e;
      }
      w->n_cand++;
      (*blk_c)++;
      if (min_e <= buf->mini) buffer_add(buf, det, min_e);
    }
  }
}

static void select_singles_and_doubles(selw *w, int i_generator, const bk_t *hole_mask,
                                       const bk_t *particle_mask, const double *fock, const double *E0,
                                       double *pt2, oq_sel_buffer *buf, int subset) {
  oq_sel_ctx *S = w->S;
  const oq_bilinear *B = S->B;
  const int nint = w->nint, M = w->M;
  const bk_t *gen = S->psi_det_sorted + (int64_t)(i_generator - 1) * 2 * nint;
  bk_t hole[2 * OQ_NINT_MAX], particle[2 * OQ_NINT_MAX];
  for (int k = 0; k < 2 * nint; ++k) { hole[k] = gen[k] & hole_mask[k]; particle[k] = ~gen[k] & particle_mask[k]; }
  int hole_list[2][64 * OQ_NINT_MAX], particle_list[2][64 * OQ_NINT_MAX], N_holes[2], N_particles[2];
  for (int s = 0; s < 2; ++s) {
    N_holes[s] = list_sel(hole + s * nint, nint, hole_list[s]);
    N_particles[s] = list_sel(particle + s * nint, nint, particle_list[s]);
  }
  int64_t N_det = S->n_det;
  /* prefilter via the bilinear structure (selection.irp.f:320-366) */
  int *indices = (int *)malloc(sizeof(int) * (N_det + 1));
  int maxab = B->n_alpha > B->n_beta ? B->n_alpha : B->n_beta;
  int *exc_degree = (int *)malloc(sizeof(int) * (maxab + 1));
  int64_t k = 0;
  for (int i = 1; i <= B->n_alpha; ++i)
    exc_degree[i - 1] = oq_excitation_degree_spin(B->alpha_unique + (int64_t)(i - 1) * nint, gen, nint);
  for (int j = 1; j <= B->n_beta; ++j) {
    int nt = oq_excitation_degree_spin(B->beta_unique + (int64_t)(j - 1) * nint, gen + nint, nint);
    if (nt > 2) continue;
    for (int l_a = B->columns_loc[j - 1]; l_a < B->columns_loc[j]; ++l_a) {
      int i = B->rows[l_a - 1];
      if (nt + exc_degree[i - 1] <= 4) indices[k++] = B->order[l_a - 1];
    }
  }
  for (int i = 1; i <= B->n_beta; ++i)
    exc_degree[i - 1] = oq_excitation_degree_spin(B->beta_unique + (int64_t)(i - 1) * nint, gen + nint, nint);
  for (int j = 1; j <= B->n_alpha; ++j) {
    int nt = oq_excitation_degree_spin(B->alpha_unique + (int64_t)(j - 1) * nint, gen, nint);
    if (nt > 1) continue;
    for (int l_a = B->transp_rows_loc[j - 1]; l_a < B->transp_rows_loc[j]; ++l_a) {
      int i = B->transp_columns[l_a - 1];
      if (exc_degree[i - 1] < 3) continue;
      if (nt + exc_degree[i - 1] <= 4) indices[k++] = B->order[B->transp_order[l_a - 1] - 1];
    }
  }
  free(exc_degree);
  int64_t nmax = k;
  /* isort */
  for (int64_t a = 1; a < nmax; ++a) {
    int v = indices[a];
    int64_t b = a - 1;
    while (b >= 0 && indices[b] > v) { indices[b + 1] = indices[b]; --b; }
    indices[b + 1] = v;
  }
  int *preint = (int *)malloc(sizeof(int) * (nmax + 1));
  int *prefull = (int *)malloc(sizeof(int) * (nmax + 1));
  bk_t *preint_det = (bk_t *)malloc(sizeof(bk_t) * 2 * nint * (nmax + 1));
  int npre = 0, nprefull = 0;
  bk_t neg[2 * OQ_NINT_MAX];
  for (int j = 0; j < 2 * nint; ++j) neg[j] = ~gen[j];
  for (int64_t kk = 0; kk < nmax; ++kk) {
    int i = indices[kk];
    const bk_t *d = S->psi_det_sorted + (int64_t)(i - 1) * 2 * nint;
    int nt = 0;
    for (int j = 0; j < 2 * nint; ++j) nt += __builtin_popcountll(neg[j] & d[j]);
    if (nt <= 4) {
      if (i <= S->n_sel) {
        preint[npre] = i;
        memcpy(preint_det + (int64_t)npre * 2 * nint, d, sizeof(bk_t) * 2 * nint);
        ++npre;
      } else if (nt <= 2) {
        prefull[nprefull++] = i;
      }
    }
  }
  free(indices);
  int *interesting = (int *)malloc(sizeof(int) * (npre + 1));
  int *fullint = (int *)malloc(sizeof(int) * (npre + nprefull + 1));
  bk_t *minilist = (bk_t *)malloc(sizeof(bk_t) * 2 * nint * (npre + 1));
  bk_t *fullmini = (bk_t *)malloc(sizeof(bk_t) * 2 * nint * (npre + nprefull + 1));
  int maskInd = -1;
  int monoAdo = 1, monoBdo = 1;
  bk_t pmask[2 * OQ_NINT_MAX], mask[2 * OQ_NINT_MAX];
  for (int s1 = 1; s1 <= 2; ++s1) {
    for (int i1 = N_holes[s1 - 1]; i1 >= 1; --i1) {
      int h1 = hole_list[s1 - 1][i1 - 1];
      apply_hole(gen, s1, h1, pmask, nint);
      for (int j = 0; j < 2 * nint; ++j) neg[j] = ~pmask[j];
      int nint_ = 0, nfull = 0;
      for (int ii = 0; ii < npre; ++ii) {
        const bk_t *d = preint_det + (int64_t)ii * 2 * nint;
        int nt = 0;
        for (int j = 0; j < 2 * nint; ++j) nt += __builtin_popcountll(neg[j] & d[j]);
        if (nt <= 4) {
          interesting[nint_] = preint[ii];
          memcpy(minilist + (int64_t)nint_ * 2 * nint, d, sizeof(bk_t) * 2 * nint);
          ++nint_;
          if (nt <= 2) {
            fullint[nfull] = preint[ii];
            memcpy(fullmini + (int64_t)nfull * 2 * nint, d, sizeof(bk_t) * 2 * nint);
            ++nfull;
          }
        }
      }
      for (int ii = 0; ii < nprefull; ++ii) {
        int i = prefull[ii];
        const bk_t *d = S->psi_det_sorted + (int64_t)(i - 1) * 2 * nint;
        int nt = 0;
        for (int j = 0; j < 2 * nint; ++j) nt += __builtin_popcountll(neg[j] & d[j]);
        if (nt <= 2) {
          fullint[nfull] = i;
          memcpy(fullmini + (int64_t)nfull * 2 * nint, d, sizeof(bk_t) * 2 * nint);
          ++nfull;
        }
      }
      for (int s2 = s1; s2 <= 2; ++s2) {
        int sp = (s1 != s2) ? 3 : s1;
        int ib = (s1 == s2) ? i1 + 1 : 1;
        monoAdo = 1;
        for (int i2 = N_holes[s2 - 1]; i2 >= ib; --i2) {
          int h2 = hole_list[s2 - 1][i2 - 1];
          apply_hole(pmask, s2, h2, mask, nint);
          memset(w->banned, 0, (size_t)M * M * 2);
          memset(w->bannedOrb, 1, (size_t)M * 2);
          for (int s3 = 1; s3 <= 2; ++s3)
            for (int i = 0; i < N_particles[s3 - 1]; ++i) BORB(w, particle_list[s3 - 1][i], s3) = 0;
          if (s1 != s2) {
            if (monoBdo) BORB(w, h1, s1) = 0;
            if (monoAdo) { BORB(w, h2, s2) = 0; monoAdo = 0; }
          }
          maskInd += 1;
#pragma omp atomic
          S->n_blocks_total++;
          int keep = (subset == 0 || (maskInd % S->fragment_count) == (subset - 1));
          if (keep && S->block_keep)
            keep = S->block_keep[S->block_keep_offsets[i_generator - 1] + maskInd] != 0;
          if (keep) {
            if (spot_isinwf(w, mask, fullmini, fullint, nfull, i_generator)) continue;
#pragma omp atomic
            S->n_blocks++;
            splash_pq(w, mask, sp, minilist, interesting, nint_, i_generator);
            double blk_e = 0.0;
            int64_t blk_c = 0;
            fill_buffer_double(w, gen, sp, h1, h2, fock, E0, pt2, buf, &blk_e, &blk_c);
            if (S->block_log) {
#pragma omp critical(oq_blocklog)
              {
                if (S->block_log_n < S->block_log_cap) {
                  double *r = S->block_log + 8 * S->block_log_n;
                  r[0] = i_generator; r[1] = s1; r[2] = h1; r[3] = s2; r[4] = h2;
                  r[5] = blk_e; r[6] = (double)blk_c; r[7] = maskInd;
                  S->block_log_n++;
                }
              }
            }
          }
        }
        if (s1 != s2) monoBdo = 0;
      }
    }
  }
  free(preint); free(prefull); free(preint_det); free(interesting); free(fullint);
  free(minilist); free(fullmini);
}

static selw *selw_create(const oq_ham *H, oq_sel_ctx *S) {
  selw *w = (selw *)calloc(1, sizeof(selw));
  w->H = H; w->S = S; w->M = H->M; w->NS = S->n_states; w->nint = H->nint;
  w->banned = (char *)malloc((size_t)w->M * w->M * 2);
  w->bannedOrb = (char *)malloc((size_t)w->M * 2);
  w->lbanned = (char *)malloc((size_t)w->M * 2);
  w->mat = (double *)malloc(sizeof(double) * w->NS * w->M * w->M);
  w->tmp_row = (double *)malloc(sizeof(double) * w->NS * w->M);
  w->tmp_row2 = (double *)malloc(sizeof(double) * w->NS * w->M);
  w->phasemask = (int *)malloc(sizeof(int) * 2 * 64 * w->nint);
  return w;
}
static void selw_destroy(selw *w) {
  free(w->banned); free(w->bannedOrb); free(w->lbanned); free(w->mat);
  free(w->tmp_row); free(w->tmp_row2); free(w->phasemask); free(w);
}

void oq_select_connected(const oq_ham *H, oq_sel_ctx *S, int i_generator, const double *E0,
                         double *pt2, oq_sel_buffer *buf, int subset) {
  selw *w = selw_create(H, S);
  double *fock = (double *)malloc(sizeof(double) * 2 * (H->M + 1));
  const bk_t *gen = S->psi_det_sorted + (int64_t)(i_generator - 1) * 2 * H->nint;
  oq_build_fock_tmp(H, gen, fock);
  bk_t hm[2 * OQ_NINT_MAX], pm[2 * OQ_NINT_MAX];
  for (int k = 0; k < 2 * H->nint; ++k) {
    hm[k] = S->gen_hole_mask[k] & gen[k];
    pm[k] = S->gen_part_mask[k] & ~gen[k];
  }
  select_singles_and_doubles(w, i_generator, hm, pm, fock, E0, pt2, buf, subset);
#pragma omp atomic
  S->n_candidates += w->n_cand;
#pragma omp atomic
  S->n_gathers += w->n_gathers;
  free(fock);
  selw_destroy(w);
}

void oq_select_all(const oq_ham *H, oq_sel_ctx *S, const double *E0, double *pt2, oq_sel_buffer *out,
                   int nthreads) {
  oq_select_tasks(H, S, E0, pt2, out, nthreads, 0, NULL, NULL, NULL);
}

/* The task farm of ZMQ_selection (zmq_selection.irp.f:20-40): one task is a
   (subset, i_generator) pair; subset 0 = every block of the generator,
   subset k = the blocks with maskInd mod fragment_count == k-1
   (selection.irp.f:561-562).  n_tasks == 0 runs every generator whole.
   task_pt2 (NS x n_tasks, may be NULL) receives each task's own EN-PT2, the
   pt2 a slave sends back for one task (run_pt2_slave.irp.f:66-80). */
void oq_select_tasks(const oq_ham *H, oq_sel_ctx *S, const double *E0, double *pt2, oq_sel_buffer *out,
                     int nthreads, int n_tasks, const int *task_gen, const int *task_subset, double *task_pt2) {
  int NS = S->n_states, nint = H->nint;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  for (int st = 0; st < NS; ++st) pt2[st] = 0.0;
  int N = out->N;
  oq_sel_buffer **bufs = (oq_sel_buffer **)calloc(nthreads, sizeof(oq_sel_buffer *));
  double *pt2t = (double *)calloc((size_t)nthreads * NS, sizeof(double));
#pragma omp parallel num_threads(nthreads)
  {
    int t = omp_get_thread_num();
    bufs[t] = oq_sel_buffer_create(N, 2 * N, nint);
    if (n_tasks > 0) {
#pragma omp for schedule(dynamic, 1)
      for (int k = 0; k < n_tasks; ++k) {
        double tp[OQ_MAXNS];
        for (int st = 0; st < NS; ++st) tp[st] = 0.0;
        oq_select_connected(H, S, task_gen[k], E0, tp, bufs[t], task_subset[k]);
        for (int st = 0; st < NS; ++st) {
          pt2t[(int64_t)t * NS + st] += tp[st];
          if (task_pt2) task_pt2[(int64_t)k * NS + st] = tp[st];
        }
      }
    } else {
#pragma omp for schedule(dynamic, 1)
      for (int g = 1; g <= S->n_gen; ++g) oq_select_connected(H, S, g, E0, pt2t + (int64_t)t * NS, bufs[t], 0);
    }
    oq_sel_buffer_sort(bufs[t]);
  }
  /* collector: gather every slave buffer and keep the N smallest values */
  int total = 0;
  for (int t = 0; t < nthreads; ++t) total += bufs[t]->cur;
  oq_sel_buffer *all = oq_sel_buffer_create(N, total > 0 ? total : 1, nint);
  for (int t = 0; t < nthreads; ++t) {
    for (int i = 0; i < bufs[t]->cur; ++i) {
      memcpy(all->det + (int64_t)all->cur * 2 * nint, bufs[t]->det + (int64_t)i * 2 * nint, sizeof(bk_t) * 2 * nint);
      all->val[all->cur++] = bufs[t]->val[i];
    }
    for (int st = 0; st < NS; ++st) pt2[st] += pt2t[(int64_t)t * NS + st];
    oq_sel_buffer_destroy(bufs[t]);
  }
  oq_sel_buffer_sort(all);
  out->cur = all->cur < out->cap ? all->cur : out->cap;
  memcpy(out->det, all->det, sizeof(bk_t) * 2 * nint * out->cur);
  memcpy(out->val, all->val, sizeof(double) * out->cur);
  oq_sel_buffer_destroy(all);
  free(bufs); free(pt2t);
}
