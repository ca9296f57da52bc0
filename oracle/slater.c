/*
 * slater.c -- Slater-Condon rules of the oracle (test infrastructure only).
 *
 * Restates src/Determinants/slater_rules.irp.f:
 *   get_excitation_degree            :1-56
 *   get_excitation                   :59-103
 *   decode_exc                       :105-168
 *   get_double_excitation            :171-313
 *   get_mono_excitation              :315-401
 *   i_H_j                            :478-571
 *   diag_H_mat_elem_fock             :1800-1867
 *   diag_H_mat_elem, a/ac_operator   :1869-2022
 *   get_excitation_degree_spin       :2305-2357
 *   get_{double,mono}_excitation_spin :2783-2987
 *   i_H_j_{mono_spin,double_spin,double_alpha_beta} :2989-3069
 * src/Determinants/mono_excitations.irp.f:97-155 (get_mono_excitation_from_fock)
 * src/Determinants/s2.irp.f:1-63 (diag_S_mat_elem, get_s2)
 * src/Determinants/Fock_diag.irp.f:1-93 (build_fock_tmp)
 *
 * exc is laid out as Fortran exc(0:2,2,2): exc[n + 3*hp + 6*spin],
 * n = 0 (count), 1, 2; hp = 0 hole, 1 particle; spin = 0 alpha, 1 beta.
 * Single-spin variants use exc(0:2,2): exc[n + 3*hp].
 */
#include <string.h>
#include "qp_oracle.h"

#define EXC(e, n, hp, s) (e)[(n) + 3 * (hp) + 6 * (s)]
#define EXCS(e, n, hp) (e)[(n) + 3 * (hp)]

static inline int pc(bk_t x) { return __builtin_popcountll(x); }

int oq_excitation_degree(const bk_t *k1, const bk_t *k2, int nint) {
  int d = 0;
  for (int l = 0; l < 2 * nint; ++l) d += pc(k1[l] ^ k2[l]);
  return d >> 1;
}

int oq_excitation_degree_spin(const bk_t *k1, const bk_t *k2, int nint) {
  int d = 0;
  for (int l = 0; l < nint; ++l) d += pc(k1[l] ^ k2[l]);
  return d >> 1;
}

/* number of occupied orbitals of string s strictly between orbitals a and b
 * (1-based), as the masked popcounts of slater_rules.irp.f:236-262 */
static int nperm_between(const bk_t *s, int a, int b) {
  int high = (a > b ? a : b) - 1;
  int low = a < b ? a : b;
  int k = high >> 6, j = low >> 6;
  int m = high & 63, n = low & 63;
  int np = 0;
  if (j == k) {
    bk_t mask = ((((bk_t)1) << m) - 1) & (~(((bk_t)1) << n) + 1);
    np += pc(s[j] & mask);
  } else {
    np += pc(s[j] & (~(((bk_t)1) << n) + 1));
    np += pc(s[k] & ((((bk_t)1) << m) - 1));
    for (int i = j + 1; i < k; ++i) np += pc(s[i]);
  }
  return np;
}

/* holes/particles of one spin: particles = bits of d2 not in d1 */
static void spin_exc_lists(const bk_t *d1, const bk_t *d2, int nint, int *exc3 /* exc(0:2,2) */) {
  int nh = 0, np = 0;
  for (int l = 0; l < nint; ++l) {
    if (d1[l] == d2[l]) continue;
    bk_t t = d1[l] ^ d2[l];
    bk_t part = t & d2[l], hole = t & d1[l];
    while (part) {
      int tz = __builtin_ctzll(part);
      ++np;
      if (np <= 2) EXCS(exc3, np, 1) = tz + 1 + 64 * l;
      part &= part - 1;
    }
    while (hole) {
      int tz = __builtin_ctzll(hole);
      ++nh;
      if (nh <= 2) EXCS(exc3, nh, 0) = tz + 1 + 64 * l;
      hole &= hole - 1;
    }
  }
  EXCS(exc3, 0, 0) = nh;
  EXCS(exc3, 0, 1) = np;
}

/* permutation count of the excitation in one spin (1 or 2 pairs) */
static int spin_nperm(const bk_t *d1, const int *exc3) {
  int nperm = 0;
  int nh = EXCS(exc3, 0, 0);
  for (int l = 1; l <= nh; ++l) nperm += nperm_between(d1, EXCS(exc3, l, 0), EXCS(exc3, l, 1));
  if (nh == 2) {
    int h1 = EXCS(exc3, 1, 0), p1 = EXCS(exc3, 1, 1), h2 = EXCS(exc3, 2, 0), p2 = EXCS(exc3, 2, 1);
    int a = h1 < p1 ? h1 : p1, b = h1 < p1 ? p1 : h1;
    int c = h2 < p2 ? h2 : p2, d = h2 < p2 ? p2 : h2;
    if (a < c && c < b && b < d) nperm += 1;
  }
  return nperm;
}

void oq_get_double_excitation(const bk_t *d1, const bk_t *d2, int *exc, double *phase, int nint) {
  int nperm = 0;
  memset(exc, 0, sizeof(int) * 12);
  for (int s = 0; s < 2; ++s) {
    int e3[6] = {0, 0, 0, 0, 0, 0};
    spin_exc_lists(d1 + s * nint, d2 + s * nint, nint, e3);
    for (int n = 0; n < 3; ++n) { EXC(exc, n, 0, s) = EXCS(e3, n, 0); EXC(exc, n, 1, s) = EXCS(e3, n, 1); }
    if (EXCS(e3, 0, 0) == 0) continue;
    nperm += spin_nperm(d1 + s * nint, e3);
    if (EXCS(e3, 0, 0) == 2) break;
  }
  *phase = (nperm & 1) ? -1.0 : 1.0;
}

void oq_get_mono_excitation(const bk_t *d1, const bk_t *d2, int *exc, double *phase, int nint) {
  memset(exc, 0, sizeof(int) * 12);
  *phase = 1.0;
  for (int s = 0; s < 2; ++s) {
    int e3[6] = {0, 0, 0, 0, 0, 0};
    spin_exc_lists(d1 + s * nint, d2 + s * nint, nint, e3);
    if (EXCS(e3, 0, 0) != 1 || EXCS(e3, 0, 1) != 1) continue;
    for (int n = 0; n < 3; ++n) { EXC(exc, n, 0, s) = EXCS(e3, n, 0); EXC(exc, n, 1, s) = EXCS(e3, n, 1); }
    int nperm = spin_nperm(d1 + s * nint, e3);
    *phase = (nperm & 1) ? -1.0 : 1.0;
    return;
  }
}

void oq_get_excitation(const bk_t *d1, const bk_t *d2, int *exc, int *degree, double *phase, int nint) {
  int d = oq_excitation_degree(d1, d2, nint);
  *degree = d;
  if (d >= 3) { *degree = -1; return; }
  if (d == 2) oq_get_double_excitation(d1, d2, exc, phase, nint);
  else if (d == 1) oq_get_mono_excitation(d1, d2, exc, phase, nint);
}

void oq_get_mono_excitation_spin(const bk_t *d1, const bk_t *d2, int *exc, double *phase, int nint) {
  memset(exc, 0, sizeof(int) * 6);
  spin_exc_lists(d1, d2, nint, exc);
  *phase = (spin_nperm(d1, exc) & 1) ? -1.0 : 1.0;
}

void oq_get_double_excitation_spin(const bk_t *d1, const bk_t *d2, int *exc, double *phase, int nint) {
  memset(exc, 0, sizeof(int) * 6);
  spin_exc_lists(d1, d2, nint, exc);
  *phase = (spin_nperm(d1, exc) & 1) ? -1.0 : 1.0;
}

void oq_decode_exc(const int *exc, int degree, int *h1, int *p1, int *h2, int *p2, int *s1, int *s2) {
  *h1 = *p1 = *h2 = *p2 = *s1 = *s2 = 0;
  if (degree == 2) {
    if (EXC(exc, 0, 0, 0) == 2) {
      *h1 = EXC(exc, 1, 0, 0); *h2 = EXC(exc, 2, 0, 0);
      *p1 = EXC(exc, 1, 1, 0); *p2 = EXC(exc, 2, 1, 0); *s1 = 1; *s2 = 1;
    } else if (EXC(exc, 0, 0, 1) == 2) {
      *h1 = EXC(exc, 1, 0, 1); *h2 = EXC(exc, 2, 0, 1);
      *p1 = EXC(exc, 1, 1, 1); *p2 = EXC(exc, 2, 1, 1); *s1 = 2; *s2 = 2;
    } else {
      *h1 = EXC(exc, 1, 0, 0); *h2 = EXC(exc, 1, 0, 1);
      *p1 = EXC(exc, 1, 1, 0); *p2 = EXC(exc, 1, 1, 1); *s1 = 1; *s2 = 2;
    }
  } else if (degree == 1) {
    if (EXC(exc, 0, 0, 0) == 1) {
      *h1 = EXC(exc, 1, 0, 0); *p1 = EXC(exc, 1, 1, 0); *s1 = 1;
    } else {
      *h1 = EXC(exc, 1, 0, 1); *p1 = EXC(exc, 1, 1, 1); *s1 = 2;
    }
  }
}

static int to_list(const bk_t *s, int nint, int *list) {
  int n = 0;
  for (int w = 0; w < nint; ++w) {
    bk_t x = s[w];
    while (x) { list[n++] = __builtin_ctzll(x) + 1 + 64 * w; x &= x - 1; }
  }
  return n;
}

#define BIGC(H, a, b, c) (H)->big_coul[((a) - 1) + (int64_t)(H)->M * (((b) - 1) + (int64_t)(H)->M * ((c) - 1))]
#define BIGX(H, a, b, c) (H)->big_exch[((a) - 1) + (int64_t)(H)->M * (((b) - 1) + (int64_t)(H)->M * ((c) - 1))]
#define MAT(H, A, i, j) (H)->A[((i) - 1) + (int64_t)(H)->M * ((j) - 1)]

double oq_mono_excitation_from_fock(const oq_ham *H, const bk_t *d1, int h, int p, int spin, double phase) {
  int nint = H->nint;
  bk_t hole[2 * OQ_NINT_MAX], partcl[2 * OQ_NINT_MAX];
  for (int i = 0; i < 2 * nint; ++i) {
    bk_t diff = d1[i] ^ H->ref_cs[i];
    hole[i] = diff & H->ref_cs[i];
    partcl[i] = diff & d1[i];
  }
  int oh[2][256], op[2][256], nh[2], np[2];
  for (int s = 0; s < 2; ++s) {
    nh[s] = to_list(hole + s * nint, nint, oh[s]);
    np[s] = to_list(partcl + s * nint, nint, op[s]);
  }
  double hij = MAT(H, fock_cs, h, p);
  int sp = spin - 1;
  for (int s = 0; s < 2; ++s)
    for (int i0 = 0; i0 < nh[s]; ++i0) hij -= BIGC(H, oh[s][i0], h, p);
  for (int i0 = 0; i0 < nh[sp]; ++i0) hij += BIGX(H, oh[sp][i0], h, p);
  for (int s = 0; s < 2; ++s)
    for (int i0 = 0; i0 < np[s]; ++i0) hij += BIGC(H, op[s][i0], h, p);
  for (int i0 = 0; i0 < np[sp]; ++i0) hij -= BIGX(H, op[sp][i0], h, p);
  return hij * phase;
}

double oq_i_H_j_mono_spin(const oq_ham *H, const bk_t *ki, const bk_t *kj, int spin) {
  int exc[6];
  double phase;
  int off = (spin - 1) * H->nint;
  oq_get_mono_excitation_spin(ki + off, kj + off, exc, &phase, H->nint);
  return oq_mono_excitation_from_fock(H, ki, EXCS(exc, 1, 0), EXCS(exc, 1, 1), spin, phase);
}

double oq_i_H_j_double_spin(const oq_ham *H, const bk_t *ki, const bk_t *kj) {
  int exc[6];
  double phase;
  oq_get_double_excitation_spin(ki, kj, exc, &phase, H->nint);
  int h1 = EXCS(exc, 1, 0), h2 = EXCS(exc, 2, 0), p1 = EXCS(exc, 1, 1), p2 = EXCS(exc, 2, 1);
  return phase * (oq_get_integral(H, h1, h2, p1, p2) - oq_get_integral(H, h1, h2, p2, p1));
}

double oq_i_H_j_double_alpha_beta(const oq_ham *H, const bk_t *ki, const bk_t *kj) {
  int ea[6], eb[6];
  double pa, pb;
  int nint = H->nint;
  oq_get_mono_excitation_spin(ki, kj, ea, &pa, nint);
  oq_get_mono_excitation_spin(ki + nint, kj + nint, eb, &pb, nint);
  double phase = pa * pb;
  int ha = EXCS(ea, 1, 0), pa_ = EXCS(ea, 1, 1), hb = EXCS(eb, 1, 0), pb_ = EXCS(eb, 1, 1);
  if (ha == pb_) return phase * BIGX(H, ha, hb, pa_);
  if (pa_ == hb) return phase * BIGX(H, pa_, ha, pb_);
  return phase * oq_get_integral(H, ha, hb, pa_, pb_);
}

/* a_operator / ac_operator, slater_rules.irp.f:1931-2022 */
static void a_operator(const oq_ham *H, int iorb, int ispin, bk_t *key, double *hjj, int *na, int *nb) {
  int nint = H->nint, other = 1 - ispin;
  key[ispin * nint + ((iorb - 1) >> 6)] &= ~(((bk_t)1) << ((iorb - 1) & 63));
  int occ[2][256];
  to_list(key, nint, occ[0]);
  to_list(key + nint, nint, occ[1]);
  *na -= 1;
  *hjj -= MAT(H, h, iorb, iorb);
  for (int i = 0; i < *na; ++i) *hjj -= MAT(H, jj_anti, occ[ispin][i], iorb);
  for (int i = 0; i < *nb; ++i) *hjj -= MAT(H, jj, occ[other][i], iorb);
}

static void ac_operator(const oq_ham *H, int iorb, int ispin, bk_t *key, double *hjj, int *na, int *nb) {
  int nint = H->nint, other = 1 - ispin;
  int occ[2][256];
  to_list(key, nint, occ[0]);
  to_list(key + nint, nint, occ[1]);
  key[ispin * nint + ((iorb - 1) >> 6)] |= ((bk_t)1) << ((iorb - 1) & 63);
  *hjj += MAT(H, h, iorb, iorb);
  for (int i = 0; i < *na; ++i) *hjj += MAT(H, jj_anti, occ[ispin][i], iorb);
  for (int i = 0; i < *nb; ++i) *hjj += MAT(H, jj, occ[other][i], iorb);
  *na += 1;
}

double oq_diag_H_mat_elem(const oq_ham *H, const bk_t *det) {
  int nint = H->nint;
  bk_t hole[2 * OQ_NINT_MAX], particle[2 * OQ_NINT_MAX];
  int nexc[2] = {0, 0};
  for (int s = 0; s < 2; ++s)
    for (int i = 0; i < nint; ++i) {
      bk_t x = det[s * nint + i] ^ H->ref[s * nint + i];
      particle[s * nint + i] = x & det[s * nint + i];
      hole[s * nint + i] = x & H->ref[s * nint + i];
      nexc[s] += __builtin_popcountll(hole[s * nint + i]);
    }
  double e = H->ref_energy;
  if (nexc[0] + nexc[1] == 0) return e;
  int op[2][256], oh[2][256];
  for (int s = 0; s < 2; ++s) {
    to_list(particle + s * nint, nint, op[s]);
    to_list(hole + s * nint, nint, oh[s]);
  }
  bk_t tmp[2 * OQ_NINT_MAX];
  memcpy(tmp, H->ref, sizeof(bk_t) * 2 * nint);
  int elec[2] = {H->na, H->nb};
  for (int s = 0; s < 2; ++s) {
    int na = elec[s], nb = elec[1 - s];
    for (int i = 0; i < nexc[s]; ++i) {
      ac_operator(H, op[s][i], s, tmp, &e, &na, &nb);
      a_operator(H, oh[s][i], s, tmp, &e, &na, &nb);
    }
  }
  return e;
}

double oq_diag_S_mat_elem(const bk_t *det, int nint) {
  int nup = 0;
  for (int i = 0; i < nint; ++i) nup += pc((det[i] ^ det[nint + i]) & det[i]);
  return (double)nup;
}

double oq_get_s2(const bk_t *ki, const bk_t *kj, int nint) {
  int degree = oq_excitation_degree(ki, kj, nint);
  if (degree == 2) {
    int exc[12];
    double ph;
    oq_get_double_excitation(kj, ki, exc, &ph, nint);
    if (EXC(exc, 0, 0, 0) == 1) {
      if (EXC(exc, 1, 0, 0) == EXC(exc, 1, 1, 1) && EXC(exc, 1, 0, 1) == EXC(exc, 1, 1, 0)) return -ph;
    }
    return 0.0;
  }
  if (degree == 0) return oq_diag_S_mat_elem(ki, nint);
  return 0.0;
}

double oq_i_H_j(const oq_ham *H, const bk_t *ki, const bk_t *kj) {
  int nint = H->nint, exc[12], degree;
  double phase;
  degree = oq_excitation_degree(ki, kj, nint);
  if (degree == 2) {
    oq_get_double_excitation(ki, kj, exc, &phase, nint);
    if (EXC(exc, 0, 0, 0) == 1) {
      int ha = EXC(exc, 1, 0, 0), hb = EXC(exc, 1, 0, 1), pa = EXC(exc, 1, 1, 0), pb = EXC(exc, 1, 1, 1);
      if (ha == pb) return phase * BIGX(H, ha, hb, pa);
      if (pa == hb) return phase * BIGX(H, pa, ha, pb);
      return phase * oq_get_integral(H, ha, hb, pa, pb);
    } else if (EXC(exc, 0, 0, 0) == 2) {
      int h1 = EXC(exc, 1, 0, 0), h2 = EXC(exc, 2, 0, 0), p1 = EXC(exc, 1, 1, 0), p2 = EXC(exc, 2, 1, 0);
      return phase * (oq_get_integral(H, h1, h2, p1, p2) - oq_get_integral(H, h1, h2, p2, p1));
    } else if (EXC(exc, 0, 0, 1) == 2) {
      int h1 = EXC(exc, 1, 0, 1), h2 = EXC(exc, 2, 0, 1), p1 = EXC(exc, 1, 1, 1), p2 = EXC(exc, 2, 1, 1);
      return phase * (oq_get_integral(H, h1, h2, p1, p2) - oq_get_integral(H, h1, h2, p2, p1));
    }
    return 0.0;
  } else if (degree == 1) {
    oq_get_mono_excitation(ki, kj, exc, &phase, nint);
    int m, p, spin;
    if (EXC(exc, 0, 0, 0) == 1) { m = EXC(exc, 1, 0, 0); p = EXC(exc, 1, 1, 0); spin = 1; }
    else { m = EXC(exc, 1, 0, 1); p = EXC(exc, 1, 1, 1); spin = 2; }
    /* the reference passes (p, m) as (h, p): the Fock-based element is
     * symmetric in its two orbital arguments */
    return oq_mono_excitation_from_fock(H, ki, p, m, spin, phase);
  } else if (degree == 0) {
    return oq_diag_H_mat_elem(H, ki);
  }
  return 0.0;
}

void oq_build_fock_tmp(const oq_ham *H, const bk_t *det, double *f) {
  int M = H->M, nint = H->nint;
  int occ[2][256];
  int ne0 = to_list(det, nint, occ[0]);
  int ne1 = to_list(det + nint, nint, occ[1]);
  (void)ne0; (void)ne1;
  for (int i = 0; i < 2 * (M + 1); ++i) f[i] = 0.0;
  double E0 = 0.0;
#define F(s, i) f[(s) + 2 * ((i) - 1)]
  for (int ii = 0; ii < H->na; ++ii) {
    int i = occ[0][ii];
    F(0, i) += MAT(H, h, i, i);
    E0 += MAT(H, h, i, i);
    for (int jj = 0; jj < H->na; ++jj) {
      int j = occ[0][jj];
      if (i == j) continue;
      F(0, i) += MAT(H, jj_anti, i, j);
      E0 += 0.5 * MAT(H, jj_anti, i, j);
    }
    for (int jj = 0; jj < H->nb; ++jj) {
      int j = occ[1][jj];
      F(0, i) += MAT(H, jj, i, j);
      E0 += MAT(H, jj, i, j);
    }
  }
  for (int ii = 0; ii < H->nb; ++ii) {
    int i = occ[1][ii];
    F(1, i) += MAT(H, h, i, i);
    E0 += MAT(H, h, i, i);
    for (int jj = 0; jj < H->nb; ++jj) {
      int j = occ[1][jj];
      if (i == j) continue;
      F(1, i) += MAT(H, jj_anti, i, j);
      E0 += 0.5 * MAT(H, jj_anti, i, j);
    }
    for (int jj = 0; jj < H->na; ++jj) {
      int j = occ[0][jj];
      F(1, i) += MAT(H, jj, i, j);
    }
  }
  for (int i = 1; i <= M; ++i) {
    if (F(0, i) != 0.0) continue;
    F(0, i) += MAT(H, h, i, i);
    for (int jj = 0; jj < H->na; ++jj) F(0, i) += MAT(H, jj_anti, i, occ[0][jj]);
    for (int jj = 0; jj < H->nb; ++jj) F(0, i) += MAT(H, jj, i, occ[1][jj]);
  }
  for (int i = 1; i <= M; ++i) {
    if (F(1, i) != 0.0) continue;
    F(1, i) += MAT(H, h, i, i);
    for (int jj = 0; jj < H->nb; ++jj) F(1, i) += MAT(H, jj_anti, i, occ[1][jj]);
    for (int jj = 0; jj < H->na; ++jj) F(1, i) += MAT(H, jj, i, occ[0][jj]);
  }
  F(0, M + 1) = E0;
  F(1, M + 1) = E0;
#undef F
}

double oq_diag_H_mat_elem_fock(const oq_ham *H, const bk_t *ref, const bk_t *pert, const double *f) {
  int nint = H->nint, exc[12], h1, p1, h2, p2, s1, s2;
  double phase;
  int degree = oq_excitation_degree(ref, pert, nint);
  int M = H->M;
#define F(s, i) f[((s) - 1) + 2 * ((i) - 1)]
  double E0 = F(1, M + 1);
  if (degree == 2) {
    oq_get_double_excitation(ref, pert, exc, &phase, nint);
    oq_decode_exc(exc, 2, &h1, &p1, &h2, &p2, &s1, &s2);
    if (s1 == 1 && s2 == 1) {
      return E0 - F(1, h1) + (F(1, p1) - MAT(H, jj_anti, h1, p1))
             - (F(1, h2) - MAT(H, jj_anti, h1, h2) + MAT(H, jj_anti, p1, h2))
             + (F(1, p2) - MAT(H, jj_anti, h1, p2) + MAT(H, jj_anti, p1, p2) - MAT(H, jj_anti, h2, p2));
    } else if (s1 == 2 && s2 == 2) {
      return E0 - F(2, h1) + (F(2, p1) - MAT(H, jj_anti, h1, p1))
             - (F(2, h2) - MAT(H, jj_anti, h1, h2) + MAT(H, jj_anti, p1, h2))
             + (F(2, p2) - MAT(H, jj_anti, h1, p2) + MAT(H, jj_anti, p1, p2) - MAT(H, jj_anti, h2, p2));
    } else {
      return E0 - F(1, h1) + (F(1, p1) - MAT(H, jj_anti, h1, p1))
             - (F(2, h2) - MAT(H, jj, h1, h2) + MAT(H, jj, p1, h2))
             + (F(2, p2) - MAT(H, jj, h1, p2) + MAT(H, jj, p1, p2) - MAT(H, jj_anti, h2, p2));
    }
  } else if (degree == 1) {
    oq_get_mono_excitation(ref, pert, exc, &phase, nint);
    oq_decode_exc(exc, 1, &h1, &p1, &h2, &p2, &s1, &s2);
    if (s1 == 1) return E0 - F(1, h1) + (F(1, p1) - MAT(H, jj_anti, h1, p1));
    return E0 - F(2, h1) + (F(2, p1) - MAT(H, jj_anti, h1, p1));
  } else if (degree == 0) {
    return E0;
  }
  return 0.0; /* the reference stops here: 'Bug in diag_H_mat_elem_fock' */
#undef F
}
