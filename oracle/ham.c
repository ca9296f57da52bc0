/*
 * ham.c -- Hamiltonian providers of the oracle (test infrastructure only).
 *
 * Restates:
 *   bielec_integrals_index            src/Integrals_Bielec/map_integrals.irp.f:19-34
 *   get_mo_bielec_integral            src/Integrals_Bielec/map_integrals.irp.f:383-428
 *   mo_bielec_integral_jj{,_anti}     src/Integrals_Bielec/mo_bi_integrals.irp.f:1344-1369
 *   big_array_{coulomb,exchange}      src/Integrals_Bielec/integrals_3_index.irp.f:1-24
 *   HF_bitmask / ref_bitmask          src/Bitmask/bitmasks.irp.f:71-95
 *   ref_bitmask_energy                src/Determinants/ref_bitmask.irp.f:1-57
 *   ref_closed_shell_bitmask,
 *   fock_operator_closed_shell_ref_bitmask
 *                                     src/Determinants/mono_excitations.irp.f:1-95
 */
#include <stdlib.h>
#include <string.h>
#include "qp_oracle.h"

static inline int64_t tri1(int64_t a, int64_t b) {
  int64_t p = a < b ? a : b, r = a < b ? b : a;
  return p + ((r * r - r) >> 1);
}

int64_t oq_integral_key(int i, int j, int k, int l) {
  /* map_integrals.irp.f:19-34: p = tri(i,k), q = tri(j,l), key = tri(p,q) */
  return tri1(tri1(i, k), tri1(j, l));
}

double oq_get_integral(const oq_ham *H, int i, int j, int k, int l) {
  return H->eri[oq_integral_key(i, j, k, l) - 1];
}

static void set_bit(bk_t *s, int orb) { s[(orb - 1) >> 6] |= ((bk_t)1) << ((orb - 1) & 63); }
static void clr_bit(bk_t *s, int orb) { s[(orb - 1) >> 6] &= ~(((bk_t)1) << ((orb - 1) & 63)); }

static int to_list(const bk_t *s, int nint, int *list) {
  int n = 0;
  for (int w = 0; w < nint; ++w) {
    bk_t x = s[w];
    while (x) {
      int tz = __builtin_ctzll(x);
      list[n++] = tz + 1 + 64 * w;
      x &= x - 1;
    }
  }
  return n;
}

oq_ham *oq_ham_create(int M, int na, int nb, const double *h, const double *eri, int64_t n8) {
  oq_ham *H = (oq_ham *)calloc(1, sizeof(oq_ham));
  H->M = M; H->na = na; H->nb = nb;
  H->nint = (M + 63) / 64;
  H->npair = (int64_t)M * (M + 1) / 2;
  H->n8 = H->npair * (H->npair + 1) / 2;
  if (n8 != H->n8 || H->nint > OQ_NINT_MAX) { free(H); return NULL; }
  H->h = (double *)malloc(sizeof(double) * M * M);
  memcpy(H->h, h, sizeof(double) * M * M);
  H->eri = (double *)malloc(sizeof(double) * n8);
  memcpy(H->eri, eri, sizeof(double) * n8);
  int64_t M2 = (int64_t)M * M, M3 = M2 * M;
  H->jj = (double *)malloc(sizeof(double) * M2);
  H->jj_exch = (double *)malloc(sizeof(double) * M2);
  H->jj_anti = (double *)malloc(sizeof(double) * M2);
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= M; ++i) {
      int64_t x = (i - 1) + (int64_t)(j - 1) * M;
      H->jj[x] = oq_get_integral(H, i, j, i, j);
      H->jj_exch[x] = oq_get_integral(H, i, j, j, i);
      H->jj_anti[x] = H->jj[x] - H->jj_exch[x];
    }
  H->big_coul = (double *)malloc(sizeof(double) * M3);
  H->big_exch = (double *)malloc(sizeof(double) * M3);
  for (int k = 1; k <= M; ++k)
    for (int i = 1; i <= M; ++i)
      for (int j = 1; j <= M; ++j) {
        int64_t x = (j - 1) + (int64_t)M * ((i - 1) + (int64_t)M * (k - 1));
        H->big_coul[x] = oq_get_integral(H, i, j, k, j);
        H->big_exch[x] = oq_get_integral(H, i, j, j, k);
      }
  /* HF bitmask: orbitals 1..na alpha, 1..nb beta */
  memset(H->ref, 0, sizeof(H->ref));
  for (int i = 1; i <= na; ++i) set_bit(H->ref, i);
  for (int i = 1; i <= nb; ++i) set_bit(H->ref + H->nint, i);
  /* ref_bitmask_energy, ref_bitmask.irp.f:1-57 */
  {
    int occa[256], occb[256];
    to_list(H->ref, H->nint, occa);
    to_list(H->ref + H->nint, H->nint, occb);
    double e = 0.0;
    for (int i = 0; i < nb; ++i)
      e += h[(occa[i] - 1) * (M + 1)] + h[(occb[i] - 1) * (M + 1)];
    for (int i = nb; i < na; ++i) e += h[(occa[i] - 1) * (M + 1)];
    for (int j = 0; j < na; ++j)
      for (int i = j + 1; i < na; ++i) e += H->jj_anti[(occa[i] - 1) + (int64_t)(occa[j] - 1) * M];
    for (int j = 0; j < nb; ++j) {
      for (int i = j + 1; i < nb; ++i) e += H->jj_anti[(occb[i] - 1) + (int64_t)(occb[j] - 1) * M];
      for (int i = 0; i < na; ++i) e += H->jj[(occa[i] - 1) + (int64_t)(occb[j] - 1) * M];
    }
    H->ref_energy = e;
  }
  /* ref_closed_shell_bitmask: clear alpha orbitals occ(nb+1..na) */
  memcpy(H->ref_cs, H->ref, sizeof(H->ref));
  {
    int occ[256];
    to_list(H->ref, H->nint, occ);
    for (int i0 = nb; i0 < na; ++i0) clr_bit(H->ref_cs, occ[i0]);
  }
  /* fock_operator_closed_shell_ref_bitmask, mono_excitations.irp.f:22-95.
   * Every (i,j) pair is covered by the docc/virt x docc/virt loops. */
  H->fock_cs = (double *)malloc(sizeof(double) * M2);
  {
    int occ[256], nocc = to_list(H->ref_cs, H->nint, occ);
    for (int i = 1; i <= M; ++i)
      for (int j = 1; j <= M; ++j) {
        double accu = 0.0;
        for (int k0 = 0; k0 < nocc; ++k0) {
          int k = occ[k0];
          accu += 2.0 * oq_get_integral(H, i, k, j, k) - oq_get_integral(H, i, k, k, j);
        }
        H->fock_cs[(i - 1) + (int64_t)(j - 1) * M] = accu + h[(i - 1) + (int64_t)(j - 1) * M];
      }
  }
  return H;
}

void oq_ham_destroy(oq_ham *H) {
  if (!H) return;
  free(H->h); free(H->eri); free(H->jj); free(H->jj_exch); free(H->jj_anti);
  free(H->big_coul); free(H->big_exch); free(H->fock_cs);
  free(H);
}

double oq_ref_energy(const oq_ham *H) { return H->ref_energy; }
void oq_get_fock_cs(const oq_ham *H, double *out) {
  memcpy(out, H->fock_cs, sizeof(double) * H->M * H->M);
}
