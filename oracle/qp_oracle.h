/*
 * qp_oracle.h -- CPU restatement of Quantum Package v1's determinant-driven
 * CI hot path (Davidson sigma H.Psi / S^2.Psi and CIPSI selection with
 * Epstein-Nesbet PT2).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle: only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark load
 * it.  The product library (libqpx) never links or calls it.
 *
 * Conventions follow the reference:
 *   - bit_kind = 64-bit words; orbital p (1-based) is bit (p-1)%64 of word
 *     (p-1)/64 (src/Bitmask/bitmasks_module.f90:1-11,
 *     src/Bitmask/bitmasks_routines.irp.f:1-63);
 *   - a determinant is det(N_int,2): alpha words then beta words, i.e.
 *     det[w + s*nint];
 *   - orbital indices in exc/h/p are 1-based, as in the reference.
 *   - <ij|kl> = (ik|jl) (physicist), stored 8-fold packed with the key of
 *     src/Integrals_Bielec/map_integrals.irp.f:19-34.
 */
#ifndef QP_ORACLE_H
#define QP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OQ_NINT_MAX 4
typedef uint64_t bk_t;

typedef struct {
  int M, na, nb, nint;
  int64_t npair, n8;
  double *h;          /* M*M column-major: mo_mono_elec_integral */
  double *eri;        /* n8 packed, index = key-1 */
  double *jj, *jj_exch, *jj_anti;    /* M*M, mo_bi_integrals.irp.f:1344-1369 */
  double *big_coul, *big_exch;       /* M^3, integrals_3_index.irp.f:1-24 */
  double *fock_cs;    /* M*M fock_operator_closed_shell_ref_bitmask */
  bk_t ref[2 * OQ_NINT_MAX];         /* ref_bitmask = HF */
  bk_t ref_cs[2 * OQ_NINT_MAX];      /* ref_closed_shell_bitmask */
  double ref_energy;                 /* ref_bitmask_energy */
} oq_ham;

/* ---- Hamiltonian providers (oracle/ham.c) ---- */
int64_t oq_integral_key(int i, int j, int k, int l);
oq_ham *oq_ham_create(int M, int na, int nb, const double *h, const double *eri, int64_t n8);
void oq_ham_destroy(oq_ham *H);
double oq_get_integral(const oq_ham *H, int i, int j, int k, int l);
double oq_ref_energy(const oq_ham *H);
void oq_get_fock_cs(const oq_ham *H, double *out);

/* ---- Slater rules (oracle/slater.c) ---- */
int oq_excitation_degree(const bk_t *k1, const bk_t *k2, int nint);
int oq_excitation_degree_spin(const bk_t *k1, const bk_t *k2, int nint);
void oq_get_double_excitation(const bk_t *d1, const bk_t *d2, int *exc, double *phase, int nint);
void oq_get_mono_excitation(const bk_t *d1, const bk_t *d2, int *exc, double *phase, int nint);
void oq_get_excitation(const bk_t *d1, const bk_t *d2, int *exc, int *degree, double *phase, int nint);
void oq_get_mono_excitation_spin(const bk_t *d1, const bk_t *d2, int *exc, double *phase, int nint);
void oq_get_double_excitation_spin(const bk_t *d1, const bk_t *d2, int *exc, double *phase, int nint);
void oq_decode_exc(const int *exc, int degree, int *h1, int *p1, int *h2, int *p2, int *s1, int *s2);
double oq_i_H_j(const oq_ham *H, const bk_t *ki, const bk_t *kj);
double oq_diag_H_mat_elem(const oq_ham *H, const bk_t *det);
double oq_diag_S_mat_elem(const bk_t *det, int nint);
double oq_get_s2(const bk_t *ki, const bk_t *kj, int nint);
double oq_i_H_j_mono_spin(const oq_ham *H, const bk_t *ki, const bk_t *kj, int spin);
double oq_i_H_j_double_spin(const oq_ham *H, const bk_t *ki, const bk_t *kj);
double oq_i_H_j_double_alpha_beta(const oq_ham *H, const bk_t *ki, const bk_t *kj);
double oq_mono_excitation_from_fock(const oq_ham *H, const bk_t *d1, int h, int p, int spin, double phase);
void oq_build_fock_tmp(const oq_ham *H, const bk_t *det, double *fock_diag_tmp /* 2*(M+1) */);
double oq_diag_H_mat_elem_fock(const oq_ham *H, const bk_t *ref, const bk_t *pert,
                               const double *fock_diag_tmp);

/* ---- bilinear alpha x beta structure (oracle/bilinear.c) ---- */
typedef struct {
  int64_t n_det;
  int nint;
  int n_alpha, n_beta;
  bk_t *alpha_unique, *beta_unique;  /* (nint, n) */
  /* 1-based arrays as in the reference, length n_det (+1 for loc) */
  int *rows, *columns, *order, *order_reverse, *columns_loc;
  int *transp_rows, *transp_columns, *transp_order, *transp_rows_loc;
  int *order_transp_reverse;
} oq_bilinear;
oq_bilinear *oq_bilinear_create(const bk_t *psi_det, int64_t n_det, int nint);
void oq_bilinear_destroy(oq_bilinear *B);

/* ---- sigma (oracle/sigma.c) ---- */
/* H_S2_u_0_nstates_openmp: v_0,s_0,u_0 are (sze, N_st) column-major in
 * psi_det order.  u_0 is not modified. */
void oq_H_S2_u_0_nstates(const oq_ham *H, const oq_bilinear *B, double *v_0, double *s_0,
                         const double *u_0, int N_st, int64_t sze, int nthreads);
/* The task-level work routine on a strided subset of rows (bilinear order). */
void oq_H_S2_u_0_work(const oq_ham *H, const oq_bilinear *B, double *v_t, double *s_t,
                      const double *u_t, int N_st, int64_t istart, int64_t iend, int64_t ishift,
                      int64_t istep, int nthreads);
/* count of connected off-diagonal pairs visited by the sigma traversal */
int64_t oq_sigma_count_pairs(const oq_ham *H, const oq_bilinear *B, int nthreads);

/* ---- dense small helpers (oracle/davidson.c) ---- */
void oq_jacobi_eigh(int n, const double *A, double *w, double *V);
typedef struct {
  double threshold;   /* threshold_davidson, on the squared residual norm */
  int sze_max;        /* davidson_sze_max */
  int s2_eig;         /* s2_eig */
  int state_following;
  double expected_s2, s_z2_sz;
  int nthreads;
  int max_iter;       /* safety cap on total inner iterations (0 = none) */
  int verbose;
} oq_davidson_params;
/* davidson_diag_hs2.  u_in (dim_in, N_st_diag) column-major: guess on entry,
 * eigenvectors on exit.  Returns number of sigma calls, or <0 on failure. */
int oq_davidson_diag_hs2(const oq_ham *H, const oq_bilinear *B, const bk_t *dets, double *u_in,
                         int64_t dim_in, int64_t sze, int N_st, int N_st_diag,
                         const oq_davidson_params *P, double *energies, double *s2_out);

/* ---- selection (oracle/selection.c) ---- */
typedef struct {
  int N;            /* requested size */
  int cur;
  int cap;          /* allocated size (2N) */
  int nint;
  bk_t *det;        /* (nint,2,cap) */
  double *val;
  double mini;
} oq_sel_buffer;
oq_sel_buffer *oq_sel_buffer_create(int N, int siz, int nint);
void oq_sel_buffer_destroy(oq_sel_buffer *b);
void oq_sel_buffer_sort(oq_sel_buffer *b);

typedef struct {
  /* psi sorted by decreasing state-averaged weight (psi_det_sorted) */
  int64_t n_det;
  int n_states;
  int nint;
  const bk_t *psi_det_sorted;      /* (nint,2,n_det) */
  const double *coef_transp;       /* (n_states, n_det): psi_selectors_coef_transp */
  int n_gen, n_sel;
  const oq_bilinear *B;            /* built on psi_det_sorted order (order = identity) */
  int fragment_count;              /* (elec_alpha_num - n_core)^2 */
  /* generator bitmask (hole mask, particle mask): one bitmask set */
  bk_t gen_hole_mask[2 * OQ_NINT_MAX], gen_part_mask[2 * OQ_NINT_MAX];
  /* instrumented counters */
  int64_t n_candidates;            /* C: candidates evaluated in fill_buffer_double */
  int64_t n_gathers;               /* G: integral + coefficient gathers in splash_pq */
  int64_t n_blocks;                /* blocks that reached splash_pq */
  int64_t n_blocks_total;          /* blocks visited (maskInd+1) */
  /* optional per-block log: (i_generator, s1, h1, s2, h2, sum_e_pt2(state 0), n_cand) */
  double *block_log;
  int64_t block_log_cap, block_log_n;
  /* optional block mask: if non-NULL, per generator g the blocks to keep are
   * those with keep[g_offsets[g] + maskInd_local] != 0 */
  const uint8_t *block_keep;
  const int64_t *block_keep_offsets;
} oq_sel_ctx;

void oq_select_connected(const oq_ham *H, oq_sel_ctx *S, int i_generator, const double *E0,
                         double *pt2, oq_sel_buffer *buf, int subset);
/* ZMQ_selection-equivalent over all generators (OpenMP over generators,
 * per-thread buffers merged as the collector does). */
void oq_select_all(const oq_ham *H, oq_sel_ctx *S, const double *E0, double *pt2,
                   oq_sel_buffer *out, int nthreads);
/* At most OQ_MAXNS states in a selection. */
#define OQ_MAXNS 8
void oq_select_tasks(const oq_ham *H, oq_sel_ctx *S, const double *E0, double *pt2, oq_sel_buffer *out,
                     int nthreads, int n_tasks, const int *task_gen, const int *task_subset, double *task_pt2);

#ifdef __cplusplus
}
#endif
#endif
