/*
 * qpx.h -- C-ABI of libqpx, the MI355X-native determinant-driven CI hot path
 * (Davidson sigma H.Psi / S^2.Psi and CIPSI selection with Epstein-Nesbet
 * PT2) of Quantum Package v1.
 *
 * Plain C types only.  Every entry point returns 0 on success and a nonzero
 * code on failure; qpx_last_error() gives the message.  One context drives
 * one GPU; call it from one host thread.  Host arrays are owned by the
 * caller and only read unless documented; device memory is owned by the
 * context.
 *
 * Conventions (as the reference's providers):
 *   - determinants are det(N_int,2) 64-bit words, alpha then beta; orbital p
 *     (1-based) is bit (p-1)%64 of word (p-1)/64
 *     (src/Bitmask/bitmasks_module.f90:1-11);
 *   - matrices are column-major (Fortran);
 *   - two-electron integrals are 8-fold packed: <ij|kl> = (ik|jl) at index
 *     key-1, key = tri(tri(i,k),tri(j,l)), tri(a,b) = min+max(max-1)/2
 *     (src/Integrals_Bielec/map_integrals.irp.f:19-34).
 */
#ifndef QPX_H
#define QPX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qpx_ctx qpx_ctx;

enum { QPX_FP64 = 0, QPX_FP32 = 1 };

/* Create a context on HIP device `device`.
 * No reference counterpart: replaces the implicit provider state. */
int qpx_create(int device, qpx_ctx **ctx);
int qpx_destroy(qpx_ctx *ctx);
/* Copies the last error message (NUL-terminated) into buf. */
int qpx_last_error(qpx_ctx *ctx, char *buf, int len);
/* level 1: qpx_davidson writes one line per iteration (iteration, energy,
 * S^2 and squared residual of each state) to stderr, as the reference
 * writes its iteration table (src/Davidson/diagonalization_hs2.irp.f:402).
 * 0 (default): silent. */
int qpx_set_verbosity(qpx_ctx *ctx, int level);

/* Hamiltonian: h_core (M x M, column-major) = mo_mono_elec_integral
 * (src/Integrals_Monoelec/mo_mono_ints.irp.f:1-25); eri_packed (n8) holds
 * the MO integrals keyed as mo_bielec_integral
 * (src/Integrals_Bielec/map_integrals.irp.f:383-428).  Derives the
 * device-side Coulomb/exchange tables (mo_bi_integrals.irp.f:1344-1369).
 * Invalidates the wavefunction.  */
int qpx_set_hamiltonian(qpx_ctx *ctx, int mo_num, int n_alpha, int n_beta, const double *h_core,
                        const double *eri_packed, int64_t n8);

/* Wavefunction determinants psi_det (N_int,2,n_det).  Builds the unique
 * spin strings and the alpha x beta bilinear structure
 * (src/Determinants/spindeterminants.irp.f:56-136,387-597) plus the
 * spin-string connectivity used by sigma.  Call again after psi changes
 * (SOFT_TOUCH psi_det, src/Determinants/H_apply.irp.f:193). */
int qpx_set_wavefunction(qpx_ctx *ctx, int n_int, int64_t n_det, const uint64_t *psi_det);

/* H_S2_u_0_nstates_openmp (src/Davidson/u0Hu0.irp.f:15-66):
 * v = H u, s = S^2 u for n_st vectors.  u, v, s are (sze, n_st)
 * column-major in psi_det order, sze >= n_det.  u is not modified (the
 * reference permutes it in place and restores it).  prec = QPX_FP64 or
 * QPX_FP32 (FP32 u, integrals and accumulation; results returned as
 * double). */
int qpx_sigma(qpx_ctx *ctx, int prec, int n_st, int64_t sze, const double *u, double *v, double *s);

/* H_S2_u_0_nstates_openmp_work (src/Davidson/u0Hu0.irp.f:69-96) on device
 * memory: u_t, v_t, s_t are (n_st, n_det) state-fastest in the context's
 * bilinear (beta-column-major) order, of type double (QPX_FP64) or float
 * (QPX_FP32).  v_t and s_t are overwritten.  stream may be NULL (the
 * context's stream). */
int qpx_sigma_device(qpx_ctx *ctx, int prec, int n_st, const void *u_t, void *v_t, void *s_t, void *stream);

/* Permutation of the context's bilinear order: perm[k] = psi_det index
 * (0-based) of bilinear position k (psi_bilinear_matrix_order - 1). */
int qpx_get_bilinear_order(qpx_ctx *ctx, int64_t *perm);

/* Diagonal <I|H|I> for every determinant, psi_det order
 * (diag_H_mat_elem, src/Determinants/slater_rules.irp.f:1869-1929). */
int qpx_diag_h(qpx_ctx *ctx, double *h_jj);

/* davidson_diag_hs2 (src/Davidson/diagonalization_hs2.irp.f:1-444).
 * u_in (dim_in, n_st_diag) column-major: guess vectors on entry (columns
 * n_st..n_st_diag-1 are replaced by seeded random vectors as the reference
 * does with random_number), eigenvectors on exit.  energies and s2_out have
 * n_st_diag entries.  mixed = 0: FP64 sigma throughout; mixed = 1: FP32
 * sigma for the correction vectors until the squared residual falls below
 * switch_threshold, then FP64 sigma (Rayleigh quotient, orthogonalisation,
 * residuals always FP64).  n_sigma_out (may be NULL) receives
 * {number of FP64 sigma calls, number of FP32 sigma calls}.  The
 * reference iterates until converged; here the iteration count is capped at
 * 1000 (an error), so a threshold below the attainable precision fails
 * instead of looping. */
int qpx_davidson(qpx_ctx *ctx, int n_st, int n_st_diag, double threshold, int sze_max, int s2_eig,
                 int state_following, int mixed, double switch_threshold, uint64_t seed, double *u_in,
                 int64_t dim_in, double *energies, double *s2_out, int *n_sigma_out);

/* ZMQ_selection (plugins/Full_CI_ZMQ/zmq_selection.irp.f:1-100) over the
 * current wavefunction.  psi must have been set in psi_det_sorted order
 * (decreasing state-averaged weight, src/Determinants/determinants.irp.f:258-293);
 * coef (n_det, n_states) column-major are psi_coef_sorted.  The first
 * n_gen determinants are the generators, the first n_sel the selectors.
 * E0 (n_states) is pt2_E0_denominator.  On exit pt2 (n_states) holds the
 * EN-PT2 sum and out_dets (N_int,2,n_keep) / out_vals (n_keep) the n_keep
 * external determinants with the most negative min-state contribution
 * (ascending values), *n_out their count.  subset = 0 processes every
 * block; subset > 0 keeps the blocks with maskInd mod fragment_count ==
 * subset-1 (selection.irp.f:561-562). */
int qpx_select(qpx_ctx *ctx, int n_states, const double *coef, int n_gen, int n_sel, const double *E0,
               int64_t n_keep, int subset, double *pt2, uint64_t *out_dets, double *out_vals, int64_t *n_out);

/* Counters of the last qpx_select call: [0] candidates evaluated (unbanned,
 * nonzero coupling), [1] blocks processed, [2] blocks visited, [3] integral
 * and coefficient gathers. */
int qpx_select_stats(qpx_ctx *ctx, int64_t *stats4);

/* The EN-PT2 contribution of each listed generator, as select_connected
 * sums it over the generator's blocks (plugins/Full_CI_ZMQ/selection.irp.f:50-79,
 * the unit of the reference's PT2 tasks, run_pt2_slave.irp.f:2-136).  gens
 * (n_list) are 1-based indices into psi_det_sorted, in any order, without
 * repeats, each <= n_sel; the first n_sel determinants are the selectors.
 * gen_pt2 (n_states, n_list) column-major receives the contributions.  The
 * wavefunction, coef and E0 are as for qpx_select; nothing is selected. */
int qpx_select_generators(qpx_ctx *ctx, int n_states, const double *coef, int n_sel, const double *E0, int n_list,
                          const int *gens, double *gen_pt2);

/* The reference's selection task farm over an explicit task list
 * (plugins/Full_CI_ZMQ/zmq_selection.irp.f:54-100 queues "i_generator subset"
 * tasks; run_selection_slave.irp.f:46-106 runs each with
 * select_connected(i_generator, E0, pt2, buf, subset)).  Task t processes
 * the blocks of generator task_gen[t] (1-based index into psi_det_sorted,
 * <= n_sel) that pass the subset filter task_subset[t] (0 = every block,
 * s > 0 = maskInd mod fragment_count == s-1, fragment_count = n_alpha^2,
 * selection.irp.f:3-10,561-562); task_subset may be NULL (all 0).  A
 * generator may appear in several tasks with distinct nonzero subsets.
 * Ownership of blocks and bans follow the generator indices as in
 * qpx_select, whichever tasks are listed.  pt2 (n_states) receives the sum
 * over the tasks and out_dets / out_vals / *n_out the n_keep best external
 * determinants of the merged buffer (as qpx_select); task_pt2
 * (n_states, n_tasks) column-major, may be NULL, each task's EN-PT2. */
int qpx_select_tasks(qpx_ctx *ctx, int n_states, const double *coef, int n_sel, const double *E0, int64_t n_keep,
                     int n_tasks, const int *task_gen, const int *task_subset, double *pt2, uint64_t *out_dets,
                     double *out_vals, int64_t *n_out, double *task_pt2);

/* make_s2_eigenfunction (src/Determinants/occ_pattern.irp.f:251-309): the
 * determinants that share an occupation pattern with a determinant of
 * psi_det (N_int,2,n_det) and are not in it, for n_alpha alpha electrons.
 * Host code, no context.  *n_out receives their count; with out_dets
 * non-null and max_out >= *n_out they are written there (N_int,2,*n_out)
 * in a fixed order (patterns grouped by open shells and alpha electrons
 * among them, then alpha subsets in lexicographic order, then patterns in
 * word order).  Returns 0, 1 for bad arguments, 2 when max_out is too small
 * (nothing written, *n_out set). */
int qpx_make_s2_eigenfunction(int n_int, int64_t n_det, const uint64_t *psi_det, int n_alpha, int64_t max_out,
                              uint64_t *out_dets, int64_t *n_out);

/* make_s2_eigenfunction (src/Determinants/occ_pattern.irp.f:251-309) on the
 * context's GPU: the same determinants, in the same order, as
 * qpx_make_s2_eigenfunction for psi_det (N_int,2,n_det) (host arrays in and
 * out).  *n_out receives their count; with out_dets null only the count is
 * computed; with max_out < *n_out nothing is written and 4 is returned.
 * Patterns with more than 64 open shells, or groups whose alpha subsets
 * number above 2^40, are rejected. */
int qpx_s2_complete(qpx_ctx *ctx, int n_int, int64_t n_det, const uint64_t *psi_det, int n_alpha, int64_t max_out,
                    uint64_t *out_dets, int64_t *n_out);

#ifdef __cplusplus
}
#endif
#endif
