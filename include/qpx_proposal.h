/*
 * qpx_proposal.h -- block-level selection and the block-proposal network of
 * libqpx.  These entry points have no counterpart in the reference; they
 * serve the learned block proposal of CIPSI selection.  Conventions,
 * return codes and the context are those of qpx.h.
 */
#ifndef QPX_PROPOSAL_H
#define QPX_PROPOSAL_H
#include <stdint.h>
#include "qpx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* qpx_select restricted to a set of (generator, hole-pair) blocks, with
 * per-block results.  A block is one iteration of the (s1,i1,s2,i2) hole
 * loops of select_singles_and_doubles (plugins/Full_CI_ZMQ/selection.irp.f:412-420)
 * for one generator; blocks are numbered generator-major in that loop
 * order (qpx_select_count_blocks gives their number).  keep_mask (one byte
 * per block, NULL = every block) generalises the reference's subset
 * filter (selection.irp.f:561-562).  block_pt2 (n_blocks x n_states,
 * states fastest) receives each block's EN-PT2 sum and block_info
 * (n_blocks x 6) {generator, s1, h1, s2, h2, status} with status 1 =
 * enumerated, 2 = skipped by the reference's spot_isinwf test (an earlier
 * determinant of psi already owns the block), 0 = not kept.  Both may be
 * NULL. */
int qpx_select_blocks(qpx_ctx *ctx, int n_states, const double *coef, int n_gen, int n_sel, const double *E0,
                      int64_t n_keep, const uint8_t *keep_mask, double *pt2, uint64_t *out_dets, double *out_vals,
                      int64_t *n_out, double *block_pt2, int64_t *block_info);

/* Number of blocks of the first n_gen generators. */
int qpx_select_count_blocks(qpx_ctx *ctx, int n_gen, int64_t *n_blocks);

/* Block-proposal network: embedding (32 per hole orbital) + features ->
 * 3 ReLU layers of width 256 (BF16 weights, BF16 MFMA, FP32 accumulation)
 * -> log|e_PT2| of the block.  One network drives one GPU. */
typedef struct qpx_mlp qpx_mlp;

int qpx_mlp_create(int device, int n_orb, uint64_t seed, qpx_mlp **mlp);
int qpx_mlp_destroy(qpx_mlp *mlp);
int qpx_mlp_last_error(qpx_mlp *mlp, char *buf, int len);

/* desc (n x 4) int {h1, h2 (0-based), spin class 0 aa / 1 bb / 2 ab,
 * excitation degree of the generator from HF, clipped to 4}; fdesc (n x 3)
 * float {h_11, h_22, log|c_gen|} standardised by the caller.  One Adam step
 * (FP32 master weights) on the mean squared error against target; lr = 0
 * computes loss and gradients only. */
int qpx_mlp_train_step(qpx_mlp *mlp, int n, const int *desc, const float *fdesc, const float *target, float lr,
                       float *loss);
int qpx_mlp_predict(qpx_mlp *mlp, int n, const int *desc, const float *fdesc, float *out);

/* qpx_mlp_predict without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The descriptors are
 * staged and the activations kept in a file-scope device array of the
 * library (about 10 MB per GPU, held until the process ends); n must be in
 * 1..2048.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (several host threads or networks) race. */
int qpx_mlp_predict_static(qpx_mlp *mlp, int n, const int *desc, const float *fdesc, float *out);

/* `epochs` epochs of Adam steps over a training set of n samples (same
 * layouts as qpx_mlp_train_step) uploaded once: each epoch orders the
 * samples by a permutation drawn from a SplitMix64 stream seeded with seed
 * and takes consecutive batches of `batch` (a last partial batch is
 * dropped; 1 <= batch <= n), gathered on the device.  epoch_loss (epochs)
 * receives each epoch's mean batch loss. */
int qpx_mlp_train_epochs(qpx_mlp *mlp, int n, const int *desc, const float *fdesc, const float *target, float lr,
                         int epochs, int batch, uint64_t seed, float *epoch_loss);

/* Read (set = 0) or write (set = 1) a parameter tensor or, with grad = 1,
 * its last gradient: which = 0 W1, 1 W2, 2 W3, 3 b1, 4 b2, 5 b3, 6 w4,
 * 7 embedding, 8 b4; n elements, row-major. */
int qpx_mlp_param(qpx_mlp *mlp, int which, int grad, int set, float *data, int64_t n);

/* C (M x N) = bf16(A) (M x K) x bf16(B) (K x N), row-major, through the
 * network's MFMA GEMM kernel (layout check). */
int qpx_mlp_gemm_check(qpx_mlp *mlp, int M, int N, int K, const float *A, const float *B, float *C);

#ifdef __cplusplus
}
#endif
#endif
