/*
 * qpx_debug.h -- test and benchmark hooks of libqpx.  None of these has a
 * counterpart in the reference: they expose which sigma kernel ran, the
 * per-kernel and per-block timings the benchmark reports, the sparse-sigma
 * visit counts its byte model uses, and a layout check of the network's
 * GEMM.  Conventions, return codes and the context are those of qpx.h.
 * Production callers do not need them.
 */
#ifndef QPX_DEBUG_H
#define QPX_DEBUG_H
#include <stdint.h>
#include "qpx.h"
#include "qpx_proposal.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Force the sigma path of later qpx_sigma / qpx_sigma_device / qpx_davidson
 * calls: 0 automatic, 1 dense-map kernels (full-CI-like spaces), 2 scanning
 * kernel, 3 sub-string bucket kernel.  *last_path (may be NULL) receives the
 * path the previous sigma call took.  A forced path that does not apply to
 * the wavefunction makes the sigma call fail. */
int qpx_debug_force_sigma_path(qpx_ctx *ctx, int path, int *last_path);

/* One qpx_sigma_device call per repetition on `stream`; ms (2) receives the
 * average duration of its two kernel phases (dense path: column kernel, row
 * kernel; sparse path: the bucket kernel, 0), measured with HIP events on
 * that stream. */
int qpx_debug_sigma_kernel_times(qpx_ctx *ctx, int prec, int n_st, const void *u_t, void *v_t, void *s_t,
                                 void *stream, float *ms, int reps);

/* LDS read bandwidth of the context's device (bytes/s): two 1024-thread
 * workgroups per CU read 16 bytes per lane per iteration from LDS in the
 * conflict-free pattern, timed with HIP events (the peak the dense-map
 * sigma kernel's LDS roofline is quoted against). */
int qpx_debug_lds_bandwidth(qpx_ctx *ctx, double *bytes_per_s);

/* Entries the sparse sigma walk visits per call in each bucket class (sum
 * of squared bucket sizes): visited (5) = ab, a1, a2, b1, b2.  Builds the
 * buckets when they do not exist. */
int qpx_debug_sparse_stats(qpx_ctx *ctx, int64_t *visited);

/* Per-block workgroup durations of the last selection call, in 100 MHz
 * wall-clock ticks, in block enumeration order (0 for blocks not run); n
 * must equal that call's block count. */
int qpx_debug_select_block_ticks(qpx_ctx *ctx, int64_t *ticks, int64_t n);

/* C (M x N) = bf16(A) bf16(B) through the network's GEMM kernel with A
 * stored transposed (ta = 1, K x M row-major) and/or B transposed (tb = 1,
 * N x K) and the split-K path forced (split = 1). */
int qpx_debug_mlp_gemm(qpx_mlp *mlp, int M, int N, int K, const float *A, const float *B, int ta, int tb, int split,
                       float *C);

#ifdef __cplusplus
}
#endif
#endif
