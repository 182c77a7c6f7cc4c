/*
 * MI355X engine selection and extensions.
 *
 * hmat_init_mi355x_interface() fills an hmat_interface_t exactly like
 * hmat_init_default_interface() (c_default_interface.cpp:242-252 of hmat-oss)
 * but with the GPU engine behind every slot it provides.
 *
 * Assembly on the GPU needs the matrix entries as a device-evaluable kernel.
 * A caller keeps the ordinary hmat_assemble_context_t and sets
 *     ctx.simple_compute = hmat_mi355x_kernel_compute;
 *     ctx.user_context   = &kernel;      (an hmat_mi355x_kernel_t)
 * hmat_mi355x_kernel_compute is a plain host implementation of the same
 * interaction (so the context also works with any other engine); this engine
 * recognises it and evaluates the kernel on the device instead.
 *
 * The hmx_* functions take and return plain pointers and sizes.  The *_device
 * variants work on device pointers in hmat numbering on the engine's stream
 * and synchronise before returning.
 */
#ifndef HMAT_MI355X_EXT_H
#define HMAT_MI355X_EXT_H

#include "hmat/hmat.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces hmat_init_default_interface (c_default_interface.cpp:242).
 * HMAT_DOUBLE_COMPLEX (HMAT_SIMPLE_COMPLEX) matrices are held as the 2 x 2
 * real representation [[Re, -Im], [Im, Re]] of every entry (real FP64 (FP32)
 * kernels throughout); complex arrays are passed as the reference passes them
 * (interleaved re, im).  They are stored whole and factorised by H-LU;
 * logdet, extract_diagonal, serialization of a factorised complex matrix and
 * complex scalars with a nonzero imaginary part in scale / add_identity /
 * axpy / gemm are not provided for them. */
HMAT_API void hmat_init_mi355x_interface(hmat_interface_t *i, hmat_value_t type);

typedef enum {
  HMAT_MI355X_KERNEL_EXPONENTIAL = 0, /* exp(-r / length) + nugget * delta_ij */
  HMAT_MI355X_KERNEL_GAUSSIAN = 1,    /* exp(-r^2 / length^2) + nugget * delta_ij */
  HMAT_MI355X_KERNEL_SHIFTED_EXPONENTIAL = 2,
  /* exp(-|x_i - x_j - shift e_z| / length) + nugget * delta_ij: non-symmetric */
  HMAT_MI355X_KERNEL_EXPONENTIAL_DLENGTH = 3,
  /* d/dlength of the exponential: (r / length^2) exp(-r / length) + nugget * delta_ij */
  HMAT_MI355X_KERNEL_GAUSSIAN_DLENGTH = 4,
  /* d/dlength of the Gaussian: (2 r^2 / length^3) exp(-r^2 / length^2) + nugget * delta_ij */
  HMAT_MI355X_KERNEL_HELMHOLTZ = 5
  /* complex (HMAT_DOUBLE_COMPLEX matrices): exp(i k r) / (4 pi r) with
   * r = sqrt(|x_i - x_j|^2 + h^2), k = wavenumber, h = regularization;
   * 1 / (4 pi h) + nugget on the diagonal (the interaction of
   * examples/c-cylinder.c:68-78 with a regularised distance) */
} hmat_mi355x_kernel_kind_t;

typedef struct {
  int kind;              /* hmat_mi355x_kernel_kind_t */
  int dimension;         /* spatial dimension of points (1..3) */
  int n;                 /* number of points (== DoFs) */
  const double *points;  /* n * dimension, user numbering */
  double length;         /* correlation length */
  double nugget;         /* added on the diagonal (user indices equal) */
  double shift;          /* HMAT_MI355X_KERNEL_SHIFTED_EXPONENTIAL only */
  double wavenumber;     /* HMAT_MI355X_KERNEL_HELMHOLTZ only */
  double regularization; /* HMAT_MI355X_KERNEL_HELMHOLTZ only (> 0) */
} hmat_mi355x_kernel_t;

/* hmat_interaction_func_t for an hmat_mi355x_kernel_t user context (the
 * simple_compute callback of assembly.cpp:75-91); writes a double, or a
 * complex double (re, im) for HMAT_MI355X_KERNEL_HELMHOLTZ. */
HMAT_API void hmat_mi355x_kernel_compute(void *user_context, int row, int col, void *result);

/* ---- flat helpers (no counterpart in hmat-oss; used by the test-suite, the
 *      benchmark and the Python binding) ---------------------------------- */

/* Last error message of a failing call on this thread ("" if none). */
HMAT_API const char *hmx_last_error(void);

/* y = alpha A x + beta y on device pointers (hmat numbering, ld = N).
 * 'precision' 0 = the FP64 matrix, 1 = the FP32 copy. */
HMAT_API int hmx_gemv_device(hmat_matrix_t *m, int precision, double alpha, const void *x,
                             double beta, void *y, int nrhs);

/* No counterpart in hmat-oss.  The same product as hmx_gemv_device for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): every temporary lives
 * in a fixed 64 MiB device array of the library, one per GPU, held until the
 * process ends; the call allocates nothing.  A matrix/nrhs whose temporaries
 * exceed 64 MiB is rejected with an error (hmx_last_error() names the bytes
 * needed).  Calls that overlap on one GPU (two streams, two threads) race on
 * the shared array; serialise them.  The matrix must be assembled. */
HMAT_API int hmx_gemv_device_noalloc(hmat_matrix_t *m, int precision, double alpha,
                                     const void *x, double beta, void *y, int nrhs);

/* Solves with the factor in place on device data (hmat numbering). */
HMAT_API int hmx_solve_device(hmat_matrix_t *m, int precision, void *b, int nrhs);

/* Mixed precision: factorise an FP32 copy of the FP64 matrix (LLt), keeping
 * the FP64 matrix for residuals. */
HMAT_API int hmx_factorize_mixed(hmat_matrix_t *m, int algo);

/* Iterative refinement on device FP64 data: x <- A^{-1} b with the FP32
 * factor as the inner solver and the FP64 H-matrix for the residual, until
 * max_col ||b - A x|| / ||b|| <= tol or max_iter sweeps.  Returns the number
 * of sweeps (< 0 on error); *resid receives the final relative residual. */
HMAT_API int hmx_solve_refined_device(hmat_matrix_t *m, double *b, int nrhs, double tol,
                                      int max_iter, double *resid);

/* No counterpart in hmat-oss, which handles one system per call
 * (assemble_generic c_wrapping.hpp:57, factorize_generic :139, solve_systems
 * :503).  Runs `count` independent systems: for each, assemble_generic(ms[i],
 * ctx[i]), factorise with `algo` (an FP32 factor refined against the FP64
 * matrix to `tol` when `mixed`, else a plain factor in the matrix's
 * precision) and solve b[i] (N_i x nrhs[i], user numbering, in place).
 * `threads` host threads take systems in turn, each system on its own
 * stream, so the device runs several systems' launches side by side.
 * seconds[i] receives the wall time of system i, status[i] 0 or 1. */
HMAT_API int hmx_run_batch(hmat_matrix_t **ms, hmat_assemble_context_t **ctx, int count,
                           int algo, int mixed, double tol, double **b, const int *nrhs,
                           int threads, double *seconds, int *status);

/* Partition in preorder: 5 ints per block node (row_offset, row_size,
 * col_offset, col_size, kind 0 full / 1 rk / 2 inner).  Returns the node
 * count; writes only when out != NULL. */
HMAT_API int hmx_partition(hmat_matrix_t *m, int *out);

HMAT_API int hmx_leaf_count(hmat_matrix_t *m);
/* info[6] = m, n, row_offset, col_offset, kind, rank (-1 for full) */
HMAT_API int hmx_leaf_info(hmat_matrix_t *m, int leaf, int precision, int *info);
/* All leaves at once: 6 ints per leaf as in hmx_leaf_info. */
HMAT_API int hmx_leaf_table(hmat_matrix_t *m, int precision, int *out);
/* Copies a leaf to the host as doubles: full -> a (m x n); rk -> a (m x k), b (n x k). */
HMAT_API int hmx_leaf_data(hmat_matrix_t *m, int leaf, int precision, double *a, double *b);

/* Timing / scheduling counters of the last assembly / factorisation / solve:
 * out[0..7] = t_assemble, t_factor, t_solve, tasks, waves, launches,
 *             segments, temp_bytes. */
HMAT_API int hmx_stats(hmat_matrix_t *m, double *out);

/* Makes the engine launch on the caller's HIP stream (hipStream_t passed as
 * void*; NULL restores the engine's own stream).  While a caller stream is
 * set, the *_device entry points are stream-ordered and do not synchronise. */
HMAT_API int hmx_set_stream(hmat_matrix_t *m, void *stream);

/* Upper bound on Rk panel capacity used for new matrices (default 160). */
HMAT_API void hmx_set_rank_capacity(int cap);

/* No counterpart in hmat-oss (the engine's own tuning).  The solves of an
 * LLt / LDLt factor apply the dense inverses of the diagonal blocks of at most
 * R rows (default 2048).  hmx_autotune_solve measures the solve of nrhs
 * right-hand sides with R = 512, 1024, 2048 and 4096 on the factorised matrix
 * and keeps the fastest (best_rows, best_ms may be NULL);
 * hmx_set_solve_block_rows sets R directly (for instance the tuned value on a
 * new matrix of the same structure). */
HMAT_API int hmx_autotune_solve(hmat_matrix_t *m, int nrhs, int *best_rows, double *best_ms);
HMAT_API int hmx_set_solve_block_rows(hmat_matrix_t *m, int rows);

/* No counterpart in hmat-oss (the engine's own scheduling).  Admissible blocks
 * of a complex matrix with at least `rows` rows and columns (default 4096) are
 * compressed by a randomized SVD over the whole device (dense evaluation,
 * GEMM, QR and SVD, cut at eps as the SVD compression cuts) instead of one
 * workgroup's partial ACA: a block of tens of thousands of rows with a rank
 * of several hundred otherwise holds one compute unit for minutes.  Takes
 * effect at the next assembly. */
HMAT_API int hmx_set_large_block_rows(hmat_matrix_t *m, int rows);

/* Test hooks of the task kernels (no counterpart in hmat-oss): the size of
 * one task record (hmx::FTask of the engine's ftask.hpp, which the tests
 * mirror), and one batched launch of `count` records of one task type on
 * device operands, as the factorisation executor launches them (GEMM tasks
 * in 64 x 64 tiles, workspace placed per task).  precision 0: FP64 kernels,
 * 1: FP32.  flags_out[0..1]: Rk capacity overflow, failing pivot. */
HMAT_API int hmx_ftask_size(void);
HMAT_API int hmx_run_tasks(int precision, int type, const void *tasks, int count, int *flags_out);

#ifdef __cplusplus
}
#endif

#endif
