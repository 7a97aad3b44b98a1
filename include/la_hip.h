/*
 * la_hip.h -- C ABI of the MI355X (gfx950) dense matmul + factor-and-solve
 * library that stands under rust-la's `la::Matrix<f64>` hot path.
 *
 * Conventions (shared by every entry point):
 *   - Matrices are dense, ROW-MAJOR `double` (as `Matrix<T>` stores them,
 *     src/matrix/mod.rs:26-30), described by (rows, cols) and, in the
 *     device-resident variants, a leading dimension `ld` >= cols.
 *   - Every function returns an int32 status from `la_status`.  LA_ERR_SHAPE
 *     is the reference's `assert!` panic; LA_SINGULAR, LA_NOT_SQUARE,
 *     LA_NOT_SYMMETRIC and LA_NOT_SPD are the reference's `Option::None`.
 *   - Host-pointer entry points (`*_rm`) are synchronous: inputs are read from
 *     host memory, outputs are written to caller-allocated host memory before
 *     the call returns, matching the reference's synchronous semantics.
 *   - Device-pointer entry points (`*_dev`) take HBM pointers, enqueue on the
 *     context's stream and return without synchronising unless stated.
 *   - One `la_ctx` per host thread; a context is not safe to share between
 *     threads without external locking.
 *   - `piv` is the reference's permutation VECTOR (row i of P*A is row piv[i]
 *     of A, src/decomp/lu.rs:95-101), never LAPACK's swap list.
 */
#ifndef LA_HIP_H
#define LA_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum la_status {
    LA_OK = 0,
    LA_ERR_SHAPE = 1,       /* reference: assert! panic                       */
    LA_SINGULAR = 2,        /* reference: LUDecomposition::solve -> None      */
    LA_NOT_SQUARE = 3,      /* reference: CholeskyDecomposition::new -> None  */
    LA_NOT_SYMMETRIC = 4,   /* reference: CholeskyDecomposition::new -> None  */
    LA_NOT_SPD = 5,         /* reference: CholeskyDecomposition::new -> None  */
    LA_ERR_HIP = 6,         /* a HIP runtime call failed                      */
    LA_ERR_NOMEM = 7,       /* device allocation failed or workspace cap hit  */
    LA_NOT_CONVERGED = 8,   /* iterative refinement did not reach tolerance   */
    LA_ERR_ARG = 9,         /* null pointer, bad enum, bad leading dimension  */
    LA_ERR_TIMEOUT = 10,    /* a cross-workgroup wait exceeded its spin limit */
    LA_ERR_IO = 11          /* reference: unwrap() panic on a missing file or
                               an item that does not parse as f64            */
};

/* Factorisation precision for the mixed-precision solvers. */
enum la_prec {
    LA_PREC_FP64 = 0,       /* plain FP64 factor (no refinement needed)         */
    LA_PREC_FP32 = 1,       /* FP32 factor, FP32 MFMA trailing update           */
    LA_PREC_BF16 = 2,       /* BF16 operands with per-column 2^k scaling, FP32 acc */
    LA_PREC_FP16 = 3        /* FP16 operands with per-matrix/column 2^k scaling  */
};

typedef struct la_ctx la_ctx;

/* ---------------------------------------------------------------- context */
int         la_ctx_create(int device, la_ctx **out);
void        la_ctx_destroy(la_ctx *ctx);
int         la_ctx_synchronize(la_ctx *ctx);
void       *la_ctx_stream(la_ctx *ctx);          /* the hipStream_t in use  */
int         la_ctx_set_stream(la_ctx *ctx, void *hip_stream);
const char *la_status_string(int status);
const char *la_ctx_last_error(la_ctx *ctx);
int         la_device_info(la_ctx *ctx, int *num_cus, int64_t *hbm_bytes, int *clock_khz);
/* Cap (bytes) on the context's grow-only device workspace; 0 = no cap. */
int         la_ctx_set_workspace_limit(la_ctx *ctx, int64_t bytes);

/* ------------------------------------------------ kernel timing harness
 * When enabled, every launch of an instrumented kernel class is bracketed
 * by HIP events on the context stream; la_ctx_kernel_stats() resolves them
 * (synchronising the stream) and returns launch count, summed GPU time and
 * the ALGORITHMIC flop and byte counts of those launches (what the math
 * requires, not what the hardware moved). */
enum la_kernel_class {
    LA_K_GEMM64 = 0,    /* FP64 MFMA GEMM (Mul, trailing updates)            */
    LA_K_GEMM32 = 1,    /* FP32 MFMA GEMM (FP32 trailing updates)            */
    LA_K_GEMM16 = 2,    /* FP16/BF16-operand MFMA GEMM                       */
    LA_K_PANEL = 3,     /* LU panel factorisation (both panel kernels)       */
    LA_K_TRSV = 4,      /* blocked triangular solves                         */
    LA_K_RESID = 5,     /* FP64 residual r = b - A x                         */
    LA_K_TRSM = 6,      /* unit-lower TRSM of the LU U12 block               */
    LA_K_ROWMOVE = 7,   /* row interchanges                                  */
    LA_K_BATCH = 8,     /* batched persistent kernels                        */
    LA_K_PIVSEL = 9,    /* tournament pivot selection + row moves (CALU)     */
    LA_K_COUNT = 10
};
int la_ctx_set_profiling(la_ctx *ctx, int enable);
int la_ctx_reset_kernel_stats(la_ctx *ctx);
int la_ctx_kernel_stats(la_ctx *ctx, int kclass, int64_t *calls, double *total_ms,
                        double *flop, double *bytes);

/* ----------------------------------------------- drop-in host entry points */

/* C[m x p] = A[m x k] * B[k2 x p].
 * Replaces `impl Mul<&Matrix<T>> for &Matrix<T>` (src/matrix/mod.rs:957-980)
 * and its by-value forwards (mod.rs:982-998) for T = f64.
 * LA_ERR_SHAPE if k != k2 (mod.rs:961) or any extent is 0 (mod.rs:209). */
int la_dgemm_rm(la_ctx *ctx, int64_t m, int64_t k, int64_t k2, int64_t p,
                const double *A, const double *B, double *C);

/* dst = A * B into a caller-provided dst of (dst_rows x dst_cols).
 * Replaces `Matrix::mmul` (src/matrix/mmatrix.rs:82-98); LA_ERR_SHAPE on the
 * three asserts of mmatrix.rs:83-85. */
int la_dmmul_rm(la_ctx *ctx, int64_t m, int64_t k, int64_t k2, int64_t p,
                int64_t dst_rows, int64_t dst_cols,
                const double *A, const double *B, double *dst);

/* Partial-pivot LU of an m x n matrix.  LU receives the packed factors
 * (strict lower = L with unit diagonal implied, diagonal+upper = U),
 * piv the permutation vector (length m), *pospivsign the sign flag.
 * Replaces `LUDecomposition::new` (src/decomp/lu.rs:104-168).  Never fails
 * for a well-formed input (lu.rs:104: no Option). */
int la_dgetrf_rm(la_ctx *ctx, int64_t m, int64_t n, const double *A,
                 double *LU, uint64_t *piv, int *pospivsign);

/* X (m x nx) solving L*U*X = B(piv,:) from a packed LU.
 * Replaces `LUDecomposition::solve` (src/decomp/lu.rs:237-278):
 * LA_ERR_SHAPE if b_rows != m (lu.rs:240); then the diagonal is scanned in
 * the order of is_non_singular (lu.rs:174-182): LA_SINGULAR (-> None) at the
 * first exactly-zero entry lu[j][j], and, when m < n and none of the m
 * in-bounds entries is zero, LA_ERR_SHAPE for the out-of-bounds index the
 * reference panics on (j = m). */
int la_dgetrs_rm(la_ctx *ctx, int64_t m, int64_t n, const double *LU,
                 const uint64_t *piv, int64_t b_rows, int64_t nx,
                 const double *B, double *X);

/* getrf + getrs in one device round trip: `Matrix::solve`
 * (src/matrix/mod.rs:1030-1032), with la_dgetrs_rm's status order.  With
 * B = identity it is `Matrix::inverse` (mod.rs:1034-1037). */
int la_dgesv_rm(la_ctx *ctx, int64_t n_rows, int64_t n_cols, const double *A,
                int64_t b_rows, int64_t nx, const double *B, double *X);

/* Ainv = A^-1: `Matrix::inverse` (src/matrix/mod.rs:1034-1037), i.e.
 * LUDecomposition::new(A).solve(id): LA_ERR_SHAPE if A is not square
 * (mod.rs:1035), LA_SINGULAR (-> None) for an exact zero on U's diagonal.
 * The identity's row permutation is formed on the device and both
 * triangular solves with n right-hand sides run as the blocked MFMA TRSM
 * (64 x 64 diagonal blocks inverted, then one GEMM per block). */
int la_dinverse_rm(la_ctx *ctx, int64_t rows, int64_t cols, const double *A, double *Ainv);

/* Householder QR.  Replaces `QRDecomposition::new` (src/decomp/qr.rs:26-106):
 * QR (m x n) receives the reference's packed `qr` (R strictly above the
 * diagonal, Householder vector u_k = x - a_k e_k on and below the diagonal
 * of column k, columns with a_k = 0 left as they were) and rdiag
 * (min(m, n)) its `rdiag` (a_k = -sign(x_kk)|x|, +|x| when x_kk <= 0).
 * Blocked: 32-column panels, the panel's reflections applied to the rest as
 * I - V T V^T on MFMA.  get_h, get_r and is_full_rank (qr.rs:108-148) are
 * host post-processing of these two outputs; LA_ERR_SHAPE for an empty
 * matrix (mod.rs:209). */
int la_dgeqrf_rm(la_ctx *ctx, int64_t m, int64_t n, const double *A, double *QR,
                 double *rdiag);
/* `QRDecomposition::get_q` (qr.rs:150-194): Q (m x m) = the reflections
 * applied, last first, to the identity restricted to its first min(m, n)
 * diagonal entries. */
int la_dqr_q_rm(la_ctx *ctx, int64_t m, int64_t n, const double *QR, const double *rdiag,
                double *Q);
/* `QRDecomposition::solve` (qr.rs:196-237): X (n x nx) minimising
 * ||Q R X - B||_2.  LA_ERR_SHAPE if b_rows != m (qr.rs:200) or n > m
 * (is_full_rank reads rdiag past its end there, a panic in the reference);
 * LA_SINGULAR (-> None) if some rdiag[j] is exactly zero (qr.rs:201-203).
 * Two deliberate differences, both defects of the reference: its "Y = Q^T B"
 * step applies x - (<u,x>/u_kk) u, JAMA's formula for vectors normalised to
 * |u|^2 = 2 u_kk, to the unnormalised vectors of qr.rs:76 (|u|^2 = -2 a u_kk),
 * which is not the reflection; here Q^T B applies the reflections of
 * qr.rs:92-104.  And for m > n the reference builds an n x nx matrix from
 * m x nx values (a panic); here X is the n x nx least-squares solution. */
int la_dqrs_rm(la_ctx *ctx, int64_t m, int64_t n, const double *QR, const double *rdiag,
               int64_t b_rows, int64_t nx, const double *B, double *X);
/* `Matrix::pinverse` (src/matrix/mod.rs:1049-1057): P (n x m) =
 * (R^T R)^-1 A^T with R = get_r() of A's QR, the inverse by LU with the
 * identity (mod.rs:1034-1037).  LA_SINGULAR where the reference's
 * inverse().unwrap() panics (an exact zero pivot of R^T R). */
int la_dpinverse_rm(la_ctx *ctx, int64_t m, int64_t n, const double *A, double *P);

/* Cholesky factor L (full n x n, strict upper triangle zeroed).
 * Replaces `CholeskyDecomposition::new` (src/decomp/cholesky.rs:56-110):
 * LA_NOT_SQUARE (cholesky.rs:57-59), LA_NOT_SYMMETRIC (exact A[k][j] !=
 * A[j][k], cholesky.rs:89-91) and LA_NOT_SPD (d <= 0, cholesky.rs:97-100)
 * all map to the reference's None. */
int la_dpotrf_rm(la_ctx *ctx, int64_t rows, int64_t cols, const double *A, double *L);

/* X (n x nx) solving L*L^T*X = B.  Replaces `CholeskyDecomposition::solve`
 * (src/decomp/cholesky.rs:116-144); LA_ERR_SHAPE if b_rows != n (:118). */
int la_dpotrs_rm(la_ctx *ctx, int64_t n, const double *L, int64_t b_rows,
                 int64_t nx, const double *B, double *X);

/* ------------------------------------------- device-resident entry points
 * Same computations on HBM pointers with explicit leading dimensions; they
 * enqueue on the context stream.  Used by the benchmarks and by callers
 * that keep matrices resident. */

int la_dgemm_dev(la_ctx *ctx, int64_t m, int64_t n, int64_t k,
                 double alpha, const double *A, int64_t lda,
                 const double *B, int64_t ldb,
                 double beta, double *C, int64_t ldc);

/* In-place LU of the m x n matrix at A (ld lda).  d_piv (device, length m)
 * receives the permutation vector; d_info (device int[2]) receives
 * {pospivsign, status}; status is LA_OK or LA_ERR_TIMEOUT.  Size limit (no
 * counterpart in the reference, which has none): the column panels are
 * factored by at most min(256, CUs) co-resident workgroups of at most 192
 * rows each (FP64; 256 for the FP32 factor of la_dgesv_ir_dev), so
 * m <= 49152 on a 256-CU MI355X (65536 for the FP32 factor); beyond it the
 * call returns LA_ERR_ARG before launching anything.  Outer blocks whose rows
 * fit 96 CUs run with look-ahead on two process-wide streams with
 * complementary CU masks (panels of block j+1 beside the update of block j),
 * ordered after earlier work on ctx's stream and before later work on it by
 * events, so the call stays stream-ordered; the factor is bit-identical to
 * the single-stream schedule's, which LAHIP_LOOKAHEAD=0 selects. */
int la_dgetrf_dev(la_ctx *ctx, int64_t m, int64_t n, double *A, int64_t lda,
                  uint64_t *d_piv, int32_t *d_info);

/* la_dgetrf_dev without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * factorisation, its kernels and its results are those of la_dgetrf_dev;
 * its workspace (panel exchange slots, swap map, L11^-1 blocks, swap list)
 * is a fixed per-device scratch array of about 1.2 MiB instead of the
 * context's grow-only workspace.  Limit: min(m, n) <= 65536 (LA_ERR_ARG
 * beyond).  The scratch is shared by every call on one device and keeps the
 * last call's contents (each call resets what it reads first): calls that
 * overlap on one device, from any context or stream, race.  The timing
 * harness is suspended for the call.  It always runs the single-stream
 * schedule (no look-ahead), whose factor is bit-identical. */
int la_dgetrf_noalloc_dev(la_ctx *ctx, int64_t m, int64_t n, double *A, int64_t lda,
                          uint64_t *d_piv, int32_t *d_info);

/* Solve with a device-resident packed LU; X (m x nx, ld ldx) must not
 * alias B.  If d_status is non-NULL it receives LA_SINGULAR when U has an
 * exact zero on its diagonal (the reference's None, lu.rs:241-243). */
int la_dgetrs_dev(la_ctx *ctx, int64_t m, int64_t n, const double *LU, int64_t ldlu,
                  const uint64_t *d_piv, int64_t nx, const double *B, int64_t ldb,
                  double *X, int64_t ldx, int32_t *d_status);

/* Device-resident QR (la_dgeqrf_rm's factor in place of A, ld lda; d_rdiag
 * device, min(m, n)), Q (m x m, ld ldq), and least-squares solve (X n x nx;
 * B is not modified; status as la_dqrs_rm, checked before anything is
 * launched). */
int la_dgeqrf_dev(la_ctx *ctx, int64_t m, int64_t n, double *A, int64_t lda, double *d_rdiag);
int la_dqr_q_dev(la_ctx *ctx, int64_t m, int64_t n, const double *QR, int64_t ldqr,
                 const double *d_rdiag, double *Q, int64_t ldq);
int la_dqrs_dev(la_ctx *ctx, int64_t m, int64_t n, const double *QR, int64_t ldqr,
                const double *d_rdiag, int64_t nx, const double *B, int64_t ldb, double *X,
                int64_t ldx);

/* Device-resident inverse: A (ld lda) is overwritten with its packed LU,
 * d_piv receives the permutation, X (ld must equal n) the inverse; d_info
 * (device int[4]) = {pospivsign, status, singular flag (LA_SINGULAR if U
 * has an exact zero on its diagonal, X then undefined)}. */
int la_dinverse_dev(la_ctx *ctx, int64_t n, double *A, int64_t lda, uint64_t *d_piv,
                    double *X, int64_t ldx, int32_t *d_info);

/* In-place Cholesky of the n x n matrix at A (lower factor, upper zeroed);
 * d_status (device int) receives LA_OK / LA_NOT_SYMMETRIC / LA_NOT_SPD. */
int la_dpotrf_dev(la_ctx *ctx, int64_t n, double *A, int64_t lda, int32_t *d_status);

int la_dpotrs_dev(la_ctx *ctx, int64_t n, const double *L, int64_t ldl,
                  int64_t nx, const double *B, int64_t ldb, double *X, int64_t ldx);

/* ------------------------------------------ mixed-precision LU with FP64 IR
 * No reference counterpart: solves A x = b to FP64 accuracy by factoring a
 * low-precision copy of A and refining with FP64 residuals r = b - A x,
 * corrections and a normwise backward-error check
 *   eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)   (per RHS, max).
 * A is never modified. */
/* Option bit OR-ed into la_dgesv_ir_dev's prec (LA_PREC_FP32 only): factor
 * a randomised butterfly transform U^T A V of A (depth-2 random butterflies)
 * without pivoting instead of P A with partial pivoting.  The factorisation
 * then needs no pivot search across the column; refinement runs on A
 * itself, and a refinement that does not converge is repeated with the
 * partial-pivoting factor. */
#define LA_IR_RBT 0x100
/* Option bit OR-ed into la_dgesv_ir_dev's prec (any precision, combinable
 * with LA_IR_RBT): each refinement correction solves A d = r by GMRES on the
 * system preconditioned with the low-precision factor (FP64 Krylov basis,
 * FP64 products with A; GMRES-IR, Carson and Higham 2017/2018) instead of
 * applying the factor once.  Converges where the factor alone is too
 * inaccurate for plain refinement, e.g. BF16 factors at kappa = 1e6. */
#define LA_IR_GMRES 0x200
/* Option bit OR-ed into la_dgesv_ir_dev's prec (LA_PREC_FP32): the FP32
 * factor's pivots are chosen per 64-column panel by a tournament (CALU:
 * partial pivoting on candidate rows along a reduction tree, then the panel
 * factored without pivoting) instead of a search per column.  A different
 * permutation from partial pivoting, with the same growth in practice; a
 * refinement that does not converge is repeated with partial pivoting. */
#define LA_IR_TOURNAMENT 0x400

/* la_ir_report.method */
enum la_ir_method {
    LA_IR_METHOD_PIVOTED = 0,          /* partial-pivoting low-precision factor  */
    LA_IR_METHOD_RBT = 1,              /* butterfly transform, no pivoting       */
    LA_IR_METHOD_RBT_THEN_PIVOTED = 2, /* butterfly attempt did not converge;
                                          the result is the pivoted attempt's  */
    LA_IR_METHOD_TOURNAMENT = 3,       /* tournament-pivoted factor            */
    LA_IR_METHOD_TOURNAMENT_THEN_PIVOTED = 4
};

typedef struct la_ir_report {
    int32_t status;          /* LA_OK, LA_NOT_CONVERGED, LA_SINGULAR ...      */
    int32_t iterations;      /* refinement steps taken                        */
    int32_t prec;            /* la_prec used for the factorisation            */
    int32_t method;          /* la_ir_method                                  */
    double  backward_error;  /* final eta (max over RHS columns)              */
    double  initial_backward_error; /* eta of the first low-precision solve   */
    double  anorm_inf;       /* ||A||_inf                                     */
    double  factor_ms;       /* GPU time of the low-precision factorisation   */
    double  refine_ms;       /* GPU time of the refinement loop               */
    double  total_ms;        /* convert + factor + refine (all attempts)      */
    int32_t perturbed_pivots;/* butterfly form: pivots raised to the static
                                threshold 2^-20 max|U^T A V|                  */
    int32_t inner_iterations;/* LA_IR_GMRES: GMRES iterations, all attempts   */
} la_ir_report;

int la_dgesv_ir_dev(la_ctx *ctx, int64_t n, int64_t nrhs,
                    const double *A, int64_t lda, const double *B, int64_t ldb,
                    double *X, int64_t ldx, int prec, double tol, int max_iter,
                    la_ir_report *report);

/* Host-pointer form of the same solver. */
int la_dgesv_ir_rm(la_ctx *ctx, int64_t n, int64_t nrhs, const double *A,
                   const double *B, double *X, int prec, double tol, int max_iter,
                   la_ir_report *report);

/* ------------------------------------------------- batched uneven solvers
 * Problem i has order n[i]; its matrix starts at A + offA[i] (row-major,
 * ld = n[i]) and its right-hand sides at B + offB[i] (n[i] x nrhs[i],
 * row-major).  All arrays are device pointers.  The whole batch runs as ONE
 * persistent kernel launch (largest problems first from an atomic queue),
 * replayed from a captured hipGraph when the plan is reused. */

typedef struct la_batch_plan la_batch_plan;

/* Build a plan for a batch (host copies of the sizes are required to sort
 * and bucket the problems).  kind 0 = Cholesky solve (FP64; n <= 256,
 * nrhs 1, h_nrhs may be NULL); kind 1 = LU solve with FP64 refinement
 * (n <= 1024, 1 <= nrhs <= 8, prec = LA_PREC_FP64 / FP32 / FP16 working
 * precision).  The plan owns its device queue, the working copies of a
 * kind-1 batch and the captured graph. */
int la_batch_plan_create(la_ctx *ctx, int kind, int64_t count,
                         const int32_t *h_n, const int32_t *h_nrhs, int prec,
                         la_batch_plan **out);
/* Waits for the context stream, then frees the plan.  A plan must be
 * destroyed before the context it was created on. */
void la_batch_plan_destroy(la_batch_plan *plan);
/* Human-readable tile choice of the plan (autotuned or default). */
const char *la_batch_plan_describe(la_batch_plan *plan);

/* Batched Cholesky + two triangular solves, FP64:
 * L_i = chol(A_i) written to L + offA[i] (full, upper zeroed), X_i solves
 * A_i X_i = B_i.  status[i] gets LA_OK / LA_NOT_SYMMETRIC / LA_NOT_SPD per
 * problem; a failing problem never affects its neighbours.  No reference
 * counterpart as a batch: each problem is CholeskyDecomposition::new + solve
 * (src/decomp/cholesky.rs:56-144). */
int la_dpotrf_potrs_batched_dev(la_ctx *ctx, la_batch_plan *plan,
                                const int64_t *d_offA, const int64_t *d_offB,
                                const double *A, double *L,
                                const double *B, double *X, int32_t *d_status);

/* Batched LU with partial pivoting in the plan's working precision and FP64
 * iterative refinement, one persistent launch (graph-replayed):
 * X_i (n_i x nrhs_i at X + offB[i]) solves A_i X_i = B_i to normwise
 * backward error tol (max over the problem's columns).  d_status[i] =
 * LA_OK / LA_NOT_CONVERGED / LA_SINGULAR, d_iters[i] = refinement steps,
 * d_berr[i] = final eta.  With LA_PREC_FP64 the factorisation is plain FP64
 * LU, one solve, eta reported (tol and max_iter unused).  Each problem is
 * LUDecomposition::new + solve (src/decomp/lu.rs:104-278) with refinement
 * added; no batched counterpart in the reference. */
int la_dgesv_ir_batched_dev(la_ctx *ctx, la_batch_plan *plan,
                            const int64_t *d_offA, const int64_t *d_offB,
                            const double *A, const double *B, double *X,
                            double tol, int max_iter, int32_t *d_status,
                            int32_t *d_iters, double *d_berr);

/* Run the autotuner for `kind` and store the winners in `cache_path`
 * (JSON); later plans of this context use them.  kind 0 (Cholesky, orders
 * in [nmin, nmax] <= 256): panel width and MFMA row tiles per size bucket.
 * kind 1 (LU + refinement, orders in [nmin, nmax] <= 1024): workgroups per
 * CU of the persistent launch for each working precision (FP64, FP32,
 * FP16), timed on a synthetic batch of 4096 kappa = 1e3 systems.  cache_path may be NULL
 * to tune in memory only. */
int la_autotune(la_ctx *ctx, int kind, int nmin, int nmax, const char *cache_path);
/* Load a previously written autotune cache (either kind). */
int la_autotune_load(la_ctx *ctx, const char *cache_path);

/* ------------------------------------------------------- test generators
 * Seeded counter-based inputs (SplitMix64 finaliser over a counter).
 * Element i of stream s under seed q is mix(K + (i+1)*G) with
 * K = mix(q + (s+1)*G), G = 0x9E3779B97F4A7C15.  Uniform values are
 * 2*((bits>>11)*2^-53) - 1; normals are Box-Muller over index pairs. */
int la_fill_uniform_dev(la_ctx *ctx, uint64_t seed, uint64_t stream,
                        int64_t count, double *out);
int la_fill_normal_dev(la_ctx *ctx, uint64_t seed, uint64_t stream,
                       int64_t count, double *out);

/* A = U diag(sigma) V^T (n x n, row-major, FP64) with U, V products of
 * `nrefl` Householder reflectors from normal vectors (streams u_stream+i and
 * v_stream+i) and sigma_i = kappa^(-i/(n-1)).  If xstar/b are non-NULL,
 * b (n x nrhs) is filled from stream b_stream with normals and xstar gets
 * the exact solution V diag(1/sigma) U^T b, evaluated through the
 * reflectors. */
int la_gen_svd_matrix_dev(la_ctx *ctx, int64_t n, double kappa, int nrefl,
                          uint64_t seed, uint64_t u_stream, uint64_t v_stream,
                          double *A, int64_t lda,
                          int64_t nrhs, uint64_t b_stream, double *b, double *xstar);

/* Batch of SPD test matrices A_i = G_i G_i^T / n_i + I, G_i standard normal
 * from stream stream_base + i (row-major, A_i at A + offA[i]).  G_scratch
 * must have the same layout and size as A.  All pointers are device. */
int la_gen_spd_batch_dev(la_ctx *ctx, int64_t count, const int32_t *d_n,
                         const int64_t *d_offA, uint64_t seed, uint64_t stream_base,
                         double *A, double *G_scratch);

/* Batch of general test systems: A_i = U_i diag(sigma) V_i^T (n_i x n_i,
 * U_i, V_i products of nrefl <= 128 Householder reflectors, sigma_k =
 * kappa^(-k/(n_i-1))), B_i standard normal (n_i x nrhs_i) and the exact
 * solutions Xstar_i = V_i diag(1/sigma) U_i^T B_i.  Problem p draws its
 * reflectors from streams stream_base + 256p + k (U) and
 * stream_base + 256p + 128 + k (V), and B_p from stream stream_base + 256p
 * + 255.  n_i <= 1024, nrhs_i <= 8; all pointers are device. */
int la_gen_svd_batch_dev(la_ctx *ctx, int64_t count, const int32_t *d_n,
                         const int32_t *d_nrhs, const int64_t *d_offA,
                         const int64_t *d_offB, double kappa, int nrefl, uint64_t seed,
                         uint64_t stream_base, double *A, double *B, double *Xstar);

/* ------------------------------------------------------------ f32 surface
 * Matrix<f32> operations of src/matrix/simd.rs on the GPU.
 * simd_mul (simd.rs:189-219): C[m x p] = A[m x k] B[k2 x p] in FP32 on the
 * FP32 MFMA (its sums round differently from the reference's sequential
 * f32x4 loop; FP32 rounding bound k 2^-24 |A||B|).  LA_ERR_SHAPE if k != k2
 * (simd.rs:190). */
int la_sgemm_rm(la_ctx *ctx, int64_t m, int64_t k, int64_t k2, int64_t p,
                const float *A, const float *B, float *C);
int la_sgemm_dev(la_ctx *ctx, int64_t m, int64_t n, int64_t k, float alpha,
                 const float *A, int64_t lda, const float *B, int64_t ldb,
                 float beta, float *C, int64_t ldc);

/* simd_scale, simd_add, simd_sub, simd_elem_mul, simd_elem_div, simd_neg
 * (simd.rs:7-164): out = a * scalar, a + b, a - b, a * b, a / b, -a, one
 * IEEE operation per element (bit-identical to the reference's lanes).
 * Binary operations need equal shapes (LA_ERR_SHAPE, simd.rs:32-33 and
 * alike).  Unlike the reference, matrices of fewer than four elements get
 * every element computed (its scalar tail does not run for them). */
enum la_elementwise_op {
    LA_EW_SCALE = 0, LA_EW_ADD = 1, LA_EW_SUB = 2, LA_EW_MUL = 3, LA_EW_DIV = 4, LA_EW_NEG = 5
};
int la_selementwise_rm(la_ctx *ctx, int op, int64_t rows, int64_t cols,
                       int64_t b_rows, int64_t b_cols, const float *a, const float *b,
                       float scalar, float *out);
int la_selementwise_dev(la_ctx *ctx, int op, int64_t count, const float *a,
                        const float *b, float scalar, float *out);

/* simd_dot (simd.rs:166-187) of two equal-length columns (LA_ERR_SHAPE
 * otherwise, :167-168): the sum of every a[i] * b[i], accumulated in FP64
 * and rounded to FP32 once.  The reference's tail adds a[i] instead of
 * a[i] * b[i] for the last len mod 4 elements; that defect is not kept. */
int la_sdot_rm(la_ctx *ctx, int64_t rows, int64_t cols, int64_t b_rows, int64_t b_cols,
               const float *a, const float *b, float *out);
int la_sdot_dev(la_ctx *ctx, int64_t count, const float *a, const float *b, float *d_out);

/* ------------------------------------------ input format and rendering
 * `read_csv` (src/util.rs:9-35) for T = f64 with the item parser
 * `str::parse::<f64>` applied to each trimmed item: one row per line
 * (a trailing "\r" is removed with the newline), items separated by ','.
 * LA_ERR_IO where the reference's unwrap() panics (file missing, an item
 * that is not an f64 literal: empty items, hex floats, underscores),
 * LA_ERR_SHAPE where its asserts panic (a row with a different item count,
 * util.rs:25-27; an empty file, :33).  la_csv_shape counts; the other two
 * parse into a host buffer or into device memory (row-major, ld >= cols),
 * and return LA_ERR_ARG if rows/cols do not match the file. */
int la_csv_shape(const char *path, int64_t *rows, int64_t *cols);
int la_read_csv_rm(const char *path, int64_t rows, int64_t cols, double *out);
int la_read_csv_dev(la_ctx *ctx, const char *path, int64_t rows, int64_t cols,
                    double *d_out, int64_t ld);

/* The `Debug` rendering of a Matrix<f64> (src/matrix/mod.rs:815-840, what
 * `Matrix::print` writes): "\n", then per row '|', each element right-
 * aligned to the widest element's width plus one leading blank, " |\n".
 * Elements use Rust's `{:?}` for f64: the shortest round-trip digits, plain
 * notation with at least one fractional digit for 1e-4 <= |v| < 1e16 and
 * zero, otherwise d[.ddd]e<exp>; NaN, inf, -inf.  Writes at most cap-1
 * characters and a terminating NUL to buf (buf may be NULL with cap 0) and
 * returns the full length, or -1 for an empty shape. */
int64_t la_format_debug(int64_t rows, int64_t cols, const double *data, char *buf,
                        int64_t cap);

/* ------------------------------------------------------------ utilities */
/* eta = ||B - A X||_inf / (||A||_inf ||X||_inf + ||B||_inf), FP64 on device. */
int la_backward_error_dev(la_ctx *ctx, int64_t n, int64_t nrhs,
                          const double *A, int64_t lda, const double *X, int64_t ldx,
                          const double *B, int64_t ldb, double *eta);

#ifdef __cplusplus
}
#endif
#endif /* LA_HIP_H */
