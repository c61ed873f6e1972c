/*
 * ttg_chol.h -- C ABI of the MI355X (gfx950) tiled-Cholesky kernel library
 * (libttg_chol.so).
 *
 * This is the drop-in boundary below the POTRF/TRSM/SYRK/GEMM device-task
 * bodies of examples/potrf/potrf.h (devreal/ttg).  Each per-tile entry point
 * replaces one vendor call of the reference (cited below); the batched, tiled
 * and solver entry points are the launch-grouped forms the runtime and the
 * drivers use.
 *
 * Conventions (all entry points):
 *   - Column-major storage.  Every pointer argument named A, B, C, L, X,
 *     tiles, d_* is a DEVICE pointer; descriptors passed in host arrays are
 *     staged by the library.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Calls are stream-ordered and never synchronise the host, except the
 *     drivers that document a host read-back.
 *   - The return value is a launch status: TTG_CHOL_OK (0) or a negative
 *     TTG_CHOL_E* code; ttg_chol_last_error() describes the last failure of
 *     the calling thread.  Numerical failure (a non-positive pivot) is
 *     reported LAPACK-style through a device int `d_info`: 0, or the 1-based
 *     index of the first pivot that is not positive (NaN included).
 *   - Per-tile kernels accept any size up to TTG_CHOL_MAX_TILE; the tiled
 *     drivers use the "lower tile-packed" layout: NT = ceil(N/NB) and tile
 *     (m, n), n <= m, starts at tiles + tile_index(m, n) * NB * NB with
 *     tile_index(m, n) = n*NT - n*(n-1)/2 + (m - n), each tile NB x NB with
 *     leading dimension NB (the PaRSEC sym_block_cyclic tile layout of
 *     examples/potrf/pmw.h:52-59 for a 1 x 1 process grid).
 */
#ifndef TTG_CHOL_H
#define TTG_CHOL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TTG_CHOL_OK 0
#define TTG_CHOL_EINVAL (-1)   /* bad argument (size, leading dimension, pointer) */
#define TTG_CHOL_EHIP (-2)     /* the HIP runtime reported an error */
#define TTG_CHOL_ENOTSUP (-3)  /* size or option outside what the kernels support */
#define TTG_CHOL_ENOMEM (-4)   /* device allocation of a plan or workspace failed */
#define TTG_CHOL_ESTATE (-5)   /* call out of order (e.g. group_end without begin) */
#define TTG_CHOL_ENOCONV (-6)  /* iterative refinement did not reach the tolerance */

/* Largest tile order handled by the single-tile POTRF/TRSM kernels. */
#define TTG_CHOL_MAX_TILE 512

/* ---------------------------------------------------------------- */
/* library                                                           */
/* ---------------------------------------------------------------- */

/* Version string of the library. */
const char *ttg_chol_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char *ttg_chol_last_error(void);
/* Number of HIP devices visible to the library (replaces
 * ttg::device::num_devices() of ttg/ttg/device/device.h:150). */
int ttg_chol_num_devices(void);
/* Free every cached plan and workspace of the tiled drivers, the solves,
 * the mixed-precision solve and the POINV drivers (their device memory,
 * streams, events and graphs; plans are otherwise kept per matrix until the
 * process ends).  Synchronises the device first; no other call of the
 * library may be in flight.  The next call on a matrix builds its plan
 * again.  No reference counterpart: the reference's buffers are owned by
 * the PaRSEC runtime. */
int ttg_chol_release_plans(void);

/* Schedule knobs of the planned factorizations, read when a plan is built
 * (plans are cached per shape: call ttg_chol_release_plans() after a
 * change).  A value set here wins over the environment variable named in
 * brackets, which wins over the default:
 *   GROUP       panels per bulk split-FP16 update, 2..8  [TTG_CHOL_GROUP, 3]
 *   POTRF_CUS   CUs reserved for the diagonal chain of the low-precision
 *               plans, 0 (none) or 4..64                 [TTG_CHOL_POTRF_CUS, 8]
 *   TILE_GROUP  side of the square tile groups the bulk update walks, 1..8
 *                                                        [TTG_CHOL_TILE_GROUP, 2]
 *   PAIRED      1: grouped split-FP16 schedule, 0: one panel per update
 *                                                        [TTG_CHOL_PAIRED, 1]
 *   SWEEP       triangular solves (up to 16 right-hand sides per pass):
 *               k = 1..8 runs both sweeps as one persistent launch with at
 *               most k workgroups per CU, 0 as three launches per block
 *               step replayed from a graph              [TTG_CHOL_SWEEP, 2]
 * value < 0 clears the setting.  ttg_chol_get_knob returns the value a plan
 * built now would use.  No reference counterpart: the reference's schedule
 * is PaRSEC's, steered by the priorities of examples/potrf/potrf.h:743-752;
 * these knobs are what the autotuner (ttg_amd/autotune.py) searches. */
#define TTG_CHOL_KNOB_GROUP 0
#define TTG_CHOL_KNOB_POTRF_CUS 1
#define TTG_CHOL_KNOB_TILE_GROUP 2
#define TTG_CHOL_KNOB_PAIRED 3
#define TTG_CHOL_KNOB_SWEEP 4
#define TTG_CHOL_KNOB_COUNT 5
int ttg_chol_set_knob(int knob, int value);
int ttg_chol_get_knob(int knob);

/* ---------------------------------------------------------------- */
/* per-tile operations (FP64)                                        */
/* ---------------------------------------------------------------- */

/* In-place lower Cholesky of the n x n tile A:  A = L L^T.
 * Replaces hipsolverDpotrf(h, LOWER, n, A, lda, work, lwork, devInfo)
 * (examples/potrf/potrf.h:68-72) and its workspace query (:45-48): no
 * workspace is needed.  *d_info receives 0 or info_offset + the 1-based
 * local index of the first non-positive pivot; d_info may be NULL.  The
 * strict upper triangle is not referenced. */
int ttg_chol_dpotrf_tile(void *stream, int n, double *A, int lda, int *d_info, int info_offset);

/* B := B * L^{-T}  (side Right, Lower, Trans, NonUnit, alpha = 1) for the
 * m x n tile B and the n x n lower triangular L.
 * Replaces hipblasDtrsm(h, RIGHT, LOWER, T, NON_UNIT, m, n, &1, L, ldl, B, ldb)
 * (examples/potrf/potrf.h:299-304). */
int ttg_chol_dtrsm_rlt(void *stream, int m, int n, const double *L, int ldl, double *B, int ldb);

/* ttg_chol_dtrsm_rlt without any allocation: the same kernels and bitwise
 * the same result, with the inverses of L's 32 x 32 diagonal blocks staged in
 * a fixed 128 KiB device array of the library (one per GPU, held until the
 * process ends) instead of a per-stream workspace that the first call on a
 * stream allocates.  n <= TTG_CHOL_MAX_TILE (larger n: TTG_CHOL_EINVAL); no
 * hipMalloc, hipFree or host allocation on any call; refused
 * (TTG_CHOL_ESTATE) while a launch group is open on the stream; the stream
 * must belong to the current device.  All calls on one GPU share its array:
 * calls on one stream are ordered, calls that overlap on different streams of
 * one GPU race.  No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, serving loops, device-task bodies
 * run where allocation is forbidden). */
int ttg_chol_dtrsm_rlt_noalloc(void *stream, int m, int n, const double *L, int ldl, double *B, int ldb);

/* C := alpha A A^T + beta C on the lower triangle of the n x n tile C, A n x k.
 * Replaces hipblasDsyrk(h, LOWER, N, n, k, &alpha, A, lda, &beta, C, ldc)
 * (examples/potrf/potrf.h:429-434; the graph uses alpha = -1, beta = 1). */
int ttg_chol_dsyrk_ln(void *stream, int n, int k, double alpha, const double *A, int lda, double beta,
                      double *C, int ldc);

/* C := alpha A B^T + beta C,  A m x k, B n x k, C m x n.
 * Replaces hipblasDgemm(h, N, T, m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc)
 * (examples/potrf/potrf.h:555-561; the graph uses alpha = -1, beta = 1). */
int ttg_chol_dgemm_nt(void *stream, int m, int n, int k, double alpha, const double *A, int lda,
                      const double *B, int ldb, double beta, double *C, int ldc);

/* *d_result := ||x||_2 of n contiguous doubles.  Replaces hipblasDnrm2
 * (examples/potrf/potrf.h:85, DEBUG_TILES_VALUES builds).  Inside a launch
 * group (below) the norm sees the values the group's earlier calls leave in
 * x: a norm of memory that an operation already recorded in the group writes
 * runs after the group's batched launches, any other norm runs at once. */
int ttg_chol_dnrm2(void *stream, int n, const double *x, double *d_result);

/* ---------------------------------------------------------------- */
/* per-tile operations on FP32 tiles                                 */
/* ---------------------------------------------------------------- */
/* The four operations above for float tiles: what the reference's task
 * bodies, templated on the element type (examples/potrf/potrf.h:90-631),
 * call for MatrixT<float> (hipsolverSpotrf, hipblasStrsm, hipblasSsyrk,
 * hipblasSgemm).  POTRF factors in FP64 internally (a workspace from the
 * stream's or the group's work area) and rounds the factor to FP32;
 * *d_info as for ttg_chol_dpotrf_tile. */
int ttg_chol_spotrf_tile(void *stream, int n, float *A, int lda, int *d_info, int info_offset);
int ttg_chol_strsm_rlt(void *stream, int m, int n, const float *L, int ldl, float *B, int ldb);
int ttg_chol_ssyrk_ln(void *stream, int n, int k, float alpha, const float *A, int lda, float beta, float *C,
                      int ldc);
int ttg_chol_sgemm_nt(void *stream, int m, int n, int k, float alpha, const float *A, int lda, const float *B,
                      int ldb, float beta, float *C, int ldc);

/* ---------------------------------------------------------------- */
/* per-tile operations of the POINV graphs (FP64)                    */
/* ---------------------------------------------------------------- */
/* The tile operations the reference's TRTRI and LAUUM graphs call on the
 * host through LAPACK++/BLAS++ (examples/potrf/trtri_L.h, lauum.h), as
 * stream-ordered device calls.  Same conventions as above: column-major,
 * device pointers, tile orders <= TTG_CHOL_MAX_TILE for the triangular
 * operands, no host synchronization, launch-groupable; the work tiles the
 * triangular ones need come from the stream's (or the group's) workspace. */

/* A := A^{-1} for the lower triangle of the n x n tile A (non-unit); the
 * strict upper triangle is not referenced.  *d_info receives 0 or
 * info_offset + the 1-based index of the first exact zero on the diagonal
 * (the inverse is then not meaningful).  Replaces lapack::trtri(Lower,
 * NonUnit, ...) (examples/potrf/trtri_L.h:32). */
int ttg_chol_dtrtri_tile(void *stream, int n, double *A, int lda, int *d_info, int info_offset);

/* B := alpha L^{-1} B, L m x m lower non-unit, B m x n.  Replaces
 * blas::trsm(Left, Lower, NoTrans, NonUnit, m, n, alpha, L, B)
 * (examples/potrf/trtri_L.h:64-71; the graph uses alpha = 1). */
int ttg_chol_dtrsm_lln(void *stream, int m, int n, double alpha, const double *L, int ldl, double *B, int ldb);

/* B := alpha B L^{-1}, L n x n lower non-unit, B m x n.  Replaces
 * blas::trsm(Right, Lower, NoTrans, NonUnit, m, n, alpha, L, B)
 * (examples/potrf/trtri_L.h:109-116; the graph uses alpha = -1). */
int ttg_chol_dtrsm_rln(void *stream, int m, int n, double alpha, const double *L, int ldl, double *B, int ldb);

/* C := alpha A B + beta C, A m x k, B k x n.  Replaces
 * blas::gemm(NoTrans, NoTrans, ...) (examples/potrf/trtri_L.h:179-185). */
int ttg_chol_dgemm_nn(void *stream, int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                      int ldb, double beta, double *C, int ldc);

/* Lower triangle of the n x n tile A := L^T L, L its lower triangle; the
 * strict upper triangle is neither read nor written.  Replaces
 * lapack::lauum(Lower, ...) (examples/potrf/lauum.h:36). */
int ttg_chol_dlauum_tile(void *stream, int n, double *A, int lda);

/* B := alpha L^T B, L m x m lower non-unit, B m x n.  Replaces
 * blas::trmm(Left, Lower, Trans, NonUnit, ...) (examples/potrf/lauum.h:71-77). */
int ttg_chol_dtrmm_llt(void *stream, int m, int n, double alpha, const double *L, int ldl, double *B, int ldb);

/* C := alpha A^T A + beta C on the lower triangle of the n x n tile C, A
 * k x n.  Replaces blas::syrk(Lower, Trans, ...) (examples/potrf/lauum.h:116-121). */
int ttg_chol_dsyrk_lt(void *stream, int n, int k, double alpha, const double *A, int lda, double beta, double *C,
                      int ldc);

/* C := alpha A^T B + beta C, A k x m, B k x n.  Replaces
 * blas::gemm(Trans, NoTrans, ...) (examples/potrf/lauum.h:160-166). */
int ttg_chol_dgemm_tn(void *stream, int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                      int ldb, double beta, double *C, int ldc);

/* ---------------------------------------------------------------- */
/* launch grouping                                                   */
/* ---------------------------------------------------------------- */

/* Between group_begin(stream) and group_end(stream), the per-tile
 * operations above issued by the calling thread on `stream` are recorded
 * instead of launched; group_end launches them as one batched kernel per
 * operation kind (PLGSY tiles, POTRF, TRSM, SYRK+GEMM), in that order.
 * A call that reads memory an earlier call of the group writes, or writes
 * memory an earlier call reads or writes (by the address ranges of their
 * blocks), makes the group launch what it holds first and go on empty, so
 * dependent calls keep their program order.  This is how the task runtime
 * fuses every ready device task of a scheduling round into a few launches
 * (no counterpart in the reference, which issues one vendor call per task). */
int ttg_chol_group_begin(void *stream);
int ttg_chol_group_end(void *stream);
/* Number of kernel launches issued by the last group on this thread (all
 * its parts). */
int ttg_chol_group_last_launches(void);
/* Number of times the last group on this thread launched early because a
 * call depended on an earlier one.  No reference counterpart. */
int ttg_chol_group_last_splits(void);

/* ---------------------------------------------------------------- */
/* whole-matrix drivers (lower tile-packed layout)                   */
/* ---------------------------------------------------------------- */

/* Flags for the tiled drivers. */
#define TTG_CHOL_F_LOOKAHEAD 0x1u /* overlap POTRF/TRSM(k+1) with the update of step k */
#define TTG_CHOL_F_GRAPH 0x2u     /* capture the launch sequence once in a hipGraph and replay it */
#define TTG_CHOL_F_PROFILE 0x4u   /* time every launch with HIP events (launches eagerly, no graph) */
#define TTG_CHOL_F_DEFAULT (TTG_CHOL_F_LOOKAHEAD | TTG_CHOL_F_GRAPH)

/* FP64 tiled Cholesky of the N x N SPD matrix held in `tiles`, in place.
 * Runs the task graph of examples/potrf/potrf.h:670-764 with every panel
 * step's TRSMs fused into one launch and its SYRK+GEMM updates fused into
 * one launch.  Padding of edge tiles (N % NB != 0) is set by the driver.
 * *d_info = 0 or the global 1-based index of the first failing pivot (the
 * remaining steps then become no-ops on the device). */
int ttg_chol_dpotrf_tiled(void *stream, int N, int NB, double *tiles, int *d_info, unsigned flags);

/* Solve A X = B with the factor produced by ttg_chol_dpotrf_tiled
 * (B is N x nrhs, leading dimension ldb, overwritten by X).
 * (No counterpart in the reference: POTRS for the iterative refinement.) */
int ttg_chol_dpotrs_tiled(void *stream, int N, int NB, const double *tiles, int nrhs, double *B, int ldb);

/* B := (L L^T)^{-1} B with an FP32 factor L (lower tiles, FP64 right-hand
 * sides): the refinement solve of a mixed-precision factorization made
 * elsewhere (for instance by the task graph over float tiles).  The sweeps
 * run in FP32 on B rounded to FP32; the result is widened back.  The tiles
 * must be square NB x NB blocks: the padding of the last tile row and
 * column of the diagonal tile identity, the rest zero (ttg_chol_dpotrf_tiled
 * leaves it so).  No reference counterpart. */
int ttg_chol_spotrs_tiled(void *stream, int N, int NB, const float *tiles, int nrhs, double *B, int ldb);

/* R := B - A X, with A symmetric given by its lower tiles (FP64). */
int ttg_chol_dresidual_tiled(void *stream, int N, int NB, const double *tiles, int nrhs, const double *X,
                             int ldx, const double *B, int ldb, double *R, int ldr);

/* ---------------------------------------------------------------- */
/* POINV family (FP64, lower tile-packed layout, in place)          */
/* ---------------------------------------------------------------- */

/* tiles := L^{-1} for the lower triangular, non-unit factor held in the
 * lower tiles (the strict upper triangle of each diagonal tile is not
 * referenced or changed).  Replaces the TRTRI graph
 * trtri_LOWER::make_trtri_ttg(A, Diag::NonUnit, ...) of
 * examples/potrf/trtri_L.h:263-339, whose tasks call lapack::trtri (:27),
 * blas::trsm Left/Right (:49-56, :96-104) and blas::gemm NN (:178-185), in
 * that graph's step order.  *d_info receives 0 or the 1-based index of the
 * first zero diagonal entry, in which case the tiles are left unchanged
 * (LAPACK dtrtri); d_info may be NULL.  The padding of the last tile row is
 * reset as in ttg_chol_dpotrf_tiled.  NB % 32 == 0, NB <= 512. */
int ttg_chol_dtrtri_tiled(void *stream, int N, int NB, double *tiles, int *d_info);

/* tiles := L^T L (lower part) for the lower triangular factor L.  Replaces
 * lauum::make_lauum_ttg of examples/potrf/lauum.h:247-322 (lapack::lauum
 * :37, blas::trmm Left/Lower/Trans :72-80, blas::syrk Lower/Trans :123-128,
 * blas::gemm TN :172-180).  Uses a device copy of the tiles (allocated on
 * the first call for a matrix and kept with its plan). */
int ttg_chol_dlauum_tiled(void *stream, int N, int NB, double *tiles);

/* tiles := A^{-1} (lower part) from the Cholesky factor of A (the output of
 * ttg_chol_dpotrf_tiled): TRTRI then LAUUM.  Replaces
 * potri::make_potri_ttg of examples/potrf/potri.h:17-30.  *d_info as for
 * ttg_chol_dtrtri_tiled (LAUUM is skipped when it is nonzero). */
int ttg_chol_dpotri_tiled(void *stream, int N, int NB, double *tiles, int *d_info);

/* Working precisions of ttg_chol_dposv_mixed. */
#define TTG_CHOL_PREC_FP64 0 /* FP64 factor (time-to-solution reference) */
#define TTG_CHOL_PREC_FP32 1 /* FP32 TRSM + trailing updates, FP64 diagonal POTRF */
#define TTG_CHOL_PREC_F16X2 2 /* FP32 TRSM; trailing updates on FP16 MFMA with each panel row split into
                                 hi + lo FP16 parts under a per-row power-of-two scale (3 products,
                                 FP32 accumulation); FP64 diagonal POTRF; NB % 128 == 0 */
#define TTG_CHOL_PREC_F16 3   /* as F16X2 with the hi part only: one FP16 product per update (operands
                                 rounded to FP16 after row scaling, FP32 accumulation); for systems
                                 with kappa * 2^-11 well below 1 */
#define TTG_CHOL_PREC_BF16 4  /* one BF16 product per update (FP32 accumulation); kappa * 2^-8 << 1 */

/* Solve A X = B (A SPD, FP64 lower tiles, left unchanged; B, X N x nrhs,
 * nrhs <= 32) with a factorization in working precision `prec` followed by
 * FP64 iterative refinement until the normwise backward error
 * max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) <= tol or
 * max_iter refinement steps; it stops earlier when that backward error is
 * NaN or two steps in a row did not halve it (a diverging or stagnating
 * refinement; TTG_CHOL_ENOCONV unless tol was met).  The residual r = b - A x and the update
 * x += d run in FP64 on the device; the diagonal POTRF computes in FP64.
 * Synchronizes the host once per refinement step (to test convergence).
 * *iters = refinement steps taken, *berr = final backward error;
 * *d_info = factorization info (nonzero: not positive definite in the
 * working precision; X is then not set).  Returns TTG_CHOL_ENOCONV if the
 * tolerance was not reached.  Workspaces are owned by the library and reused
 * across calls with the same (N, NB, nrhs, prec).
 * (No counterpart in the reference: the mixed-precision solve.) */
int ttg_chol_dposv_mixed(void *stream, int N, int NB, const double *A, int nrhs, const double *B, int ldb, double *X,
                         int ldx, int prec, double tol, int max_iter, int *d_info, int *iters, double *berr);

/* A X = B by preconditioned conjugate gradients in FP64, preconditioned by
 * the factor that ttg_chol_dposv_mixed would use (precision prec) of A with
 * its diagonal scaled by 1 + shift (0 <= shift <= 1; a small shift, of the
 * order of the working precision's unit roundoff, lets a low-precision
 * factorization of an ill-conditioned A finish).  One CG per right-hand
 * side from X = M^{-1} B; each iteration is one FP64 product with A and one
 * solve with the factor.  Stops when every column's backward error (as
 * ttg_chol_dposv_mixed) is <= tol on the true residual, or after max_iter
 * iterations (TTG_CHOL_ENOCONV).  *iters: CG iterations; other arguments
 * and the plan cache as ttg_chol_dposv_mixed.  No reference counterpart:
 * the Krylov refinement of the mixed-precision solve (GMRES-IR / CG-IR). */
int ttg_chol_dposv_krylov(void *stream, int N, int NB, const double *A, int nrhs, const double *B, int ldb,
                          double *X, int ldx, int prec, double shift, double tol, int max_iter, int *d_info,
                          int *iters, double *berr);

/* The working factor of the last ttg_chol_dposv_mixed call with the same
 * (N, NB, nrhs, prec), widened to FP64, into L (device, num_tiles(N, NB) *
 * NB * NB doubles, the lower tile-packed layout): L in the lower part, the
 * input's values in the strict upper part of the diagonal tiles (POTRF does
 * not touch them), the padding of the edge tiles as the factorization set
 * it.  TTG_CHOL_ESTATE if no such call has run since the plans were built.
 * (No counterpart in the reference: it exposes the factor that the
 * mixed-precision solve keeps internally, so that its backward error can
 * be stated.) */
int ttg_chol_dposv_mixed_factor(void *stream, int N, int NB, int nrhs, int prec, double *L);

/* The factorization check of examples/potrf/testing_dpotrf.cc:277-335 in
 * tiles on the device: d_out[0] = ||A0||_inf, d_out[1] = ||L L^T - A0||_inf,
 * d_out[2] = d_out[1] / (d_out[0] N eps), eps = DBL_EPSILON (the reference
 * passes when d_out[2] <= 60).  A0 and L are FP64 lower tiles (device): L is
 * read from the lower triangle of each diagonal tile only (the trmm of
 * :298-299); the residual is mirrored from its lower part (:302-310);
 * ||A0||_inf is taken over the lower tiles as stored, diagonal tiles whole,
 * as getLAPACKMatrix (pmw.h:122-139) would hold them (:312).  L L^T is formed
 * in FP64 on MFMA in one batched launch (N^3/3 flops).  Uses library-owned
 * workspaces of 2 num_tiles(N, NB) NB^2 doubles and waits for its kernels
 * before it returns.  NB <= TTG_CHOL_MAX_TILE. */
int ttg_chol_dcheck_dpotrf_tiled(void *stream, int N, int NB, const double *A0, const double *L, double *d_out);

/* ---------------------------------------------------------------- */
/* generators                                                        */
/* ---------------------------------------------------------------- */

/* PLGSY of examples/potrf/plgsy.h:7-34 / core_plgsy.h:63-150 on the device:
 * every lower tile of the N x N symmetric random matrix with `bump` on the
 * diagonal, bit-identical to the reference generator. */
int ttg_chol_plgsy_tiled(void *stream, int N, int NB, double bump, unsigned long long seed, double *tiles);

/* CORE_dplgsy (examples/potrf/core_plgsy.h:63-150) for one m x n tile at
 * global offset (m0, n0) of a gM-row matrix, on the device; bit-identical to
 * the reference's host generator.  The device-task form of the PLGSY task
 * (examples/potrf/plgsy.h:12-33).  Inside a launch group the tiles of the
 * group are generated by one batched launch. */
int ttg_chol_dplgsy_tile(void *stream, double bump, int m, int n, double *A, int lda, int gM, int m0, int n0,
                         unsigned long long seed);

/* A = B B^T / N + alpha I into the lower tiles (diagonal tiles whole), with
 * B(i,j) = k/127, k an integer in [-127, 127] drawn from the PLASMA LCG
 * stream position i + j*N.  B B^T is formed exactly (integer data in FP64
 * MFMA), so A is bit-identical to the CPU construction.  `work` must hold
 * N*N doubles (the integer B).
 * (No counterpart in the reference: the test suite's SPD construction.) */
int ttg_chol_spd_tiled(void *stream, int N, int NB, double alpha, unsigned long long seed, double *tiles,
                       double *work);

/* The same construction with B FP64 uniform on [-1, 1): B(i,j) =
 * 2 (ran >> 11) 2^-53 - 1 from LCG stream position i + j*N (bit-identical
 * to the CPU construction), A = B B^T / N + alpha I formed on FP64 MFMA
 * (rounded: equal to a CPU product up to summation order).  Diagonal tiles
 * are written whole.  `work` holds N*N doubles.  (No counterpart in the
 * reference: the benchmark configurations' "B fp64 uniform[-1,1]".) */
int ttg_chol_spd_uniform_tiled(void *stream, int N, int NB, double alpha, unsigned long long seed, double *tiles,
                               double *work);

/* ---------------------------------------------------------------- */
/* diagnostics                                                       */
/* ---------------------------------------------------------------- */

/* Kernel kinds timed by TTG_CHOL_F_PROFILE runs. */
#define TTG_CHOL_PROF_UPDATE 0 /* batched SYRK+GEMM trailing update */
#define TTG_CHOL_PROF_TRSM 1   /* batched panel TRSM */
#define TTG_CHOL_PROF_POTRF 2  /* diagonal-tile POTRF */

/* Sum of the event-timed durations (ms), the launch count and the
 * algorithmic flops (DPLASMA counts: GEMM 2b^3, SYRK b^2(b+1), TRSM b^3,
 * POTRF b^3/3 per tile) of every launch of `kind` recorded by profiled
 * driver calls since the last read.  Waits for those launches to finish.
 * (No counterpart in the reference.) */
int ttg_chol_profile_read(int kind, double *ms_total, long long *launches, double *flops);
/* Profile every tiled factorization (also inside ttg_chol_dposv_mixed) as
 * with TTG_CHOL_F_PROFILE while `on` is nonzero. */
int ttg_chol_profile_enable(int on);

/* Lane-layout self test of the MFMA instructions the kernels use
 * (kind 0: FP64 16x16x4, kind 1: FP32 16x16x4, kind 2: FP16 16x16x32,
 * kind 3: FP16 32x32x16).  `out` (device; 64*64*4 doubles for kinds 0-1,
 * 512*64*4 for kind 2, 512*64*16 for kind 3) receives, for each one-hot B
 * position, the outputs of every lane.  The kernels' operand/result index
 * maps are checked against it by the test suite.
 * (No counterpart in the reference.) */
int ttg_chol_selftest_mfma_layout(void *stream, int kind, double *out);

/* ---- batched variable-size solves (no reference counterpart: many
 * independent small SPD systems in one launch) ------------------------------ */
#define TTG_CHOL_BATCH_MAX_N 1024

/* For i < count: factor the FP32 SPD matrix A[i] (n[i] x n[i], column-major,
 * leading dimension lda[i] >= n[i], 1 <= n[i] <= 1024; lower triangle read,
 * overwritten by L) and overwrite b[i] (n[i] floats) with the solution of
 * A x = b.  n, A, lda, b are HOST arrays (A[i], b[i] are device pointers);
 * d_info (device, count ints) receives 0 or the 1-based column of the first
 * non-positive pivot (b[i] then unchanged).  One persistent launch: the
 * matrices are dealt largest first from a device work counter.  cfg selects a
 * launch configuration (0 .. ttg_chol_batched_num_configs()-1, see
 * ttg_chol_batched_config), -1 the default; ttg_amd/autotune.py picks one per
 * workload.  Stream-ordered; the descriptors are staged through pinned host
 * memory (the next call waits for the previous copy only). */
int ttg_chol_sposv_batched(void *stream, int count, const int *n, float *const *A, const int *lda, float *const *b,
                           int *d_info, int cfg);
int ttg_chol_batched_num_configs(void);
/* Waves per workgroup and persistent workgroups per CU of configuration cfg. */
int ttg_chol_batched_config(int cfg, int *waves, int *wg_per_cu);

#ifdef __cplusplus
}
#endif
#endif /* TTG_CHOL_H */
