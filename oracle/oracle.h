/*
 * CPU oracle for the tiled Cholesky path (test infrastructure only).
 *
 * This library is a plain-C restatement of the reference algorithm of
 * examples/potrf in devreal/ttg.  It exists to CHECK the HIP product path:
 * only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark load it.  Nothing in the product library links against it.
 *
 * Storage convention shared with the product library ("lower tile-packed"):
 *   NT = ceil(N / NB) tile rows/cols; only tiles (m, n) with n <= m exist.
 *   Tile (m, n) starts at  tiles + oracle_tile_index(m, n, NT) * NB * NB  and
 *   is column-major with leading dimension NB (as a PaRSEC sym_block_cyclic
 *   tile with mb = nb = NB, examples/potrf/pmw.h:52-59).  Edge tiles hold
 *   rows = N - m*NB (cols = N - n*NB) valid entries.
 *
 * Parity status of each function (how it is pinned) is written next to it.
 */
#ifndef TTG_CHOL_ORACLE_H
#define TTG_CHOL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* DPLASMA flop count, examples/potrf/potrf.h:30-33.
 * Pinned: closed form N^3/3 + N^2/2 + N/6 (tests). */
double oracle_flops_dpotrf(double n);

/* PLASMA LCG skip-ahead, examples/potrf/random.h:19-41.
 * Pinned: brute-force stepping, and the reference's own header compiled
 * into oracle/_ref. */
unsigned long long oracle_rnd64_jump(unsigned long long n, unsigned long long seed);

/* CORE_dplgsy for T=double, examples/potrf/core_plgsy.h:63-150.
 * Pinned: bit-exact against the reference header compiled into oracle/_ref. */
void oracle_core_plgsy(double bump, int m, int n, double *A, int lda,
                       int gM, int m0, int n0, unsigned long long seed);

/* Tile-packed layout helpers. */
long oracle_tile_index(int m, int n, int nt);
int oracle_num_tiles(int N, int NB);
/* Generate every lower tile with PLGSY exactly as examples/potrf/plgsy.h:7-34. */
void oracle_plgsy_tiled(int N, int NB, double bump, unsigned long long seed, double *tiles);
/* Dense column-major <-> tile-packed.  tiles_to_dense mirrors getLAPACKMatrix
 * (examples/potrf/pmw.h:122-139): entries outside the lower tiles become 0. */
void oracle_tiles_to_dense(int N, int NB, const double *tiles, double *A);
void oracle_dense_to_tiles(int N, int NB, const double *A, double *tiles);

/* Random SPD construction  A = B*B^T/N + alpha*I  (no reference counterpart).
 * B(i,j) = k/127 with k an integer in [-127,127] drawn from the PLASMA LCG
 * stream position i + j*N (so every entry of B*B^T is an integer multiple of
 * 1/127^2 and the sum is exact in FP64 in any order).
 * Pinned: integer arithmetic identity (tests compare to int64 sums). */
int oracle_spd_bint(long i, long j, long N, unsigned long long seed);
void oracle_spd_tiled(int N, int NB, double alpha, unsigned long long seed, double *tiles);
/* The same construction with B FP64 uniform on [-1, 1) (the benchmark
 * configurations' "B fp64 uniform[-1,1]"): B(i,j) = 2 (ran >> 11) 2^-53 - 1
 * from LCG stream position i + j*N (exact), A = B B^T / N + alpha I summed in
 * FP64 in k order (rounded; not bit-comparable with other summation orders).
 * Pinned: B values against a Python LCG; A against numpy's B B^T / N. */
double oracle_spd_uval(long i, long j, long N, unsigned long long seed);
void oracle_spd_uniform_tiled(int N, int NB, double alpha, unsigned long long seed, double *tiles);

/* Tile kernels, FP64, column-major (textbook definitions; pinned against
 * scipy LAPACK/BLAS fixtures and brute force in tests). */
/* examples/potrf/potrf.h:55-74,201: lower Cholesky in place; returns LAPACK info. */
int oracle_dpotrf_tile(int n, double *A, int lda);
/* examples/potrf/potrf.h:299-304,345: B := B * L^{-T}  (Right, Lower, Trans, NonUnit). */
void oracle_dtrsm_rlt(int m, int n, const double *L, int ldl, double *B, int ldb);
/* examples/potrf/potrf.h:429-434,480: C := alpha*A*A^T + beta*C, lower triangle only. */
void oracle_dsyrk_ln(int n, int k, double alpha, const double *A, int lda, double beta,
                     double *C, int ldc);
/* examples/potrf/potrf.h:555-561,609: C := alpha*A*B^T + beta*C. */
void oracle_dgemm_nt(int m, int n, int k, double alpha, const double *A, int lda,
                     const double *B, int ldb, double beta, double *C, int ldc);

/* Task graph of examples/potrf/potrf.h:90-764.  kinds: */
enum { ORACLE_POTRF = 0, ORACLE_TRSM = 1, ORACLE_SYRK = 2, ORACLE_GEMM = 3 };
/* DPLASMA priorities, potrf.h:743-752 (keys as in the reference:
 * POTRF(K), TRSM(M,K), SYRK(K,M), GEMM(M,N,K)).  Pinned: closed form. */
int oracle_priority(int kind, int nt, int k0, int k1, int k2);
/* Successor keys of a task (potrf.h:112-116,259-268,449-457,578-586).
 * Writes up to max entries of (kind, k0, k1, k2) into out[4*i..]; returns the
 * count.  Result sinks are reported with kind = -1 and key (I, J). */
int oracle_successors(int kind, int nt, int k0, int k1, int k2, int *out, int max);

/* Tiled Cholesky executing the graph above with a priority-ordered thread
 * pool (nthreads <= 0: all cores).  Each tile's update sequence is fixed by
 * the graph, so the result is bitwise identical under any schedule and equal
 * to oracle_dpotrf_tiled_seq.  Returns LAPACK info (global 1-based index). */
int oracle_dpotrf_tiled(int N, int NB, double *tiles, int nthreads);
int oracle_dpotrf_tiled_seq(int N, int NB, double *tiles);
/* Task counts seen by the last oracle_dpotrf_tiled run (tests check the
 * closed forms). */
void oracle_last_task_counts(long counts[4]);

/* check_dpotrf of examples/potrf/testing_dpotrf.cc:277-335:
 * ||L*L^T - A0||_inf / (||A0||_inf * N * eps).  A holds L in its lower part
 * (dense, col-major, lda = N; its upper part is not read); A0 is the
 * original matrix as getLAPACKMatrix returns it (pmw.h:122-139: the lower
 * tiles, diagonal tiles whole, zeros above).  The residual is mirrored from
 * its lower part (:302-310); ||A0||_inf is lange(Inf) of A0 as given (:312).
 * Pinned: planted symmetric perturbations of an exactly representable
 * L L^T give ||E||_inf / (||A0||_inf N eps) to rounding (tests). */
double oracle_check_dpotrf(const double *A, const double *A0, int N);
/* Same quantity computed tile by tile (no dense N*N work arrays), with
 * ||A0||_inf over the lower tiles as stored (diagonal tiles whole), as
 * getLAPACKMatrix would hold them.  Pinned: equal to the dense form. */
double oracle_check_dpotrf_tiled(int N, int NB, const double *Ltiles, const double *A0tiles,
                                 int nthreads);

/* Solve with a tiled factor: B := (L L^T)^{-1} B   (B dense N x nrhs, ldb). */
void oracle_dpotrs_tiled(int N, int NB, const double *Ltiles, int nrhs, double *B, int ldb);
/* R := B - A*X with A symmetric given by its lower tiles. */
void oracle_dresidual_tiled(int N, int NB, const double *Atiles, int nrhs, const double *X,
                            int ldx, const double *B, int ldb, double *R, int ldr);
/* Normwise backward error of a solve, max over columns:
 * ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), with ||A||_inf of the
 * full symmetric matrix.  Pinned: a planted error d e_j on an exact integer
 * solve gives d max_i |A_ij| / (||A||_inf ||x||_inf + ||b||_inf) (tests). */
double oracle_solve_backward_error(int N, int NB, const double *Atiles, int nrhs,
                                   const double *X, int ldx, const double *B, int ldb);
/* ||A||_inf of the symmetric matrix held by its lower tiles (upper part of
 * the diagonal tiles not read).  Pinned: numpy's norm of the mirrored dense
 * matrix, exactly on integer data. */
double oracle_norm_inf_sym_tiled(int N, int NB, const double *Atiles);

/* POINV family (examples/potrf/trtri_L.h, lauum.h, potri.h). */
void oracle_dtrsm_rln(int m, int n, double alpha, const double *L, int ldl, double *B, int ldb);
void oracle_dtrsm_lln(int m, int n, const double *L, int ldl, double *B, int ldb);
void oracle_dtrmm_llt(int m, int n, const double *L, int ldl, double *B, int ldb);
void oracle_dsyrk_lt(int n, int k, double alpha, const double *A, int lda, double beta, double *C, int ldc);
void oracle_dgemm_xn(int ta, int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                     int ldb, double beta, double *C, int ldc);
int oracle_dtrtri_tile(int n, double *A, int lda);
void oracle_dlauum_tile(int n, double *A, int lda);
int oracle_dtrtri_tiled(int N, int NB, double *tiles);
void oracle_dlauum_tiled(int N, int NB, double *tiles);
int oracle_dpotri_tiled(int N, int NB, double *tiles);
double oracle_check_dtrtri(const double *L, const double *Linv, int N);
double oracle_flops_dtrtri(double n);
double oracle_flops_dlauum(double n);
double oracle_flops_dpotri(double n);

#ifdef __cplusplus
}
#endif
#endif
