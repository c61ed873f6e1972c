/* ttg_graph.h -- C entry points of libttg_graph.so: the examples/potrf task
 * graph (POTRF/TRSM/SYRK/GEMM template tasks with the reference's keys,
 * successor rules and DPLASMA priorities) executed by the single-GPU TTG
 * runtime of ttg_amd/include/ttg.h over libttg_chol's gfx950 tile kernels.
 *
 * The matrix is host memory in the lower tile-packed layout: tile (m, n),
 * n <= m < NT = ceil(N / NB), starts at element (n*NT - n(n-1)/2 + m - n) NB^2
 * and is NB x NB column-major with leading dimension NB; edge tiles hold
 * N - m*NB valid rows / columns.
 */
#ifndef TTG_GRAPH_H
#define TTG_GRAPH_H

#ifdef __cplusplus
extern "C" {
#endif

/* Factor the tiles in place with the task graph: the body of
 * examples/potrf/testing_dpotrf.cc:227-262 (load -> make_potrf_ttg -> result
 * writer, then fence).  stats (optional, 4 longs): tasks run, scheduling
 * rounds, batched kernel launches, largest round.  *info (optional): 0, or
 * LAPACK's info (1-based column of the first non-positive pivot) when a tile
 * is not positive definite; the graph then stops cleanly after the rounds in
 * flight, leaving the tiles partly factored (the reference's POTRF task ends
 * the process instead, potrf.h:183-187).  Returns 0, -1 on a bad argument. */
int ttg_graph_dpotrf(int N, int NB, double *tiles, long *stats, int *info);

/* The POINV graphs of examples/potrf/trtri_L.h, lauum.h and potri.h as
 * device-task graphs over the tiles (load -> graph -> result writer, then
 * fence), with stats as above:
 *   ttg_graph_dtrtri  L := L^{-1} (lower, non-unit); *info (optional): 0 or
 *                     LAPACK's info of an exactly zero diagonal entry (the
 *                     graph then stops; the reference asserts, trtri_L.h:33)
 *   ttg_graph_dlauum  A := L^T L (lower)
 *   ttg_graph_dpotri  A := (L L^T)^{-1} from the factor L (TRTRI then LAUUM)
 * Return 0, -1 on a bad argument. */
int ttg_graph_dtrtri(int N, int NB, double *tiles, long *stats, int *info);
int ttg_graph_dlauum(int N, int NB, double *tiles, long *stats);
int ttg_graph_dpotri(int N, int NB, double *tiles, long *stats, int *info);

/* Mixed-precision SPD solve whose factorization runs through the task
 * graph: the potrf graph over FP32 tiles (device tasks on libttg_chol's S
 * tile kernels, the reference's bodies instantiated for float), then FP64
 * iterative refinement on the device (FP64 residual, FP32-factor solves)
 * until max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) <=
 * tol or max_iter corrections.  A: host FP64 lower tiles (unchanged); B and
 * X: host N x nrhs column-major (leading dimensions ldb, ldx).  *iters
 * (optional): corrections applied; *berr (optional): final backward error;
 * *info (optional): 0, or LAPACK's info of the FP32 factorization (X is then
 * not computed).  Returns 0, -1 on a bad argument, -2 on a device error. */
int ttg_graph_dposv_mixed(int N, int NB, const double *A, int nrhs, const double *B, int ldb, double *X, int ldx,
                          double tol, int max_iter, int *iters, double *berr, int *info);

/* PLGSY(N, bump, seed) into the tiles with one device task per tile (the
 * PLGSY template task of examples/potrf/plgsy.h:7-50; values bit-identical
 * to CORE_dplgsy).  Returns 0, -1 on a bad argument. */
int ttg_graph_dplgsy(int N, int NB, double bump, unsigned long seed, double *tiles);

#ifdef __cplusplus
}
#endif
#endif
