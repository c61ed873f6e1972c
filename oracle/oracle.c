/*
 * CPU oracle for the tiled Cholesky path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of examples/potrf (devreal/ttg).  Loaded only by the
 * test suite, the smoke check and the benchmark's CPU-baseline leg; the HIP
 * product library never links or calls it.  Citations are file:line in the
 * reference repository.
 *
 * The generators round every intermediate through a volatile so that no
 * compiler contraction into fused multiply-adds changes their bits; the tile
 * kernels may be contracted freely.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------ */
/* flops and generator                                                 */
/* ------------------------------------------------------------------ */

/* examples/potrf/potrf.h:31-33 (FMULS_POTRF + FADDS_POTRF). */
double oracle_flops_dpotrf(double n) {
  double fmuls = n * (((1. / 6.) * n + 0.5) * n + (1. / 3.));
  double fadds = n * (((1. / 6.) * n) * n - (1. / 6.));
  return fmuls + fadds;
}

#define ORACLE_RND64_A 6364136223846793005ULL /* random.h:19 */
#define ORACLE_RND64_C 1ULL                   /* random.h:20 */
#define ORACLE_RNDF_MUL 5.4210108624275222e-20f /* random.h:21 (a float) */

/* random.h:24-41: the LCG x' = a x + c advanced n steps by binary powering of
 * the affine map (a_k, c_k) -> (a_k^2, c_k (a_k + 1)). */
unsigned long long oracle_rnd64_jump(unsigned long long n, unsigned long long seed) {
  unsigned long long a_k = ORACLE_RND64_A, c_k = ORACLE_RND64_C, ran = seed;
  for (; n; n >>= 1) {
    if (n & 1ULL) ran = a_k * ran + c_k;
    c_k *= (a_k + 1ULL);
    a_k *= a_k;
  }
  return ran;
}

/* One PLGSY entry: 0.5f - ran * RndF_Mul evaluated in float
 * (core_plgsy.h:86,115: the constant is a float so the product and the
 * subtraction are single precision, then widened to double). */
static inline double plgsy_value(unsigned long long ran) {
  volatile float r = (float)ran;
  volatile float p = r * ORACLE_RNDF_MUL;
  volatile float v = 0.5f - p;
  return (double)v;
}

/* core_plgsy.h:63-150 for T = double (nbelem = 1). */
void oracle_core_plgsy(double bump, int m, int n, double *A, int lda, int gM, int m0, int n0,
                       unsigned long long seed) {
  unsigned long long jump = (unsigned long long)m0 + (unsigned long long)n0 * (unsigned long long)gM;
  if (m0 == n0) {
    /* diagonal tile: generate the lower part column by column
     * (core_plgsy.h:82-97), then bump the diagonal and mirror (:99-105). */
    for (int j = 0; j < n; j++) {
      unsigned long long ran = oracle_rnd64_jump(jump, seed);
      for (int i = j; i < m; i++) {
        A[i + (long)j * lda] = plgsy_value(ran);
        ran = ORACLE_RND64_A * ran + ORACLE_RND64_C;
      }
      jump += (unsigned long long)gM + 1ULL;
    }
    for (int j = 0; j < n; j++) {
      A[j + (long)j * lda] += bump;
      for (int i = 0; i < j; i++) A[(long)lda * j + i] = A[(long)lda * i + j];
    }
  } else if (m0 > n0) {
    /* strictly lower tile (core_plgsy.h:110-126) */
    for (int j = 0; j < n; j++) {
      unsigned long long ran = oracle_rnd64_jump(jump, seed);
      for (int i = 0; i < m; i++) {
        A[i + (long)j * lda] = plgsy_value(ran);
        ran = ORACLE_RND64_A * ran + ORACLE_RND64_C;
      }
      jump += (unsigned long long)gM;
    }
  } else {
    /* upper tile (core_plgsy.h:131-148): row-wise stream from (n0, m0) */
    jump = (unsigned long long)n0 + (unsigned long long)m0 * (unsigned long long)gM;
    for (int i = 0; i < m; i++) {
      unsigned long long ran = oracle_rnd64_jump(jump, seed);
      for (int j = 0; j < n; j++) {
        A[(long)j * lda + i] = plgsy_value(ran);
        ran = ORACLE_RND64_A * ran + ORACLE_RND64_C;
      }
      jump += (unsigned long long)gM;
    }
  }
}

long oracle_tile_index(int m, int n, int nt) {
  /* column-major over the lower tiles: column n holds tiles n..nt-1 */
  return (long)n * nt - (long)n * (n - 1) / 2 + (m - n);
}

int oracle_num_tiles(int N, int NB) {
  int nt = (N + NB - 1) / NB;
  return nt * (nt + 1) / 2;
}

static inline int tile_dim(int N, int NB, int t) {
  int r = N - t * NB;
  return r < NB ? r : NB;
}

static inline double *tile_ptr(double *tiles, int m, int n, int nt, int NB) {
  return tiles + oracle_tile_index(m, n, nt) * (long)NB * NB;
}
static inline const double *tile_cptr(const double *tiles, int m, int n, int nt, int NB) {
  return tiles + oracle_tile_index(m, n, nt) * (long)NB * NB;
}

/* plgsy.h:19-30: tempmm/tempnn are the edge-tile sizes, ldam = NB,
 * gM = N, (m0, n0) = (I*NB, J*NB). */
void oracle_plgsy_tiled(int N, int NB, double bump, unsigned long long seed, double *tiles) {
  int nt = (N + NB - 1) / NB;
  for (int n = 0; n < nt; n++)
    for (int m = n; m < nt; m++)
      oracle_core_plgsy(bump, tile_dim(N, NB, m), tile_dim(N, NB, n), tile_ptr(tiles, m, n, nt, NB),
                        NB, N, m * NB, n * NB, seed);
}

/* pmw.h:122-139 */
void oracle_tiles_to_dense(int N, int NB, const double *tiles, double *A) {
  int nt = (N + NB - 1) / NB;
  for (long j = 0; j < N; j++)
    for (long i = 0; i < N; i++) {
      int m = (int)(i / NB), n = (int)(j / NB);
      if (n <= m)
        A[i + j * N] = tile_cptr(tiles, m, n, nt, NB)[(i % NB) + (j % NB) * NB];
      else
        A[i + j * N] = 0.0;
    }
}

void oracle_dense_to_tiles(int N, int NB, const double *A, double *tiles) {
  int nt = (N + NB - 1) / NB;
  for (int n = 0; n < nt; n++)
    for (int m = n; m < nt; m++) {
      double *t = tile_ptr(tiles, m, n, nt, NB);
      int mb = tile_dim(N, NB, m), nb = tile_dim(N, NB, n);
      for (int j = 0; j < nb; j++)
        for (int i = 0; i < mb; i++) t[i + (long)j * NB] = A[(long)(m * NB + i) + (long)(n * NB + j) * N];
    }
}

int oracle_spd_bint(long i, long j, long N, unsigned long long seed) {
  unsigned long long ran = oracle_rnd64_jump((unsigned long long)(i + j * N), seed);
  unsigned long long hi = ran >> 32;
  return (int)((hi * 255ULL) >> 32) - 127;
}

static void gemm_nt_acc(int m, int n, int k, double alpha, const double *restrict A, int lda,
                        const double *restrict B, int ldb, double *restrict C, int ldc);
static int host_cores(void);

/* B(i,j) of the uniform construction: 2 u - 1 with u = (ran >> 11) 2^-53,
 * ran the LCG state at stream position i + j*N.  Both steps are exact, so B
 * takes every value k 2^-52 - 1, k < 2^53, of [-1, 1). */
double oracle_spd_uval(long i, long j, long N, unsigned long long seed) {
  unsigned long long ran = oracle_rnd64_jump((unsigned long long)(i + j * N), seed);
  return 2.0 * ((double)(ran >> 11) * 0x1p-53) - 1.0;
}

typedef struct {
  int N, NB, nt, nthreads, tid;
  double alpha, c;
  const double *B;
  double *tiles;
} spd_arg;

static void *spd_worker(void *p) {
  spd_arg *a = (spd_arg *)p;
  const int N = a->N, NB = a->NB, nt = a->nt;
  const double c = a->c;
  long idx = 0;
  for (int n = 0; n < nt; n++)
    for (int m = n; m < nt; m++, idx++) {
      if (idx % a->nthreads != a->tid) continue;
      double *t = a->tiles + oracle_tile_index(m, n, nt) * (long)NB * NB;
      int mb = N - m * NB < NB ? N - m * NB : NB, nb = N - n * NB < NB ? N - n * NB : NB;
      for (int j = 0; j < nb; j++)
        for (int i = 0; i < mb; i++) t[i + (long)j * NB] = 0.0;
      /* integer-valued operands (oracle_spd_tiled): every product and
       * partial sum is an exact integer below 2^53, so the blocked kernel's
       * order does not matter; uniform operands round in its k order */
      gemm_nt_acc(mb, nb, N, 1.0, a->B + (long)m * NB, N, a->B + (long)n * NB, N, t, NB);
      for (int j = 0; j < nb; j++)
        for (int i = 0; i < mb; i++) {
          volatile double v = t[i + (long)j * NB] * c;
          if (m == n && i == j) v = v + a->alpha;
          t[i + (long)j * NB] = v;
        }
    }
  return NULL;
}

/* tiles := c B B^T + alpha I (lower tiles), on every host core */
static void spd_run(int N, int NB, int nt, double alpha, double c, const double *B, double *tiles) {
  int nth = host_cores();
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nth);
  spd_arg *args = (spd_arg *)malloc(sizeof(spd_arg) * nth);
  for (int t = 0; t < nth; t++) {
    args[t] = (spd_arg){N, NB, nt, nth, t, alpha, c, B, tiles};
    pthread_create(&th[t], NULL, spd_worker, &args[t]);
  }
  for (int t = 0; t < nth; t++) pthread_join(th[t], NULL);
  free(th);
  free(args);
}

void oracle_spd_tiled(int N, int NB, double alpha, unsigned long long seed, double *tiles) {
  int nt = (N + NB - 1) / NB;
  /* integer B (as doubles), column-major, generated column by column from the LCG */
  double *B = (double *)malloc(sizeof(double) * (size_t)N * N);
  for (long j = 0; j < N; j++) {
    unsigned long long ran = oracle_rnd64_jump((unsigned long long)(j * N), seed);
    for (long i = 0; i < N; i++) {
      B[i + j * N] = (double)((int)((((ran >> 32) * 255ULL) >> 32)) - 127);
      ran = ORACLE_RND64_A * ran + ORACLE_RND64_C;
    }
  }
  spd_run(N, NB, nt, alpha, 1.0 / (127.0 * 127.0 * (double)N), B, tiles);
  free(B);
}

void oracle_spd_uniform_tiled(int N, int NB, double alpha, unsigned long long seed, double *tiles) {
  int nt = (N + NB - 1) / NB;
  double *B = (double *)malloc(sizeof(double) * (size_t)N * N);
  for (long j = 0; j < N; j++) {
    unsigned long long ran = oracle_rnd64_jump((unsigned long long)(j * N), seed);
    for (long i = 0; i < N; i++) {
      B[i + j * N] = 2.0 * ((double)(ran >> 11) * 0x1p-53) - 1.0;
      ran = ORACLE_RND64_A * ran + ORACLE_RND64_C;
    }
  }
  spd_run(N, NB, nt, alpha, 1.0 / (double)N, B, tiles);
  free(B);
}

/* ------------------------------------------------------------------ */
/* tile kernels                                                        */
/* ------------------------------------------------------------------ */

/* C += alpha * A * B^T on an m x n block, A m x k, B n x k (col-major).
 * Register-blocked 8 x 4 micro-kernel, k blocked by 128. */
static void gemm_nt_acc(int m, int n, int k, double alpha, const double *restrict A, int lda,
                        const double *restrict B, int ldb, double *restrict C, int ldc) {
  const int KC = 128;
  for (int k0 = 0; k0 < k; k0 += KC) {
    int kc = (k - k0) < KC ? (k - k0) : KC;
    int j0 = 0;
    for (; j0 + 4 <= n; j0 += 4) {
      int i0 = 0;
      for (; i0 + 8 <= m; i0 += 8) {
        double c[4][8];
        memset(c, 0, sizeof c);
        for (int l = k0; l < k0 + kc; l++) {
          const double *a = A + i0 + (long)l * lda;
          const double *b = B + j0 + (long)l * ldb;
          double b0 = b[0], b1 = b[1], b2 = b[2], b3 = b[3];
          for (int ii = 0; ii < 8; ii++) {
            double av = a[ii];
            c[0][ii] += av * b0;
            c[1][ii] += av * b1;
            c[2][ii] += av * b2;
            c[3][ii] += av * b3;
          }
        }
        for (int jj = 0; jj < 4; jj++)
          for (int ii = 0; ii < 8; ii++) C[i0 + ii + (long)(j0 + jj) * ldc] += alpha * c[jj][ii];
      }
      for (; i0 < m; i0++)
        for (int jj = 0; jj < 4; jj++) {
          double s = 0.0;
          for (int l = k0; l < k0 + kc; l++) s += A[i0 + (long)l * lda] * B[j0 + jj + (long)l * ldb];
          C[i0 + (long)(j0 + jj) * ldc] += alpha * s;
        }
    }
    for (; j0 < n; j0++)
      for (int i = 0; i < m; i++) {
        double s = 0.0;
        for (int l = k0; l < k0 + kc; l++) s += A[i + (long)l * lda] * B[j0 + (long)l * ldb];
        C[i + (long)j0 * ldc] += alpha * s;
      }
  }
}

void oracle_dgemm_nt(int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                     int ldb, double beta, double *C, int ldc) {
  if (beta != 1.0)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++) C[i + (long)j * ldc] = (beta == 0.0) ? 0.0 : beta * C[i + (long)j * ldc];
  if (k > 0 && alpha != 0.0) gemm_nt_acc(m, n, k, alpha, A, lda, B, ldb, C, ldc);
}

void oracle_dsyrk_ln(int n, int k, double alpha, const double *A, int lda, double beta, double *C,
                     int ldc) {
  if (beta != 1.0)
    for (int j = 0; j < n; j++)
      for (int i = j; i < n; i++) C[i + (long)j * ldc] = (beta == 0.0) ? 0.0 : beta * C[i + (long)j * ldc];
  if (k <= 0 || alpha == 0.0) return;
  for (int j0 = 0; j0 < n; j0 += 4) {
    int jn = (n - j0) < 4 ? (n - j0) : 4;
    /* lower part of the jn x jn diagonal block */
    for (int j = j0; j < j0 + jn; j++)
      for (int i = j; i < j0 + jn; i++) {
        double s = 0.0;
        for (int l = 0; l < k; l++) s += A[i + (long)l * lda] * A[j + (long)l * lda];
        C[i + (long)j * ldc] += alpha * s;
      }
    /* rectangle below it */
    if (j0 + jn < n)
      gemm_nt_acc(n - j0 - jn, jn, k, alpha, A + j0 + jn, lda, A + j0, lda, C + (j0 + jn) + (long)j0 * ldc,
                  ldc);
  }
}

/* Right-looking column Cholesky (LAPACK dpotf2 semantics: stops at the first
 * pivot that is not > 0 or is NaN, returns its 1-based index; the strict
 * upper triangle is not referenced). */
int oracle_dpotrf_tile(int n, double *A, int lda) {
  for (int j = 0; j < n; j++) {
    double ajj = A[j + (long)j * lda];
    if (!(ajj > 0.0) || isnan(ajj)) return j + 1;
    ajj = sqrt(ajj);
    A[j + (long)j * lda] = ajj;
    double r = 1.0 / ajj;
    double *cj = A + (long)j * lda;
    for (int i = j + 1; i < n; i++) cj[i] *= r;
    for (int k = j + 1; k < n; k++) {
      double f = cj[k];
      double *ck = A + (long)k * lda;
      for (int i = k; i < n; i++) ck[i] -= cj[i] * f;
    }
  }
  return 0;
}

/* Reference-BLAS order for side=R, uplo=L, trans=T: column k is scaled by
 * 1/L(k,k), then removed from the columns to its right. */
void oracle_dtrsm_rlt(int m, int n, const double *L, int ldl, double *B, int ldb) {
  for (int k = 0; k < n; k++) {
    double t = 1.0 / L[k + (long)k * ldl];
    double *bk = B + (long)k * ldb;
    for (int i = 0; i < m; i++) bk[i] *= t;
    for (int j = k + 1; j < n; j++) {
      double f = L[j + (long)k * ldl];
      if (f != 0.0) {
        double *bj = B + (long)j * ldb;
        for (int i = 0; i < m; i++) bj[i] -= f * bk[i];
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* task graph                                                          */
/* ------------------------------------------------------------------ */

int oracle_priority(int kind, int nt, int k0, int k1, int k2) {
  long long NT = nt, a = k0, b = k1, c = k2;
  switch (kind) {
    case ORACLE_POTRF: return (int)((NT - a) * (NT - a) * (NT - a));
    case ORACLE_TRSM: return (int)((NT - a) * (NT - a) * (NT - a) + 3 * ((2 * NT) - b - a - 1) * (a - b));
    case ORACLE_SYRK: return (int)((NT - a) * (NT - a) * (NT - a) + 3 * (a - b));
    case ORACLE_GEMM:
      return (int)((NT - a) * (NT - a) * (NT - a) + 3 * ((2 * NT) - a - b - 3) * (a - b) + 6 * (a - c));
  }
  return 0;
}

static int push_key(int *out, int cnt, int max, int kind, int a, int b, int c) {
  if (cnt < max) {
    out[4 * cnt + 0] = kind;
    out[4 * cnt + 1] = a;
    out[4 * cnt + 2] = b;
    out[4 * cnt + 3] = c;
  }
  return cnt + 1;
}

int oracle_successors(int kind, int nt, int k0, int k1, int k2, int *out, int max) {
  int cnt = 0;
  (void)k2;
  if (kind == ORACLE_POTRF) { /* potrf.h:112-116,180 broadcast<0,1> */
    int K = k0;
    cnt = push_key(out, cnt, max, -1, K, K, 0);
    for (int m = K + 1; m < nt; m++) cnt = push_key(out, cnt, max, ORACLE_TRSM, m, K, 0);
  } else if (kind == ORACLE_TRSM) { /* potrf.h:259-268,320 broadcast<0,1,2,3> */
    int M = k0, K = k1;
    cnt = push_key(out, cnt, max, -1, M, K, 0);
    cnt = push_key(out, cnt, max, ORACLE_SYRK, K, M, 0);
    for (int n = K + 1; n < M; n++) cnt = push_key(out, cnt, max, ORACLE_GEMM, M, n, K);
    for (int m = M + 1; m < nt; m++) cnt = push_key(out, cnt, max, ORACLE_GEMM, m, M, K);
  } else if (kind == ORACLE_SYRK) { /* potrf.h:449-457 */
    int K = k0, M = k1;
    if (M == K + 1)
      cnt = push_key(out, cnt, max, ORACLE_POTRF, K + 1, 0, 0);
    else
      cnt = push_key(out, cnt, max, ORACLE_SYRK, K + 1, M, 0);
  } else if (kind == ORACLE_GEMM) { /* potrf.h:578-586 */
    int M = k0, N = k1, K = k2;
    if (N == K + 1)
      cnt = push_key(out, cnt, max, ORACLE_TRSM, M, N, 0);
    else
      cnt = push_key(out, cnt, max, ORACLE_GEMM, M, N, K + 1);
  }
  return cnt;
}

typedef struct {
  int kind, a, b, c, prio;
} otask;

typedef struct {
  int N, NB, nt;
  double *tiles;
  /* pending input counts */
  int *dep_potrf, *dep_trsm, *dep_syrk, *dep_gemm;
  otask *heap;
  long heap_n, heap_cap;
  long remaining;
  int info;
  int abort;
  int idle;
  int nthreads;
  long counts[4];
  pthread_mutex_t mu;
  pthread_cond_t cv;
} osched;

static void heap_push(osched *s, otask t) {
  if (s->heap_n == s->heap_cap) {
    s->heap_cap = s->heap_cap ? 2 * s->heap_cap : 1024;
    s->heap = (otask *)realloc(s->heap, sizeof(otask) * s->heap_cap);
  }
  long i = s->heap_n++;
  s->heap[i] = t;
  while (i > 0) {
    long p = (i - 1) / 2;
    if (s->heap[p].prio >= s->heap[i].prio) break;
    otask tmp = s->heap[p];
    s->heap[p] = s->heap[i];
    s->heap[i] = tmp;
    i = p;
  }
}

static otask heap_pop(osched *s) {
  otask top = s->heap[0];
  s->heap[0] = s->heap[--s->heap_n];
  long i = 0;
  for (;;) {
    long l = 2 * i + 1, r = l + 1, b = i;
    if (l < s->heap_n && s->heap[l].prio > s->heap[b].prio) b = l;
    if (r < s->heap_n && s->heap[r].prio > s->heap[b].prio) b = r;
    if (b == i) break;
    otask tmp = s->heap[b];
    s->heap[b] = s->heap[i];
    s->heap[i] = tmp;
    i = b;
  }
  return top;
}

static void make_ready(osched *s, int kind, int a, int b, int c) {
  otask t = {kind, a, b, c, oracle_priority(kind, s->nt, a, b, c)};
  heap_push(s, t);
}

/* decrement the pending count of a successor; push it when it reaches 0 */
static void deliver(osched *s, int kind, int a, int b, int c) {
  int nt = s->nt;
  int *d = NULL;
  switch (kind) {
    case ORACLE_POTRF: d = &s->dep_potrf[a]; break;
    case ORACLE_TRSM: d = &s->dep_trsm[a * nt + b]; break;
    case ORACLE_SYRK: d = &s->dep_syrk[a * nt + b]; break;
    case ORACLE_GEMM: d = &s->dep_gemm[((long)a * nt + b) * nt + c]; break;
    default: return;
  }
  if (--(*d) == 0) make_ready(s, kind, a, b, c);
}

static int run_task(osched *s, const otask *t) {
  int nt = s->nt, NB = s->NB, N = s->N;
  double *T = s->tiles;
  switch (t->kind) {
    case ORACLE_POTRF: {
      int K = t->a;
      int info = oracle_dpotrf_tile(tile_dim(N, NB, K), tile_ptr(T, K, K, nt, NB), NB);
      return info ? K * NB + info : 0;
    }
    case ORACLE_TRSM: {
      int M = t->a, K = t->b;
      oracle_dtrsm_rlt(tile_dim(N, NB, M), tile_dim(N, NB, K), tile_ptr(T, K, K, nt, NB), NB,
                       tile_ptr(T, M, K, nt, NB), NB);
      return 0;
    }
    case ORACLE_SYRK: {
      int K = t->a, M = t->b;
      oracle_dsyrk_ln(tile_dim(N, NB, M), tile_dim(N, NB, K), -1.0, tile_ptr(T, M, K, nt, NB), NB, 1.0,
                      tile_ptr(T, M, M, nt, NB), NB);
      return 0;
    }
    case ORACLE_GEMM: {
      int M = t->a, Nn = t->b, K = t->c;
      oracle_dgemm_nt(tile_dim(N, NB, M), tile_dim(N, NB, Nn), tile_dim(N, NB, K), -1.0,
                      tile_ptr(T, M, K, nt, NB), NB, tile_ptr(T, Nn, K, nt, NB), NB, 1.0,
                      tile_ptr(T, M, Nn, nt, NB), NB);
      return 0;
    }
  }
  return 0;
}

static void *worker(void *arg) {
  osched *s = (osched *)arg;
  int succ[4 * 1024];
  int *big = NULL;
  pthread_mutex_lock(&s->mu);
  for (;;) {
    while (s->heap_n == 0 && s->remaining > 0 && !s->abort) {
      s->idle++;
      if (s->idle == s->nthreads) { /* nothing ready and nobody running: stuck */
        s->abort = 1;
        pthread_cond_broadcast(&s->cv);
        s->idle--;
        break;
      }
      pthread_cond_wait(&s->cv, &s->mu);
      s->idle--;
    }
    if (s->remaining == 0 || s->abort) break;
    otask t = heap_pop(s);
    pthread_mutex_unlock(&s->mu);
    int info = run_task(s, &t);
    int cap = 1024, *out = succ;
    int n = oracle_successors(t.kind, s->nt, t.a, t.b, t.c, out, cap);
    if (n > cap) {
      big = (int *)realloc(big, sizeof(int) * 4 * n);
      out = big;
      oracle_successors(t.kind, s->nt, t.a, t.b, t.c, out, n);
    }
    pthread_mutex_lock(&s->mu);
    s->counts[t.kind]++;
    s->remaining--;
    if (info) {
      if (s->info == 0) s->info = info;
      s->abort = 1; /* potrf.h:183-186: a failed tile stops the DAG */
    } else {
      for (int i = 0; i < n; i++)
        if (out[4 * i] >= 0) deliver(s, out[4 * i], out[4 * i + 1], out[4 * i + 2], out[4 * i + 3]);
    }
    pthread_cond_broadcast(&s->cv);
  }
  pthread_cond_broadcast(&s->cv);
  pthread_mutex_unlock(&s->mu);
  free(big);
  return NULL;
}

static long g_last_counts[4];

static int host_cores(void) {
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof set, &set) == 0) return CPU_COUNT(&set);
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? (int)n : 1;
}

int oracle_dpotrf_tiled(int N, int NB, double *tiles, int nthreads) {
  osched s;
  memset(&s, 0, sizeof s);
  s.N = N;
  s.NB = NB;
  s.nt = (N + NB - 1) / NB;
  s.tiles = tiles;
  int nt = s.nt;
  if (nthreads <= 0) nthreads = host_cores();
  s.nthreads = nthreads;
  s.dep_potrf = (int *)calloc(nt, sizeof(int));
  s.dep_trsm = (int *)calloc((size_t)nt * nt, sizeof(int));
  s.dep_syrk = (int *)calloc((size_t)nt * nt, sizeof(int));
  s.dep_gemm = (int *)calloc((size_t)nt * nt * nt, sizeof(int));
  /* inputs per task (potrf.h:189,323,459,588); the dispatcher
   * (potrf.h:637-664) has delivered every initial tile already. */
  long total = 0;
  for (int K = 0; K < nt; K++) {
    s.dep_potrf[K] = K > 0 ? 1 : 0;
    total++;
    for (int M = K + 1; M < nt; M++) {
      s.dep_trsm[M * nt + K] = K > 0 ? 2 : 1;
      s.dep_syrk[K * nt + M] = K > 0 ? 2 : 1;
      total += 2;
      for (int Nn = K + 1; Nn < M; Nn++) {
        s.dep_gemm[((long)M * nt + Nn) * nt + K] = K > 0 ? 3 : 2;
        total++;
      }
    }
  }
  s.remaining = total;
  pthread_mutex_init(&s.mu, NULL);
  pthread_cond_init(&s.cv, NULL);
  make_ready(&s, ORACLE_POTRF, 0, 0, 0);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, worker, &s);
  for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
  memcpy(g_last_counts, s.counts, sizeof g_last_counts);
  pthread_mutex_destroy(&s.mu);
  pthread_cond_destroy(&s.cv);
  free(th);
  free(s.heap);
  free(s.dep_potrf);
  free(s.dep_trsm);
  free(s.dep_syrk);
  free(s.dep_gemm);
  return s.info;
}

void oracle_last_task_counts(long counts[4]) { memcpy(counts, g_last_counts, sizeof g_last_counts); }

int oracle_dpotrf_tiled_seq(int N, int NB, double *T) {
  int nt = (N + NB - 1) / NB;
  for (int K = 0; K < nt; K++) {
    int info = oracle_dpotrf_tile(tile_dim(N, NB, K), tile_ptr(T, K, K, nt, NB), NB);
    if (info) return K * NB + info;
    for (int M = K + 1; M < nt; M++)
      oracle_dtrsm_rlt(tile_dim(N, NB, M), tile_dim(N, NB, K), tile_ptr(T, K, K, nt, NB), NB,
                       tile_ptr(T, M, K, nt, NB), NB);
    for (int M = K + 1; M < nt; M++) {
      oracle_dsyrk_ln(tile_dim(N, NB, M), tile_dim(N, NB, K), -1.0, tile_ptr(T, M, K, nt, NB), NB, 1.0,
                      tile_ptr(T, M, M, nt, NB), NB);
      for (int Nn = K + 1; Nn < M; Nn++)
        oracle_dgemm_nt(tile_dim(N, NB, M), tile_dim(N, NB, Nn), tile_dim(N, NB, K), -1.0,
                        tile_ptr(T, M, K, nt, NB), NB, tile_ptr(T, Nn, K, nt, NB), NB, 1.0,
                        tile_ptr(T, M, Nn, nt, NB), NB);
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* checks                                                              */
/* ------------------------------------------------------------------ */

/* testing_dpotrf.cc:277-335 */
double oracle_check_dpotrf(const double *A, const double *A0, int N) {
  double *LLt = (double *)malloc(sizeof(double) * (size_t)N * N);
  /* LLt = lower(A) * lower(A)^T  (the trmm of :298-299) */
  for (long j = 0; j < N; j++)
    for (long i = 0; i < N; i++) {
      long kmax = i < j ? i : j;
      double s = 0.0;
      for (long k = 0; k <= kmax; k++) s += A[i + k * N] * A[j + k * N];
      LLt[i + j * N] = s;
    }
  /* :302-310: subtract A0 on the lower part and mirror */
  for (long i = 0; i < N; i++)
    for (long j = 0; j < N; j++) {
      if (j <= i)
        LLt[i + j * N] -= A0[i + j * N];
      else
        LLt[i + j * N] = LLt[j + i * N] - A0[j + i * N];
    }
  /* :312-313: lange(Inf) = max row sum.  A0 has a zero upper part. */
  double Anorm = 0.0, Rnorm = 0.0;
  for (long i = 0; i < N; i++) {
    double sa = 0.0, sr = 0.0;
    for (long j = 0; j < N; j++) {
      sa += fabs(A0[i + j * N]);
      sr += fabs(LLt[i + j * N]);
    }
    if (sa > Anorm || isnan(sa)) Anorm = sa;
    if (sr > Rnorm || isnan(sr)) Rnorm = sr;
  }
  free(LLt);
  return Rnorm / (Anorm * N * DBL_EPSILON);
}

typedef struct {
  int N, NB, nt, nthreads, tid;
  const double *L, *A0;
  double *rows_r, *rows_a; /* per-thread row sums (N each) */
} chk_arg;

static void *chk_worker(void *p) {
  chk_arg *c = (chk_arg *)p;
  int N = c->N, NB = c->NB, nt = c->nt;
  double *R = (double *)malloc(sizeof(double) * NB * NB);
  double *D = (double *)malloc(sizeof(double) * NB * NB);
  long idx = 0;
  for (int n = 0; n < nt; n++)
    for (int m = n; m < nt; m++, idx++) {
      if (idx % c->nthreads != c->tid) continue;
      int mb = tile_dim(N, NB, m), nb = tile_dim(N, NB, n);
      memset(R, 0, sizeof(double) * NB * NB);
      /* R = sum_{k<=n} L_mk L_nk^T - A0_mn, where the diagonal factor tile
       * L_nn counts with its strict upper part as zero (the trmm of
       * testing_dpotrf.cc:298 only references the lower triangle). */
      for (int k = 0; k < n; k++) {
        int kb = tile_dim(N, NB, k);
        gemm_nt_acc(mb, nb, kb, 1.0, tile_cptr(c->L, m, k, nt, NB), NB, tile_cptr(c->L, n, k, nt, NB), NB,
                    R, NB);
      }
      {
        const double *d = tile_cptr(c->L, n, n, nt, NB);
        for (int j = 0; j < nb; j++)
          for (int i = 0; i < nb; i++) D[i + j * NB] = (i >= j) ? d[i + (long)j * NB] : 0.0;
        if (m == n) {
          gemm_nt_acc(mb, nb, nb, 1.0, D, NB, D, NB, R, NB);
        } else {
          gemm_nt_acc(mb, nb, nb, 1.0, tile_cptr(c->L, m, n, nt, NB), NB, D, NB, R, NB);
        }
      }
      const double *a0 = tile_cptr(c->A0, m, n, nt, NB);
      for (int j = 0; j < nb; j++)
        for (int i = 0; i < mb; i++) {
          long gi = (long)m * NB + i, gj = (long)n * NB + j;
          /* ||A0||_inf is lange(Inf) of getLAPACKMatrix(A0) (pmw.h:122-139,
           * testing_dpotrf.cc:312): the lower tiles as stored, diagonal tiles
           * whole (their upper part included), no mirror. */
          c->rows_a[gi] += fabs(a0[i + (long)j * NB]);
          if (m == n && j > i) continue;
          /* the residual is mirrored (testing_dpotrf.cc:302-310) */
          double r = fabs(R[i + j * NB] - a0[i + (long)j * NB]);
          c->rows_r[gi] += r;
          if (gi != gj) c->rows_r[gj] += r; /* symmetric counterpart */
        }
    }
  free(R);
  free(D);
  return NULL;
}

double oracle_check_dpotrf_tiled(int N, int NB, const double *Ltiles, const double *A0tiles, int nthreads) {
  int nt = (N + NB - 1) / NB;
  if (nthreads <= 0) nthreads = host_cores();
  chk_arg *args = (chk_arg *)calloc(nthreads, sizeof(chk_arg));
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  for (int t = 0; t < nthreads; t++) {
    args[t] = (chk_arg){N, NB, nt, nthreads, t, Ltiles, A0tiles, NULL, NULL};
    args[t].rows_r = (double *)calloc(N, sizeof(double));
    args[t].rows_a = (double *)calloc(N, sizeof(double));
    pthread_create(&th[t], NULL, chk_worker, &args[t]);
  }
  double Rn = 0.0, An = 0.0;
  double *rr = (double *)calloc(N, sizeof(double)), *ra = (double *)calloc(N, sizeof(double));
  for (int t = 0; t < nthreads; t++) {
    pthread_join(th[t], NULL);
    for (int i = 0; i < N; i++) {
      rr[i] += args[t].rows_r[i];
      ra[i] += args[t].rows_a[i];
    }
    free(args[t].rows_r);
    free(args[t].rows_a);
  }
  for (int i = 0; i < N; i++) {
    if (rr[i] > Rn || isnan(rr[i])) Rn = rr[i];
    if (ra[i] > An) An = ra[i];
  }
  free(rr);
  free(ra);
  free(args);
  free(th);
  return Rn / (An * N * DBL_EPSILON);
}

/* ------------------------------------------------------------------ */
/* solve and residual                                                  */
/* ------------------------------------------------------------------ */

void oracle_dpotrs_tiled(int N, int NB, const double *L, int nrhs, double *B, int ldb) {
  int nt = (N + NB - 1) / NB;
  for (int r = 0; r < nrhs; r++) {
    double *b = B + (long)r * ldb;
    /* forward: L y = b */
    for (int i = 0; i < nt; i++) {
      int ib = tile_dim(N, NB, i);
      double *bi = b + (long)i * NB;
      for (int j = 0; j < i; j++) {
        const double *t = tile_cptr(L, i, j, nt, NB);
        const double *bj = b + (long)j * NB;
        int jb = tile_dim(N, NB, j);
        for (int c = 0; c < jb; c++)
          for (int q = 0; q < ib; q++) bi[q] -= t[q + (long)c * NB] * bj[c];
      }
      const double *d = tile_cptr(L, i, i, nt, NB);
      for (int c = 0; c < ib; c++) {
        bi[c] /= d[c + (long)c * NB];
        for (int q = c + 1; q < ib; q++) bi[q] -= d[q + (long)c * NB] * bi[c];
      }
    }
    /* backward: L^T x = y */
    for (int i = nt - 1; i >= 0; i--) {
      int ib = tile_dim(N, NB, i);
      double *bi = b + (long)i * NB;
      for (int j = i + 1; j < nt; j++) {
        const double *t = tile_cptr(L, j, i, nt, NB);
        const double *bj = b + (long)j * NB;
        int jb = tile_dim(N, NB, j);
        for (int c = 0; c < ib; c++) {
          double s = 0.0;
          for (int q = 0; q < jb; q++) s += t[q + (long)c * NB] * bj[q];
          bi[c] -= s;
        }
      }
      const double *d = tile_cptr(L, i, i, nt, NB);
      for (int c = ib - 1; c >= 0; c--) {
        double s = bi[c];
        for (int q = c + 1; q < ib; q++) s -= d[q + (long)c * NB] * bi[q];
        bi[c] = s / d[c + (long)c * NB];
      }
    }
  }
}

void oracle_dresidual_tiled(int N, int NB, const double *A, int nrhs, const double *X, int ldx,
                            const double *B, int ldb, double *R, int ldr) {
  int nt = (N + NB - 1) / NB;
  for (int r = 0; r < nrhs; r++) {
    const double *x = X + (long)r * ldx;
    double *rr = R + (long)r * ldr;
    for (int i = 0; i < N; i++) rr[i] = B[i + (long)r * ldb];
    for (int n = 0; n < nt; n++)
      for (int m = n; m < nt; m++) {
        const double *t = tile_cptr(A, m, n, nt, NB);
        int mb = tile_dim(N, NB, m), nb = tile_dim(N, NB, n);
        const double *xm = x + (long)m * NB, *xn = x + (long)n * NB;
        double *rm = rr + (long)m * NB, *rn = rr + (long)n * NB;
        for (int j = 0; j < nb; j++)
          for (int i = 0; i < mb; i++) {
            if (m == n && i < j) continue;
            double a = t[i + (long)j * NB];
            rm[i] -= a * xn[j];
            if (m != n || i != j) rn[j] -= a * xm[i];
          }
      }
  }
}

double oracle_norm_inf_sym_tiled(int N, int NB, const double *A) {
  int nt = (N + NB - 1) / NB;
  double *rs = (double *)calloc(N, sizeof(double));
  for (int n = 0; n < nt; n++)
    for (int m = n; m < nt; m++) {
      const double *t = tile_cptr(A, m, n, nt, NB);
      int mb = tile_dim(N, NB, m), nb = tile_dim(N, NB, n);
      for (int j = 0; j < nb; j++)
        for (int i = 0; i < mb; i++) {
          if (m == n && i < j) continue;
          double a = fabs(t[i + (long)j * NB]);
          long gi = (long)m * NB + i, gj = (long)n * NB + j;
          rs[gi] += a;
          if (gi != gj) rs[gj] += a;
        }
    }
  double mx = 0.0;
  for (int i = 0; i < N; i++)
    if (rs[i] > mx) mx = rs[i];
  free(rs);
  return mx;
}

double oracle_solve_backward_error(int N, int NB, const double *A, int nrhs, const double *X, int ldx,
                                   const double *B, int ldb) {
  double *R = (double *)malloc(sizeof(double) * (size_t)N * nrhs);
  oracle_dresidual_tiled(N, NB, A, nrhs, X, ldx, B, ldb, R, N);
  double an = oracle_norm_inf_sym_tiled(N, NB, A);
  double worst = 0.0;
  for (int r = 0; r < nrhs; r++) {
    double rn = 0.0, xn = 0.0, bn = 0.0;
    for (int i = 0; i < N; i++) {
      double v = fabs(R[i + (long)r * N]);
      if (v > rn || isnan(v)) rn = v;
      v = fabs(X[i + (long)r * ldx]);
      if (v > xn) xn = v;
      v = fabs(B[i + (long)r * ldb]);
      if (v > bn) bn = v;
    }
    double e = rn / (an * xn + bn);
    if (e > worst || isnan(e)) worst = e;
  }
  free(R);
  return worst;
}

/* ------------------------------------------------------------------ */
/* POINV family: TRTRI (lower, non-unit), LAUUM (lower), POTRI         */
/* ------------------------------------------------------------------ */

/* B := alpha B L^{-1}  (Right, Lower, NoTrans, NonUnit): X L = alpha B is
 * solved column by column from the right. */
void oracle_dtrsm_rln(int m, int n, double alpha, const double *L, int ldl, double *B, int ldb) {
  for (int k = n - 1; k >= 0; k--) {
    double *bk = B + (long)k * ldb;
    for (int i = 0; i < m; i++) bk[i] *= alpha;
    for (int j = k + 1; j < n; j++) {
      double f = L[j + (long)k * ldl];
      if (f != 0.0) {
        const double *bj = B + (long)j * ldb;
        for (int i = 0; i < m; i++) bk[i] -= f * bj[i];
      }
    }
    double t = 1.0 / L[k + (long)k * ldl];
    for (int i = 0; i < m; i++) bk[i] *= t;
  }
}

/* B := L^{-1} B  (Left, Lower, NoTrans, NonUnit): forward substitution. */
void oracle_dtrsm_lln(int m, int n, const double *L, int ldl, double *B, int ldb) {
  for (int j = 0; j < n; j++) {
    double *b = B + (long)j * ldb;
    for (int i = 0; i < m; i++) {
      double s = b[i];
      for (int l = 0; l < i; l++) s -= L[i + (long)l * ldl] * b[l];
      b[i] = s / L[i + (long)i * ldl];
    }
  }
}

/* B := L^T B  (Left, Lower, Trans, NonUnit); row i of the result only reads
 * rows >= i of B, so rows are overwritten top down. */
void oracle_dtrmm_llt(int m, int n, const double *L, int ldl, double *B, int ldb) {
  for (int j = 0; j < n; j++) {
    double *b = B + (long)j * ldb;
    for (int i = 0; i < m; i++) {
      double s = 0.0;
      for (int k = i; k < m; k++) s += L[k + (long)i * ldl] * b[k];
      b[i] = s;
    }
  }
}

/* C := alpha A^T A + beta C, lower triangle of the n x n C; A is k x n. */
void oracle_dsyrk_lt(int n, int k, double alpha, const double *A, int lda, double beta, double *C, int ldc) {
  for (int j = 0; j < n; j++)
    for (int i = j; i < n; i++) {
      double s = 0.0;
      for (int l = 0; l < k; l++) s += A[l + (long)i * lda] * A[l + (long)j * lda];
      double c = (beta == 0.0) ? 0.0 : beta * C[i + (long)j * ldc];
      C[i + (long)j * ldc] = c + alpha * s;
    }
}

/* C := alpha op(A) B + beta C with op(A) = A (ta = 0, A m x k) or A^T (ta = 1,
 * A k x m); B is k x n. */
void oracle_dgemm_xn(int ta, int m, int n, int k, double alpha, const double *A, int lda, const double *B, int ldb,
                     double beta, double *C, int ldc) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      double s = 0.0;
      for (int l = 0; l < k; l++)
        s += (ta ? A[l + (long)i * lda] : A[i + (long)l * lda]) * B[l + (long)j * ldb];
      double c = (beta == 0.0) ? 0.0 : beta * C[i + (long)j * ldc];
      C[i + (long)j * ldc] = c + alpha * s;
    }
}

/* In-place inverse of a lower triangular, non-unit n x n tile (LAPACK dtrti2,
 * lower: columns from the right, each multiplied by the already inverted
 * trailing block).  Returns 0, or the 1-based index of the first zero on the
 * diagonal, in which case A is left untouched (LAPACK dtrtri's check). */
int oracle_dtrtri_tile(int n, double *A, int lda) {
  for (int j = 0; j < n; j++)
    if (A[j + (long)j * lda] == 0.0) return j + 1;
  for (int j = n - 1; j >= 0; j--) {
    A[j + (long)j * lda] = 1.0 / A[j + (long)j * lda];
    double ajj = -A[j + (long)j * lda];
    /* x := T x, T = inverted block (j+1.., j+1..), x = column j below the diagonal */
    for (int i = n - 1; i > j; i--) {
      double s = 0.0;
      for (int l = j + 1; l <= i; l++) s += A[i + (long)l * lda] * A[l + (long)j * lda];
      A[i + (long)j * lda] = s;
    }
    for (int i = j + 1; i < n; i++) A[i + (long)j * lda] *= ajj;
  }
  return 0;
}

/* Lower part of L^T L in place (LAPACK dlauum, lower).  Row i of the result
 * reads rows >= i, so each row is formed in a work row, then stored. */
void oracle_dlauum_tile(int n, double *A, int lda) {
  double *row = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) {
    for (int j = 0; j <= i; j++) {
      double s = 0.0;
      for (int k = i; k < n; k++) s += A[k + (long)i * lda] * A[k + (long)j * lda];
      row[j] = s;
    }
    for (int j = 0; j <= i; j++) A[i + (long)j * lda] = row[j];
  }
  free(row);
}

/* L := L^{-1} over the lower tiles, in the task order of
 * examples/potrf/trtri_L.h: at step K, TRSMR(K, M) A(M,K) := -A(M,K) L_KK^{-1}
 * (:96-104), GEMM(K, M, N) A(M,N) += A(M,K) A(K,N) for N < K (:178-185),
 * TRSML(K, N) A(K,N) := L_KK^{-1} A(K,N) (:49-56), TRTRI(K) (:27).  The
 * GEMMs read A(K,N) before its TRSML and both TRSMs read the original
 * L_KK, as in the graph.  Returns 0 or the global 1-based index of the first
 * zero diagonal entry (nothing is changed then). */
int oracle_dtrtri_tiled(int N, int NB, double *T) {
  int nt = (N + NB - 1) / NB;
  for (int K = 0; K < nt; K++) {
    const double *d = tile_ptr(T, K, K, nt, NB);
    for (int j = 0; j < tile_dim(N, NB, K); j++)
      if (d[j + (long)j * NB] == 0.0) return K * NB + j + 1;
  }
  for (int K = 0; K < nt; K++) {
    int dk = tile_dim(N, NB, K);
    double *Lkk = tile_ptr(T, K, K, nt, NB);
    for (int M = K + 1; M < nt; M++)
      oracle_dtrsm_rln(tile_dim(N, NB, M), dk, -1.0, Lkk, NB, tile_ptr(T, M, K, nt, NB), NB);
    for (int M = K + 1; M < nt; M++)
      for (int Nn = 0; Nn < K; Nn++)
        oracle_dgemm_xn(0, tile_dim(N, NB, M), tile_dim(N, NB, Nn), dk, 1.0, tile_ptr(T, M, K, nt, NB), NB,
                        tile_ptr(T, K, Nn, nt, NB), NB, 1.0, tile_ptr(T, M, Nn, nt, NB), NB);
    for (int Nn = 0; Nn < K; Nn++)
      oracle_dtrsm_lln(dk, tile_dim(N, NB, Nn), Lkk, NB, tile_ptr(T, K, Nn, nt, NB), NB);
    oracle_dtrtri_tile(dk, Lkk, NB);
  }
  return 0;
}

/* A := L^T L (lower tiles), in the task order of examples/potrf/lauum.h:
 * tile (M, N), N < M, is TRMM(M, N) A(M,N) := L_MM^T A(M,N) (:72-80) followed
 * by GEMM(K, N, M) A(M,N) += L_KM^T L_KN for K = M+1, M+2, ... (:172-180);
 * tile (M, M) is LAUUM(M) (:37) followed by SYRK(K, M) A(M,M) += L_KM^T L_KM
 * (:123-128).  Every task reads the original L (the dispatcher's tiles are
 * read-only), so a tile row is formed in work tiles and stored after the
 * whole row is done; later rows never read it. */
void oracle_dlauum_tiled(int N, int NB, double *T) {
  int nt = (N + NB - 1) / NB;
  size_t per = (size_t)NB * NB;
  double *row = (double *)malloc(sizeof(double) * per * (size_t)(nt > 0 ? nt : 1));
  for (int M = 0; M < nt; M++) {
    int dm = tile_dim(N, NB, M);
    const double *Lmm = tile_ptr(T, M, M, nt, NB);
    for (int Nn = 0; Nn <= M; Nn++) {
      int dn = tile_dim(N, NB, Nn);
      double *X = row + per * Nn;
      const double *Lmn = tile_ptr(T, M, Nn, nt, NB);
      for (int j = 0; j < dn; j++)
        for (int i = 0; i < dm; i++) X[i + (long)j * NB] = Lmn[i + (long)j * NB];
      if (Nn < M) {
        oracle_dtrmm_llt(dm, dn, Lmm, NB, X, NB);
        for (int K = M + 1; K < nt; K++)
          oracle_dgemm_xn(1, dm, dn, tile_dim(N, NB, K), 1.0, tile_ptr(T, K, M, nt, NB), NB,
                          tile_ptr(T, K, Nn, nt, NB), NB, 1.0, X, NB);
      } else {
        oracle_dlauum_tile(dm, X, NB);
        for (int K = M + 1; K < nt; K++)
          oracle_dsyrk_lt(dm, tile_dim(N, NB, K), 1.0, tile_ptr(T, K, M, nt, NB), NB, 1.0, X, NB);
      }
    }
    for (int Nn = 0; Nn <= M; Nn++) {
      int dn = tile_dim(N, NB, Nn);
      double *D = tile_ptr(T, M, Nn, nt, NB);
      const double *X = row + per * Nn;
      for (int j = 0; j < dn; j++)
        for (int i = (Nn == M ? j : 0); i < dm; i++) D[i + (long)j * NB] = X[i + (long)j * NB];
    }
  }
  free(row);
}

/* POTRI (examples/potrf/potri.h:17-30): TRTRI then LAUUM of the factor. */
int oracle_dpotri_tiled(int N, int NB, double *T) {
  int info = oracle_dtrtri_tiled(N, NB, T);
  if (info) return info;
  oracle_dlauum_tiled(N, NB, T);
  return 0;
}

/* check_dtrtri of examples/potrf/testing_dtrtri.cc:269-325 (lower, non-unit):
 * ||I - L Linv||_1 / (||L||_1 ||Linv||_1 N eps), one-norms of the lower
 * triangles; dense col-major N x N, lda = N, only lower parts read. */
double oracle_check_dtrtri(const double *L, const double *Linv, int N) {
  double *P = (double *)calloc((size_t)N * N, sizeof(double));
  for (int j = 0; j < N; j++)
    for (int i = j; i < N; i++) {
      double s = 0.0;
      for (int k = j; k <= i; k++) s += L[i + (long)k * N] * Linv[k + (long)j * N];
      P[i + (long)j * N] = ((i == j) ? 1.0 : 0.0) - s;
    }
  double an = 0.0, bn = 0.0, rn = 0.0;
  for (int j = 0; j < N; j++) {
    double sa = 0.0, sb = 0.0, sr = 0.0;
    for (int i = j; i < N; i++) {
      sa += fabs(L[i + (long)j * N]);
      sb += fabs(Linv[i + (long)j * N]);
      sr += fabs(P[i + (long)j * N]);
    }
    if (sa > an) an = sa;
    if (sb > bn) bn = sb;
    if (sr > rn) rn = sr;
  }
  free(P);
  double eps = 2.220446049250313e-16;
  return (rn / an / bn) / (eps * N);
}

/* FLOPS_DTRTRI / FLOPS_DLAUUM / FLOPS_DPOTRI of lauum.h:11-25 (DPLASMA). */
double oracle_flops_dtrtri(double n) {
  return n * (n * (1. / 6. * n + 0.5) + 1. / 3.) + n * (n * (1. / 6. * n - 0.5) + 1. / 3.);
}
double oracle_flops_dpotri(double n) {
  return n * ((2. / 3.) + n * ((1. / 3.) * n + 1.)) + n * ((1. / 6.) + n * ((1. / 3.) * n - 0.5));
}
double oracle_flops_dlauum(double n) { return oracle_flops_dpotri(n) - oracle_flops_dtrtri(n); }
