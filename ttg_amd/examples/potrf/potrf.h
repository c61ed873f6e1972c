// The tiled Cholesky task graph (examples/potrf/potrf.h:28-766 of
// devreal/ttg) on the MI355X runtime: the same template tasks, keys,
// successor keys, dispatch and DPLASMA priorities; the device bodies call
// libttg_chol's gfx950 kernels (include/ttg_chol.h) on
// ttg::device::current_stream() instead of hipSOLVER/hipBLAS, for double
// or float tiles as the reference's templates do.  The runtime's
// launch grouping turns all the tile calls of a scheduling round into a few
// batched launches.
// Built with DEBUG_TILES_VALUES, every task also takes the device 2-norm of
// its input tiles, checks each against the norm the tile's producer set, and
// sets the norm of its output tile (potrf.h:135-173, 271-318, 403-447,
// 529-576): a task that received a tile other than the one its producer sent
// aborts the run.
#pragma once

#include <array>
#include <cassert>
#include <iostream>
#include <type_traits>
#include <vector>

#include "pmw.h"
#include "ttg.h"
#include "ttg_chol.h"

namespace potrf {

// DPLASMA flop count (potrf.h:30-33): N^3/3 + N^2/2 + N/6
inline double FMULS_POTRF(double n) { return n * (((1. / 6.) * n + 0.5) * n + (1. / 3.)); }
inline double FADDS_POTRF(double n) { return n * (((1. / 6.) * n) * n - (1. / 6.)); }
inline double FLOPS_DPOTRF(double n) { return FMULS_POTRF(n) + FADDS_POTRF(n); }

inline void check_launch(int rc) {
  if (rc != TTG_CHOL_OK) ttg::detail::fatal(std::string("tile kernel: ") + ttg_chol_last_error());
}

// The tile kernels for the element type: the D routines for double tiles,
// the S routines for float tiles (the reference's bodies are templated on
// MatrixT::element_type and reach hipsolverSpotrf/hipblasS* for float).
inline int tile_potrf(void *s, int n, double *A, int lda, int *info, int off) {
  return ttg_chol_dpotrf_tile(s, n, A, lda, info, off);
}
inline int tile_potrf(void *s, int n, float *A, int lda, int *info, int off) {
  return ttg_chol_spotrf_tile(s, n, A, lda, info, off);
}
inline int tile_trsm(void *s, int m, int n, const double *L, int ldl, double *B, int ldb) {
  return ttg_chol_dtrsm_rlt(s, m, n, L, ldl, B, ldb);
}
inline int tile_trsm(void *s, int m, int n, const float *L, int ldl, float *B, int ldb) {
  return ttg_chol_strsm_rlt(s, m, n, L, ldl, B, ldb);
}
inline int tile_syrk(void *s, int n, int k, double alpha, const double *A, int lda, double beta, double *C, int ldc) {
  return ttg_chol_dsyrk_ln(s, n, k, alpha, A, lda, beta, C, ldc);
}
inline int tile_syrk(void *s, int n, int k, double alpha, const float *A, int lda, double beta, float *C, int ldc) {
  return ttg_chol_ssyrk_ln(s, n, k, (float)alpha, A, lda, (float)beta, C, ldc);
}
inline int tile_gemm(void *s, int m, int n, int k, double alpha, const double *A, int lda, const double *B, int ldb,
                     double beta, double *C, int ldc) {
  return ttg_chol_dgemm_nt(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
inline int tile_gemm(void *s, int m, int n, int k, double alpha, const float *A, int lda, const float *B, int ldb,
                     double beta, float *C, int ldc) {
  return ttg_chol_sgemm_nt(s, m, n, k, (float)alpha, A, lda, B, ldb, (float)beta, C, ldc);
}

#ifdef DEBUG_TILES_VALUES
// device 2-norm of a whole tile (lda x cols values) into *d_norm, on the
// task's stream (potrf.h:76-87)
template <typename T>
void device_norm(const MatrixTile<T> &t, double *d_norm) {
  static_assert(std::is_same_v<T, double>, "tile norms are checked for double tiles only");
  check_launch(ttg_chol_dnrm2(ttg::device::current_stream(), (int)t.size(), t.buffer().current_device_ptr(), d_norm));
}
#endif

// POTRF(K): factor tile (K,K); broadcast it to the result and to TRSM(m,K),
// m > K (potrf.h:90-222).
template <typename MatrixT>
auto make_potrf(MatrixT &A, ttg::Edge<Key1, MatrixTile<typename MatrixT::element_type>> &input_disp,
                ttg::Edge<Key1, MatrixTile<typename MatrixT::element_type>> &input,
                ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &output_trsm,
                ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &output_result) {
  using T = typename MatrixT::element_type;
  auto f = [&A](const Key1 &key, MatrixTile<T> &&tile_kk,
                std::tuple<ttg::Out<Key2, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int K = key[0];
    std::vector<Key2> keylist;
    keylist.reserve(A.rows() - K);
    for (int m = K + 1; m < (int)A.rows(); ++m) keylist.push_back(Key2(m, K));
    int host_info = -32;
    auto dev_info = ttg::make_scratch(&host_info, ttg::scope::Allocate);
#ifdef DEBUG_TILES_VALUES
    const double expect_kk = tile_kk.norm();
    std::array<double, 2> norms{};  // input and output of tile_kk
    auto norms_s = ttg::make_scratch(norms.data(), ttg::scope::Allocate, norms.size());
    co_await ttg::device::select(tile_kk.buffer(), dev_info, norms_s);
    device_norm(tile_kk, norms_s.device_ptr() + 0);
#else
    co_await ttg::device::select(tile_kk.buffer(), dev_info);
#endif
    check_launch(tile_potrf(ttg::device::current_stream(), (int)tile_kk.cols(), tile_kk.buffer().current_device_ptr(),
                            (int)tile_kk.lda(), dev_info.device_ptr(), 0));
#ifdef DEBUG_TILES_VALUES
    device_norm(tile_kk, norms_s.device_ptr() + 1);
    co_await ttg::device::wait(dev_info, norms_s);
    expect_norm("POTRF", key, "tile_kk", expect_kk, norms[0]);
    tile_kk.set_norm(norms[1]);
#else
    co_await ttg::device::wait(dev_info);
#endif
    if (host_info != 0) {
      std::cerr << "Factorization of tile " << K << " failed: " << host_info << std::endl;
      // the reference ends the process here (potrf.h:183-187); this graph
      // stops cleanly with LAPACK's info, the 1-based column of the matrix
      ttg::cancel((long)K * (long)A.rows_in_tile() + host_info);
      co_return;
    }
    co_await ttg::device::forward(
        ttg::device::broadcast<0, 1>(std::make_tuple(Key2(K, K), std::move(keylist)), std::move(tile_kk), out));
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(ttg::fuse(input, input_disp)),
                                                ttg::edges(output_result, output_trsm), "POTRF",
                                                {"tile_kk/dispatcher"}, {"output_result", "output_trsm"});
}

// TRSM(M,K): tile (M,K) := tile (M,K) L_KK^{-T}; broadcast to the result,
// SYRK(K,M), GEMM(M,n,K) for K<n<M and GEMM(m,M,K) for m>M (potrf.h:224-378).
template <typename MatrixT>
auto make_trsm(MatrixT &A, ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input_disp,
               ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input_kk,
               ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input_mk,
               ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &output_diag,
               ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &output_row,
               ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &output_col,
               ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &output_result) {
  using T = typename MatrixT::element_type;
  auto f = [&A](const Key2 &key, const MatrixTile<T> &tile_kk, MatrixTile<T> &&tile_mk,
                std::tuple<ttg::Out<Key2, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>, ttg::Out<Key3, MatrixTile<T>>,
                           ttg::Out<Key3, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int M = key[0], K = key[1];
    assert(tile_mk.cols() == tile_kk.rows());
    std::vector<Key3> keylist_row, keylist_col;
    keylist_row.reserve(M - K);
    keylist_col.reserve(A.rows() - M - 1);
    for (int n = K + 1; n < M; ++n) keylist_row.push_back(Key3(M, n, K));
    for (int m = M + 1; m < (int)A.rows(); ++m) keylist_col.push_back(Key3(m, M, K));
#ifdef DEBUG_TILES_VALUES
    const double expect_kk = tile_kk.norm(), expect_mk = tile_mk.norm();
    std::array<double, 3> norms{};  // inputs tile_kk, tile_mk; output tile_mk
    auto norms_s = ttg::make_scratch(norms.data(), ttg::scope::Allocate, norms.size());
    co_await ttg::device::select(tile_kk.buffer(), tile_mk.buffer(), norms_s);
    device_norm(tile_kk, norms_s.device_ptr() + 0);
    device_norm(tile_mk, norms_s.device_ptr() + 1);
#else
    co_await ttg::device::select(tile_kk.buffer(), tile_mk.buffer());
#endif
    check_launch(tile_trsm(ttg::device::current_stream(), (int)tile_mk.rows(), (int)tile_mk.cols(),
                           tile_kk.buffer().current_device_ptr(), (int)tile_kk.lda(),
                           tile_mk.buffer().current_device_ptr(), (int)tile_mk.lda()));
#ifdef DEBUG_TILES_VALUES
    device_norm(tile_mk, norms_s.device_ptr() + 2);
    co_await ttg::device::wait(norms_s);
    expect_norm("TRSM", key, "tile_kk", expect_kk, norms[0]);
    expect_norm("TRSM", key, "tile_mk", expect_mk, norms[1]);
    tile_mk.set_norm(norms[2]);
#endif
    co_await ttg::device::forward(ttg::device::broadcast<0, 1, 2, 3>(
        std::make_tuple(key, Key2(K, M), std::move(keylist_row), std::move(keylist_col)), std::move(tile_mk), out));
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_kk, ttg::fuse(input_mk, input_disp)),
                                                ttg::edges(output_result, output_diag, output_row, output_col), "TRSM",
                                                {"tile_kk", "tile_mk/dispatcher"},
                                                {"output_result", "tile_mk", "output_row", "output_col"});
}

// SYRK(K,M): tile (M,M) -= tile (M,K) tile (M,K)^T (lower); then POTRF(K+1)
// when M == K+1, else SYRK(K+1,M) (potrf.h:380-500).
template <typename MatrixT>
auto make_syrk(MatrixT &A, ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input_disp,
               ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input_mk,
               ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input_kk,
               ttg::Edge<Key1, MatrixTile<typename MatrixT::element_type>> &output_potrf,
               ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &output_syrk) {
  using T = typename MatrixT::element_type;
  auto f = [](const Key2 &key, const MatrixTile<T> &tile_mk, MatrixTile<T> &&tile_kk,
              std::tuple<ttg::Out<Key1, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int K = key[0], M = key[1];
    assert(tile_kk.rows() == tile_kk.cols() && tile_mk.rows() == tile_kk.rows());
#ifdef DEBUG_TILES_VALUES
    const double expect_mk = tile_mk.norm(), expect_kk = tile_kk.norm();
    std::array<double, 3> norms{};  // inputs tile_mk, tile_kk; output tile_kk
    auto norms_s = ttg::make_scratch(norms.data(), ttg::scope::Allocate, norms.size());
    co_await ttg::device::select(tile_kk.buffer(), tile_mk.buffer(), norms_s);
    device_norm(tile_mk, norms_s.device_ptr() + 0);
    device_norm(tile_kk, norms_s.device_ptr() + 1);
#else
    co_await ttg::device::select(tile_kk.buffer(), tile_mk.buffer());
#endif
    check_launch(tile_syrk(ttg::device::current_stream(), (int)tile_mk.rows(), (int)tile_mk.cols(), -1.0,
                           tile_mk.buffer().current_device_ptr(), (int)tile_mk.lda(), 1.0,
                           tile_kk.buffer().current_device_ptr(), (int)tile_kk.lda()));
#ifdef DEBUG_TILES_VALUES
    device_norm(tile_kk, norms_s.device_ptr() + 2);
    co_await ttg::device::wait(norms_s);
    expect_norm("SYRK", key, "tile_mk", expect_mk, norms[0]);
    expect_norm("SYRK", key, "tile_kk", expect_kk, norms[1]);
    tile_kk.set_norm(norms[2]);
#endif
    if (M == K + 1)
      co_await ttg::device::send<0>(Key1(K + 1), std::move(tile_kk), out);
    else
      co_await ttg::device::send<1>(Key2(K + 1, M), std::move(tile_kk), out);
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_mk, ttg::fuse(input_kk, input_disp)),
                                                ttg::edges(output_potrf, output_syrk), "SYRK",
                                                {"tile_mk", "tile_kk/dispatcher"}, {"output_potrf", "output_syrk"});
}

// GEMM(M,N,K): tile (M,N) -= tile (M,K) tile (N,K)^T; then TRSM(M,N) when
// N == K+1, else GEMM(M,N,K+1) (potrf.h:502-631).
template <typename MatrixT>
auto make_gemm(MatrixT &A, ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &input_disp,
               ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &input_mk,
               ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &input_nk,
               ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &input_mn,
               ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &output_trsm,
               ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &output_gemm) {
  using T = typename MatrixT::element_type;
  auto f = [](const Key3 &key, const MatrixTile<T> &tile_mk, const MatrixTile<T> &tile_nk, MatrixTile<T> &&tile_mn,
              std::tuple<ttg::Out<Key2, MatrixTile<T>>, ttg::Out<Key3, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int M = key[0], N = key[1], K = key[2];
    assert(M != N && M > K && N > K);
#ifdef DEBUG_TILES_VALUES
    const double expect_mk = tile_mk.norm(), expect_nk = tile_nk.norm(), expect_mn = tile_mn.norm();
    std::array<double, 4> norms{};  // inputs tile_mk, tile_nk, tile_mn; output tile_mn
    auto norms_s = ttg::make_scratch(norms.data(), ttg::scope::Allocate, norms.size());
    co_await ttg::device::select(tile_mk.buffer(), tile_nk.buffer(), tile_mn.buffer(), norms_s);
    device_norm(tile_mk, norms_s.device_ptr() + 0);
    device_norm(tile_nk, norms_s.device_ptr() + 1);
    device_norm(tile_mn, norms_s.device_ptr() + 2);
#else
    co_await ttg::device::select(tile_mk.buffer(), tile_nk.buffer(), tile_mn.buffer());
#endif
    check_launch(tile_gemm(ttg::device::current_stream(), (int)tile_mk.rows(), (int)tile_nk.rows(),
                           (int)tile_nk.cols(), -1.0, tile_mk.buffer().current_device_ptr(), (int)tile_mk.lda(),
                           tile_nk.buffer().current_device_ptr(), (int)tile_nk.lda(), 1.0,
                           tile_mn.buffer().current_device_ptr(), (int)tile_mn.lda()));
#ifdef DEBUG_TILES_VALUES
    device_norm(tile_mn, norms_s.device_ptr() + 3);
    co_await ttg::device::wait(norms_s);
    expect_norm("GEMM", key, "tile_mk", expect_mk, norms[0]);
    expect_norm("GEMM", key, "tile_nk", expect_nk, norms[1]);
    expect_norm("GEMM", key, "tile_mn", expect_mn, norms[2]);
    tile_mn.set_norm(norms[3]);
#endif
    if (N == K + 1)
      co_await ttg::device::send<0>(Key2(M, N), std::move(tile_mn), out);
    else
      co_await ttg::device::send<1>(Key3(M, N, K + 1), std::move(tile_mn), out);
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_mk, input_nk, ttg::fuse(input_disp, input_mn)),
                                                ttg::edges(output_trsm, output_gemm), "GEMM",
                                                {"input_mk", "input_kn", "input_mn/dispatcher"},
                                                {"output_trsm", "output_gemm"});
}

// Dispatcher: routes each initial tile (potrf.h:633-668).
template <typename T>
auto make_dispatcher(ttg::Edge<Key2, MatrixTile<T>> &input, ttg::Edge<Key1, MatrixTile<T>> &to_potrf,
                     ttg::Edge<Key2, MatrixTile<T>> &to_trsm, ttg::Edge<Key2, MatrixTile<T>> &to_syrk,
                     ttg::Edge<Key3, MatrixTile<T>> &to_gemm) {
  auto f = [](const Key2 &key, const MatrixTile<T> &tile,
              std::tuple<ttg::Out<Key1, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>,
                         ttg::Out<Key3, MatrixTile<T>>> &out) {
    if (key[0] == 0 && key[1] == 0) {
      ttg::send<0>(Key1{key[0]}, tile, out);
    } else if (key[0] == key[1]) {
      ttg::send<2>(Key2{0, key[0]}, tile, out);
    } else if (key[1] == 0) {
      ttg::send<1>(key, tile, out);
    } else {
      ttg::send<3>(Key3{key[0], key[1], 0}, tile, out);
    }
  };
  return ttg::make_tt(f, ttg::edges(input), ttg::edges(to_potrf, to_trsm, to_syrk, to_gemm), "POTRF Dispatch",
                      {"Input"}, {"POTRF", "TRSM", "SYRK", "GEMM"});
}

// The whole graph (potrf.h:670-764): input = tiles (m,n) with n <= m, output =
// the final tiles of L.
template <typename MatrixT>
auto make_potrf_ttg(MatrixT &A, ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input,
                    ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &output, bool defer_write,
                    bool enable_device_map = true) {
  using T = typename MatrixT::element_type;
  ttg::Edge<Key1, MatrixTile<T>> syrk_potrf("syrk_potrf"), disp_potrf("disp_potrf");
  ttg::Edge<Key2, MatrixTile<T>> potrf_trsm("potrf_trsm"), trsm_syrk("trsm_syrk"), gemm_trsm("gemm_trsm"),
      syrk_syrk("syrk_syrk"), disp_trsm("disp_trsm"), disp_syrk("disp_syrk");
  ttg::Edge<Key3, MatrixTile<T>> gemm_gemm("gemm_gemm"), trsm_gemm_row("trsm_gemm_row"),
      trsm_gemm_col("trsm_gemm_col"), disp_gemm("disp_gemm");

  auto tt_dispatch = make_dispatcher(input, disp_potrf, disp_trsm, disp_syrk, disp_gemm);
  tt_dispatch->set_defer_writer(defer_write);
  auto tt_potrf = make_potrf(A, disp_potrf, syrk_potrf, potrf_trsm, output);
  tt_potrf->set_defer_writer(defer_write);
  auto tt_trsm = make_trsm(A, disp_trsm, potrf_trsm, gemm_trsm, trsm_syrk, trsm_gemm_row, trsm_gemm_col, output);
  tt_trsm->set_defer_writer(defer_write);
  auto tt_syrk = make_syrk(A, disp_syrk, trsm_syrk, syrk_syrk, syrk_potrf, syrk_syrk);
  tt_syrk->set_defer_writer(defer_write);
  auto tt_gemm = make_gemm(A, disp_gemm, trsm_gemm_row, trsm_gemm_col, gemm_gemm, gemm_trsm, gemm_gemm);
  tt_gemm->set_defer_writer(defer_write);
  (void)enable_device_map;  // one GPU: the row-block device map is the identity

  // DPLASMA priorities (potrf.h:743-752)
  const int nt = (int)A.cols();
  tt_potrf->set_priomap([nt](const Key1 &k) { return (nt - k[0]) * (nt - k[0]) * (nt - k[0]); });
  tt_trsm->set_priomap([nt](const Key2 &k) {
    return (nt - k[0]) * (nt - k[0]) * (nt - k[0]) + 3 * ((2 * nt) - k[1] - k[0] - 1) * (k[0] - k[1]);
  });
  tt_syrk->set_priomap([nt](const Key2 &k) { return (nt - k[0]) * (nt - k[0]) * (nt - k[0]) + 3 * (k[0] - k[1]); });
  tt_gemm->set_priomap([nt](const Key3 &k) {
    return (nt - k[0]) * (nt - k[0]) * (nt - k[0]) + 3 * ((2 * nt) - k[0] - k[1] - 3) * (k[0] - k[1]) +
           6 * (k[0] - k[2]);
  });

  std::vector<std::unique_ptr<ttg::TTBase>> ops;
  ops.push_back(std::move(tt_dispatch));
  ops.push_back(std::move(tt_potrf));
  ops.push_back(std::move(tt_syrk));
  ops.push_back(std::move(tt_trsm));
  ops.push_back(std::move(tt_gemm));
  return ttg::make_ttg(std::move(ops), std::make_tuple(), std::make_tuple(), "POTRF TTG");
}

}  // namespace potrf
