// The lower-triangular product A := L^T L as a task graph
// (examples/potrf/lauum.h of devreal/ttg): the same template tasks, keys,
// successor keys and dispatch, with device bodies on libttg_chol's gfx950
// tile operations (include/ttg_chol.h) in place of the reference's host
// LAPACK++/BLAS++ calls.  Output tile (M,N):
//   N < M: TRMM(M, N)  A(M,N) := L_MM^T L_MN                (:52-90)
//          GEMM(K, N, M) A(M,N) += L_KM^T L_KN, K = M+1 ... (:135-177)
//   N = M: LAUUM(M)    A(M,M) := L_MM^T L_MM (lower)        (:24-49)
//          SYRK(K, M)  A(M,M) += L_KM^T L_KM, K = M+1 ...   (:93-133)
// Every task reads the dispatcher's copies of the original L; the tile a
// chain accumulates is passed along it.  The reference's SYRK passes its
// read-only input tile_kn down the chain instead of the accumulated tile_nn
// (lauum.h:125-127), which would drop every earlier term of A(N,N); this
// graph passes tile_nn, the product LAPACK's dlauum defines.
#pragma once

#include <vector>

#include "pmw.h"
#include "ttg.h"
#include "ttg_chol.h"

namespace lauum {

inline void check_launch(int rc) {
  if (rc != TTG_CHOL_OK) ttg::detail::fatal(std::string("tile kernel: ") + ttg_chol_last_error());
}

template <typename T>
auto make_lauum(const MatrixT<T> &A, ttg::Edge<Key1, MatrixTile<T>> &input_disp,
                ttg::Edge<Key2, MatrixTile<T>> &to_syrk_C, ttg::Edge<Key2, MatrixTile<T>> &output_result) {
  const int nt = (int)A.rows();
  auto f = [nt](const Key1 &key, MatrixTile<T> &&tile_kk,
                std::tuple<ttg::Out<Key2, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int K = key[0];
    co_await ttg::device::select(tile_kk.buffer());
    check_launch(ttg_chol_dlauum_tile(ttg::device::current_stream(), (int)tile_kk.cols(),
                                      tile_kk.buffer().current_device_ptr(), (int)tile_kk.lda()));
    if (K < nt - 1)
      co_await ttg::device::send<0>(Key2{K + 1, K}, std::move(tile_kk), out);
    else
      co_await ttg::device::send<1>(Key2{K, K}, std::move(tile_kk), out);
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_disp), ttg::edges(to_syrk_C, output_result),
                                                "LAUUM", {"tile_kk"}, {"to_syrk_C", "output_result"});
}

// TRMM(K, N): A(K,N) := L_KK^T L_KN; then C input of GEMM(K+1, N, K), or the
// output for the last tile row
template <typename T>
auto make_trmm(const MatrixT<T> &A, ttg::Edge<Key2, MatrixTile<T>> &input_kk, ttg::Edge<Key2, MatrixTile<T>> &input_kn,
               ttg::Edge<Key3, MatrixTile<T>> &to_gemm_C, ttg::Edge<Key2, MatrixTile<T>> &output_result) {
  const int nt = (int)A.rows();
  auto f = [nt](const Key2 &key, const MatrixTile<T> &tile_kk, MatrixTile<T> &&tile_kn,
                std::tuple<ttg::Out<Key3, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int K = key[0], N = key[1];
    co_await ttg::device::select(tile_kk.buffer(), tile_kn.buffer());
    check_launch(ttg_chol_dtrmm_llt(ttg::device::current_stream(), (int)tile_kn.rows(), (int)tile_kn.cols(), 1.0,
                                    tile_kk.buffer().current_device_ptr(), (int)tile_kk.lda(),
                                    tile_kn.buffer().current_device_ptr(), (int)tile_kn.lda()));
    if (K < nt - 1)
      co_await ttg::device::send<0>(Key3{K + 1, N, K}, std::move(tile_kn), out);
    else
      co_await ttg::device::send<1>(Key2{K, N}, std::move(tile_kn), out);
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_kk, input_kn),
                                                ttg::edges(to_gemm_C, output_result), "TRMM", {"tile_kk", "tile_kn"},
                                                {"to_GEMM_C", "output_result"});
}

// SYRK(K, N): A(N,N) += L_KN^T L_KN (lower); then SYRK(K+1, N) or the output
template <typename T>
auto make_syrk(const MatrixT<T> &A, ttg::Edge<Key2, MatrixTile<T>> &input_kn,
               ttg::FusedEdges<Key2, MatrixTile<T>> input_nn, ttg::Edge<Key2, MatrixTile<T>> &to_syrk_nn,
               ttg::Edge<Key2, MatrixTile<T>> &output_result) {
  const int nt = (int)A.rows();
  auto f = [nt](const Key2 &key, const MatrixTile<T> &tile_kn, MatrixTile<T> &&tile_nn,
                std::tuple<ttg::Out<Key2, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int K = key[0], N = key[1];
    co_await ttg::device::select(tile_kn.buffer(), tile_nn.buffer());
    check_launch(ttg_chol_dsyrk_lt(ttg::device::current_stream(), (int)tile_nn.rows(), (int)tile_kn.rows(), 1.0,
                                   tile_kn.buffer().current_device_ptr(), (int)tile_kn.lda(), 1.0,
                                   tile_nn.buffer().current_device_ptr(), (int)tile_nn.lda()));
    if (K < nt - 1)
      co_await ttg::device::send<0>(Key2{K + 1, N}, std::move(tile_nn), out);
    else
      co_await ttg::device::send<1>(Key2{N, N}, std::move(tile_nn), out);
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_kn, input_nn),
                                                ttg::edges(to_syrk_nn, output_result), "SYRK", {"tile_kn", "tile_nn"},
                                                {"SYRK_nn", "output_result"});
}

// GEMM(K, N, M): A(M,N) += L_KM^T L_KN; then GEMM(K+1, N, M) or the output
template <typename T>
auto make_gemm(const MatrixT<T> &A, ttg::Edge<Key3, MatrixTile<T>> &input_A, ttg::Edge<Key3, MatrixTile<T>> &input_B,
               ttg::FusedEdges<Key3, MatrixTile<T>> input_C, ttg::Edge<Key3, MatrixTile<T>> &to_gemm_C,
               ttg::Edge<Key2, MatrixTile<T>> &output_result) {
  const int nt = (int)A.rows();
  auto f = [nt](const Key3 &key, const MatrixTile<T> &in_A, const MatrixTile<T> &in_B, MatrixTile<T> &&in_C,
                std::tuple<ttg::Out<Key3, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int K = key[0], N = key[1], M = key[2];
    co_await ttg::device::select(in_A.buffer(), in_B.buffer(), in_C.buffer());
    check_launch(ttg_chol_dgemm_tn(ttg::device::current_stream(), (int)in_A.cols(), (int)in_B.cols(),
                                   (int)in_A.rows(), 1.0, in_A.buffer().current_device_ptr(), (int)in_A.lda(),
                                   in_B.buffer().current_device_ptr(), (int)in_B.lda(), 1.0,
                                   in_C.buffer().current_device_ptr(), (int)in_C.lda()));
    if (K < nt - 1)
      co_await ttg::device::send<0>(Key3{K + 1, N, M}, std::move(in_C), out);
    else
      co_await ttg::device::send<1>(Key2{M, N}, std::move(in_C), out);
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_A, input_B, input_C),
                                                ttg::edges(to_gemm_C, output_result), "GEMM", {"A", "B", "C"},
                                                {"GEMM_C", "output result"});
}

// Dispatcher (:181-251): tile (m,m) to LAUUM(m) and, as L_mm, to TRMM(m, n)
// for n < m; tile (m,n), m > n, to SYRK(m, n), TRMM(m, n) (the tile it
// rewrites), A of GEMM(m, n', n) for n' < n and B of GEMM(m, n, n') for
// n < n' < m
template <typename T>
auto make_dispatcher(const MatrixT<T> &A, ttg::Edge<Key2, MatrixTile<T>> &input, ttg::Edge<Key1, MatrixTile<T>> &to_lauum,
                     ttg::Edge<Key2, MatrixTile<T>> &to_syrk, ttg::Edge<Key2, MatrixTile<T>> &to_trmm_A,
                     ttg::Edge<Key2, MatrixTile<T>> &to_trmm_B, ttg::Edge<Key3, MatrixTile<T>> &to_gemm_A,
                     ttg::Edge<Key3, MatrixTile<T>> &to_gemm_B) {
  auto f = [](const Key2 &key, const MatrixTile<T> &tile,
              std::tuple<ttg::Out<Key1, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>,
                         ttg::Out<Key2, MatrixTile<T>>, ttg::Out<Key3, MatrixTile<T>>, ttg::Out<Key3, MatrixTile<T>>>
                  &out) {
    const int m = key[0], n = key[1];
    std::vector<Key1> keys_lauum;
    std::vector<Key2> keys_syrk, keys_trmm_A, keys_trmm_B;
    std::vector<Key3> keys_gemm_A, keys_gemm_B;
    if (m == n) {
      keys_lauum.push_back(Key1{m});
      for (int j = 0; j < n; ++j) keys_trmm_A.push_back(Key2{n, j});
    }
    if (m > n) {
      keys_syrk.push_back(key);
      keys_trmm_B.push_back(key);
      for (int j = 0; j < n; ++j) keys_gemm_A.push_back(Key3{m, j, n});
    }
    for (int j = n + 1; j < m; ++j) keys_gemm_B.push_back(Key3{m, n, j});
    ttg::broadcast<0, 1, 2, 3, 4, 5>(std::make_tuple(std::move(keys_lauum), std::move(keys_syrk), std::move(keys_trmm_A),
                                                     std::move(keys_trmm_B), std::move(keys_gemm_A),
                                                     std::move(keys_gemm_B)),
                                     tile, out);
  };
  return ttg::make_tt(f, ttg::edges(input), ttg::edges(to_lauum, to_syrk, to_trmm_A, to_trmm_B, to_gemm_A, to_gemm_B),
                      "LAUUM Dispatch", {"Input"}, {"LAUUM", "SYRK", "TRMM_A", "TRMM_B", "GEMM_A", "GEMM_B"});
}

// The whole graph (:253-322): input = the lower tiles of L, output = the
// lower tiles of L^T L
template <typename T>
auto make_lauum_ttg(MatrixT<T> &A, ttg::Edge<Key2, MatrixTile<T>> &input, ttg::Edge<Key2, MatrixTile<T>> &output,
                    bool defer_write) {
  ttg::Edge<Key1, MatrixTile<T>> disp_lauum("disp_lauum");
  ttg::Edge<Key2, MatrixTile<T>> disp_syrk("disp_syrk"), disp_trmm_A("disp_trmm_A"), disp_trmm_B("disp_trmm_B"),
      lauum_syrk_nn("lauum_syrk_nn"), syrk_syrk_nn("syrk_syrk_nn");
  ttg::Edge<Key3, MatrixTile<T>> disp_gemm_A("disp_gemm_A"), disp_gemm_B("disp_gemm_B"), trmm_gemm_C("trmm_gemm_C"),
      gemm_gemm_C("gemm_gemm_C");

  auto tt_dispatch = make_dispatcher(A, input, disp_lauum, disp_syrk, disp_trmm_A, disp_trmm_B, disp_gemm_A, disp_gemm_B);
  tt_dispatch->set_defer_writer(defer_write);
  auto tt_lauum = make_lauum(A, disp_lauum, lauum_syrk_nn, output);
  tt_lauum->set_defer_writer(defer_write);
  auto tt_syrk = make_syrk(A, disp_syrk, ttg::fuse(lauum_syrk_nn, syrk_syrk_nn), syrk_syrk_nn, output);
  tt_syrk->set_defer_writer(defer_write);
  auto tt_trmm = make_trmm(A, disp_trmm_A, disp_trmm_B, trmm_gemm_C, output);
  tt_trmm->set_defer_writer(defer_write);
  auto tt_gemm = make_gemm(A, disp_gemm_A, disp_gemm_B, ttg::fuse(trmm_gemm_C, gemm_gemm_C), gemm_gemm_C, output);
  tt_gemm->set_defer_writer(defer_write);

  std::vector<std::unique_ptr<ttg::TTBase>> ops;
  ops.push_back(std::move(tt_dispatch));
  ops.push_back(std::move(tt_lauum));
  ops.push_back(std::move(tt_syrk));
  ops.push_back(std::move(tt_trmm));
  ops.push_back(std::move(tt_gemm));
  return ttg::make_ttg(std::move(ops), std::make_tuple(), std::make_tuple(), "LAUUM TTG");
}

}  // namespace lauum
