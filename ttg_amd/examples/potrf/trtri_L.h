// The lower triangular inverse as a task graph (examples/potrf/trtri_L.h of
// devreal/ttg): the same template tasks, keys, successor keys and dispatch,
// with device bodies on libttg_chol's gfx950 tile operations
// (include/ttg_chol.h) in place of the reference's host LAPACK++/BLAS++
// calls.  Step K of L := L^{-1}:
//   TRSMR(K, M)   A(M,K) := -A(M,K) L_KK^{-1}             M > K  (:96-145)
//   GEMM(K, M, N) A(M,N) += A(M,K) A(K,N)                 M > K > N (:160-214)
//   TRSML(K, N)   A(K,N) := L_KK^{-1} A(K,N)              N < K  (:49-79)
//   TRTRI(K)      A(K,K) := L_KK^{-1}                             (:24-40)
// where L_KK is the dispatcher's copy of the diagonal tile (read-only for
// TRSMR and TRSML, written by TRTRI: a writer of a value others read, which
// the runtime defers or copies).  A non-invertible diagonal tile stops the
// graph with LAPACK's info (the reference asserts instead, :33).
#pragma once

#include <iostream>
#include <vector>

#include "pmw.h"
#include "ttg.h"
#include "ttg_chol.h"

namespace trtri_lower {

inline void check_launch(int rc) {
  if (rc != TTG_CHOL_OK) ttg::detail::fatal(std::string("tile kernel: ") + ttg_chol_last_error());
}

template <typename MatrixT>
auto make_trtri(const MatrixT &A, ttg::Edge<Key1, MatrixTile<typename MatrixT::element_type>> &input_disp,
                ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &output_result) {
  using T = typename MatrixT::element_type;
  const long NB = (long)A.rows_in_tile();
  auto f = [NB](const Key1 &key, MatrixTile<T> &&tile_kk,
                std::tuple<ttg::Out<Key2, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int K = key[0];
    int host_info = -1;
    auto dev_info = ttg::make_scratch(&host_info, ttg::scope::Allocate);
    co_await ttg::device::select(tile_kk.buffer(), dev_info);
    check_launch(ttg_chol_dtrtri_tile(ttg::device::current_stream(), (int)tile_kk.cols(),
                                      tile_kk.buffer().current_device_ptr(), (int)tile_kk.lda(), dev_info.device_ptr(),
                                      0));
    co_await ttg::device::wait(dev_info);
    if (host_info != 0) {
      std::cerr << "TRTRI of tile " << K << " failed: " << host_info << std::endl;
      ttg::cancel(K * NB + host_info);
      co_return;
    }
    co_await ttg::device::send<0>(Key2{K, K}, std::move(tile_kk), out);
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_disp), ttg::edges(output_result), "TRTRI",
                                                {"tile_kk"}, {"output_result"});
}

// TRSML(K, N): A(K,N) := L_KK^{-1} A(K,N), then to the output
template <typename MatrixT>
auto make_trsml(const MatrixT &A, ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input_kk,
                ttg::FusedEdges<Key2, MatrixTile<typename MatrixT::element_type>> input_kn,
                ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &output_result) {
  using T = typename MatrixT::element_type;
  auto f = [](const Key2 &key, const MatrixTile<T> &tile_kk, MatrixTile<T> &&tile_kn,
              std::tuple<ttg::Out<Key2, MatrixTile<T>>> &out) -> ttg::device::Task {
    co_await ttg::device::select(tile_kk.buffer(), tile_kn.buffer());
    check_launch(ttg_chol_dtrsm_lln(ttg::device::current_stream(), (int)tile_kn.rows(), (int)tile_kn.cols(), 1.0,
                                    tile_kk.buffer().current_device_ptr(), (int)tile_kk.lda(),
                                    tile_kn.buffer().current_device_ptr(), (int)tile_kn.lda()));
    co_await ttg::device::send<0>(key, std::move(tile_kn), out);
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_kk, input_kn), ttg::edges(output_result),
                                                "TRSML", {"tile_kk", "tile_kn"}, {"output_result"});
}

// TRSMR(K, M): A(M,K) := -A(M,K) L_KK^{-1}; then A input of GEMM(K, M, k)
// for k < K, B input of GEMM(K+1, m, K) for m > M when M = K+1, and C input
// of GEMM(K+1, M, K) (M > K+1) or TRSML(K+1, K) (M = K+1)
template <typename MatrixT>
auto make_trsmr(const MatrixT &A, ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input_kk,
                ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input_mk,
                ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &to_gemm_A,
                ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &to_gemm_B,
                ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &to_gemm_C,
                ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &to_trsml_kn) {
  using T = typename MatrixT::element_type;
  const int nt = (int)A.rows();
  auto f = [nt](const Key2 &key, const MatrixTile<T> &tile_mm, MatrixTile<T> &&tile_kn,
                std::tuple<ttg::Out<Key3, MatrixTile<T>>, ttg::Out<Key3, MatrixTile<T>>, ttg::Out<Key3, MatrixTile<T>>,
                           ttg::Out<Key2, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int K = key[0], M = key[1];
    co_await ttg::device::select(tile_mm.buffer(), tile_kn.buffer());
    check_launch(ttg_chol_dtrsm_rln(ttg::device::current_stream(), (int)tile_kn.rows(), (int)tile_kn.cols(), -1.0,
                                    tile_mm.buffer().current_device_ptr(), (int)tile_mm.lda(),
                                    tile_kn.buffer().current_device_ptr(), (int)tile_kn.lda()));
    std::vector<Key3> keys_a, keys_b, keys_c;
    std::vector<Key2> keys_trsml;
    for (int k = 0; k < K; ++k) keys_a.push_back(Key3{K, M, k});
    if (M == K + 1 && K < nt - 2)
      for (int m = M + 1; m < nt; ++m) keys_b.push_back(Key3{K + 1, m, K});
    if (M > K + 1)
      keys_c.push_back(Key3{K + 1, M, K});
    else if (M == K + 1)
      keys_trsml.push_back(Key2{K + 1, K});
    co_await ttg::device::forward(ttg::device::broadcast<0, 1, 2, 3>(
        std::make_tuple(std::move(keys_a), std::move(keys_b), std::move(keys_c), std::move(keys_trsml)),
        std::move(tile_kn), out));
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_kk, input_mk),
                                                ttg::edges(to_gemm_A, to_gemm_B, to_gemm_C, to_trsml_kn), "TRSMR",
                                                {"tile_kk", "tile_mk"}, {"GEMM_A", "GEMM_B", "GEMM_C", "TRSML_kn"});
}

// GEMM(K, M, N): A(M,N) += A(M,K) A(K,N); then B input of GEMM(K+1, m, N)
// for m > M when M = K+1, C input of GEMM(K+1, M, N) (M > K+1), or
// TRSML(K+1, N) (M = K+1)
template <typename MatrixT>
auto make_gemm(const MatrixT &A, ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &input_A,
               ttg::FusedEdges<Key3, MatrixTile<typename MatrixT::element_type>> input_B,
               ttg::FusedEdges<Key3, MatrixTile<typename MatrixT::element_type>> input_C,
               ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &to_gemm_B,
               ttg::Edge<Key3, MatrixTile<typename MatrixT::element_type>> &to_gemm_C,
               ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &to_trsml_kn) {
  using T = typename MatrixT::element_type;
  const int nt = (int)A.rows();
  auto f = [nt](const Key3 &key, const MatrixTile<T> &in_A, const MatrixTile<T> &in_B, MatrixTile<T> &&in_C,
                std::tuple<ttg::Out<Key3, MatrixTile<T>>, ttg::Out<Key3, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>>
                    &out) -> ttg::device::Task {
    const int K = key[0], M = key[1], N = key[2];
    co_await ttg::device::select(in_A.buffer(), in_B.buffer(), in_C.buffer());
    check_launch(ttg_chol_dgemm_nn(ttg::device::current_stream(), (int)in_A.rows(), (int)in_B.cols(),
                                   (int)in_A.cols(), 1.0, in_A.buffer().current_device_ptr(), (int)in_A.lda(),
                                   in_B.buffer().current_device_ptr(), (int)in_B.lda(), 1.0,
                                   in_C.buffer().current_device_ptr(), (int)in_C.lda()));
    std::vector<Key3> keys_b, keys_c;
    std::vector<Key2> keys_trsml;
    if (M == K + 1) keys_trsml.push_back(Key2{K + 1, N});
    if (M == K + 1 && K < nt - 2)
      for (int m = M + 1; m < nt; ++m) keys_b.push_back(Key3{K + 1, m, N});
    if (M > K + 1) keys_c.push_back(Key3{K + 1, M, N});
    co_await ttg::device::forward(ttg::device::broadcast<0, 1, 2>(
        std::make_tuple(std::move(keys_b), std::move(keys_c), std::move(keys_trsml)), std::move(in_C), out));
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input_A, input_B, input_C),
                                                ttg::edges(to_gemm_B, to_gemm_C, to_trsml_kn), "GEMM", {"A", "B", "C"},
                                                {"GEMM_B", "GEMM_C", "TRSML_kn"});
}

// Dispatcher (:216-260): diagonal tile (K,K) to TRTRI(K), to TRSML(K, k) for
// k < K and TRSMR(K, k) for k > K as tile_kk; tile (m,n), m > n, to TRSMR(n, m)
template <typename MatrixT>
auto make_dispatcher(const MatrixT &A, ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input,
                     ttg::Edge<Key1, MatrixTile<typename MatrixT::element_type>> &to_trtri,
                     ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &to_trsml_kk,
                     ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &to_trsmr_kk,
                     ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &to_trsmr_mk) {
  using T = typename MatrixT::element_type;
  const int nt = (int)A.rows();
  auto f = [nt](const Key2 &key, MatrixTile<T> &&tile,
                std::tuple<ttg::Out<Key1, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>, ttg::Out<Key2, MatrixTile<T>>,
                           ttg::Out<Key2, MatrixTile<T>>> &out) {
    if (key[0] == key[1]) {
      std::vector<Key2> keys_trsml, keys_trsmr;
      for (int k = key[0] + 1; k < nt; ++k) keys_trsmr.push_back(Key2{key[0], k});
      for (int k = 0; k < key[0]; ++k) keys_trsml.push_back(Key2{key[0], k});
      ttg::broadcast<0, 1, 2>(std::make_tuple(Key1{key[0]}, std::move(keys_trsml), std::move(keys_trsmr)),
                              std::move(tile), out);
      return;
    }
    ttg::send<3>(Key2{key[1], key[0]}, std::move(tile), out);
  };
  return ttg::make_tt(f, ttg::edges(input), ttg::edges(to_trtri, to_trsml_kk, to_trsmr_kk, to_trsmr_mk),
                      "TRTRI Dispatch", {"Input"}, {"TRTRI", "TRSML_kk", "TRSMR_kk", "TRSMR_mk"});
}

// The whole graph (:262-339): input = the lower tiles of L, output = the
// tiles of L^{-1}
template <typename MatrixT>
auto make_trtri_ttg(const MatrixT &A, ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &input,
                    ttg::Edge<Key2, MatrixTile<typename MatrixT::element_type>> &output, bool defer_write) {
  using T = typename MatrixT::element_type;
  ttg::Edge<Key1, MatrixTile<T>> disp_trtri("disp_trtri");
  ttg::Edge<Key2, MatrixTile<T>> disp_trsml_kk("disp_trsml_kk"), disp_trsmr_kk("disp_trsmr_kk"),
      disp_trsmr_mk("disp_trsmr_mk"), trsmr_trsml("trsmr_trsml"), gemm_trsml("gemm_trsml");
  ttg::Edge<Key3, MatrixTile<T>> gemm_gemm_B("gemm_gemm_B"), gemm_gemm_C("gemm_gemm_C"), trsmr_gemm_A("trsmr_gemm_A"),
      trsmr_gemm_B("trsmr_gemm_B"), trsmr_gemm_C("trsmr_gemm_C");

  auto tt_dispatch = make_dispatcher(A, input, disp_trtri, disp_trsml_kk, disp_trsmr_kk, disp_trsmr_mk);
  tt_dispatch->set_defer_writer(defer_write);
  auto tt_trtri = make_trtri(A, disp_trtri, output);
  tt_trtri->set_defer_writer(defer_write);
  auto tt_trsml = make_trsml(A, disp_trsml_kk, ttg::fuse(gemm_trsml, trsmr_trsml), output);
  tt_trsml->set_defer_writer(defer_write);
  auto tt_trsmr =
      make_trsmr(A, disp_trsmr_kk, disp_trsmr_mk, trsmr_gemm_A, trsmr_gemm_B, trsmr_gemm_C, trsmr_trsml);
  tt_trsmr->set_defer_writer(defer_write);
  auto tt_gemm = make_gemm(A, trsmr_gemm_A, ttg::fuse(trsmr_gemm_B, gemm_gemm_B), ttg::fuse(trsmr_gemm_C, gemm_gemm_C),
                           gemm_gemm_B, gemm_gemm_C, gemm_trsml);
  tt_gemm->set_defer_writer(defer_write);

  std::vector<std::unique_ptr<ttg::TTBase>> ops;
  ops.push_back(std::move(tt_dispatch));
  ops.push_back(std::move(tt_trtri));
  ops.push_back(std::move(tt_trsml));
  ops.push_back(std::move(tt_trsmr));
  ops.push_back(std::move(tt_gemm));
  return ttg::make_ttg(std::move(ops), std::make_tuple(), std::make_tuple(), "TRTRI TTG");
}

}  // namespace trtri_lower
