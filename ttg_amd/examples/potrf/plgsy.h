// PLGSY as a device task: one task per lower tile generates the tile in HBM
// with libttg_chol's CORE_dplgsy kernel (the reference's PLGSY task,
// examples/potrf/plgsy.h:7-50, runs core_plgsy on the host; the values are
// bit-identical).
#pragma once

#include "pmw.h"
#include "ttg.h"
#include "ttg_chol.h"

template <typename T>
auto make_plgsy(MatrixT<T> &A, unsigned long bump, unsigned long seed, ttg::Edge<Key2, void> &input,
                ttg::Edge<Key2, MatrixTile<T>> &output) {
  auto f = [&A, bump, seed](const Key2 &key, std::tuple<ttg::Out<Key2, MatrixTile<T>>> &out) -> ttg::device::Task {
    const int I = key[0], J = key[1];
    auto tile = A(I, J);
    tile.buffer().discard_host();  // produced on the device
    co_await ttg::device::select(tile.buffer());
    if (ttg_chol_dplgsy_tile(ttg::device::current_stream(), (double)bump, (int)tile.rows(), (int)tile.cols(),
                             tile.buffer().current_device_ptr(), (int)tile.lda(), (int)A.rows_in_matrix(),
                             I * (int)A.rows_in_tile(), J * (int)A.cols_in_tile(), seed) != TTG_CHOL_OK)
      ttg::detail::fatal(ttg_chol_last_error());
    co_await ttg::device::send<0>(key, std::move(tile), out);
  };
  return ttg::make_tt<ttg::ExecutionSpace::HIP>(f, ttg::edges(input), ttg::edges(output), "PLGSY", {"startup"},
                                                {"output"});
}

template <typename T>
auto make_plgsy_ttg(MatrixT<T> &A, unsigned long bump, unsigned long seed, ttg::Edge<Key2, void> &startup,
                    ttg::Edge<Key2, MatrixTile<T>> &result, bool defer_write) {
  auto tt = make_plgsy(A, bump, seed, startup, result);
  tt->set_defer_writer(defer_write);
  std::vector<std::unique_ptr<ttg::TTBase>> ops;
  ops.push_back(std::move(tt));
  return ttg::make_ttg(std::move(ops), std::make_tuple(), std::make_tuple(), "PLGSY TTG");
}
