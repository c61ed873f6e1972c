// Result sink: a host task per final tile that writes it into the matrix
// when it is not already there (examples/potrf/result.h:7-44).  The runtime
// pushes the tile back from HBM before the task runs.  DEBUG_TILES_VALUES
// builds check the tile's host norm against the norm its producer set
// (result.h:14-17).
#pragma once

#include <algorithm>

#include "pmw.h"
#include "ttg.h"

template <typename T>
auto make_result(MatrixT<T> &A, ttg::Edge<Key2, MatrixTile<T>> &result) {
  auto f = [&A](const Key2 &key, MatrixTile<T> &&tile, std::tuple<> &) {
    const int I = key[0], J = key[1];
    if (ttg::tracing()) ttg::print("RESULT(", key, ")");
#ifdef DEBUG_TILES_VALUES
    expect_norm("RESULT", key, "tile", tile.norm(), host_nrm2(tile.size(), tile.data()));
#endif
    auto atile = A(I, J);
    if (atile.data() != tile.data()) std::copy_n(tile.data(), tile.rows() * tile.cols(), atile.data());
  };
  return ttg::make_tt(f, ttg::edges(result), ttg::edges(), "Final Output", {"result"}, {});
}

template <typename T>
auto make_result_ttg(MatrixT<T> &A, ttg::Edge<Key2, MatrixTile<T>> &result, bool defer_write) {
  auto tt = make_result(A, result);
  tt->set_defer_writer(defer_write);
  std::vector<std::unique_ptr<ttg::TTBase>> ops;
  ops.push_back(std::move(tt));
  return ttg::make_ttg(std::move(ops), std::make_tuple(), std::make_tuple(), "Result Writer");
}
