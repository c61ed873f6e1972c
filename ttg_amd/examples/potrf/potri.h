// POTRI as a task graph (examples/potrf/potri.h:17-30 of devreal/ttg): the
// TRTRI graph's output tiles feed the LAUUM graph's dispatcher, so
// A^{-1} = L^{-T} L^{-1} is formed from the factor L in one graph.
#pragma once

#include "lauum.h"
#include "pmw.h"
#include "trtri_L.h"
#include "ttg.h"

namespace potri {

template <typename T>
auto make_potri_ttg(MatrixT<T> &A, ttg::Edge<Key2, MatrixTile<T>> &input, ttg::Edge<Key2, MatrixTile<T>> &output,
                    bool defer_write) {
  ttg::Edge<Key2, MatrixTile<T>> trtri_to_lauum("trtri_to_lauum");
  auto ttg_trtri = trtri_lower::make_trtri_ttg(A, input, trtri_to_lauum, defer_write);
  auto ttg_lauum = lauum::make_lauum_ttg(A, trtri_to_lauum, output, defer_write);
  std::vector<std::unique_ptr<ttg::TTBase>> ops;
  ops.push_back(std::move(ttg_trtri));
  ops.push_back(std::move(ttg_lauum));
  return ttg::make_ttg(std::move(ops), std::make_tuple(), std::make_tuple(), "POTRI TTG");
}

}  // namespace potri
