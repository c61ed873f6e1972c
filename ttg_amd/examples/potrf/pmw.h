// TiledMatrix<T>: the lower tile-packed symmetric matrix the graph works on
// (in place of the PaRSEC sym_block_cyclic collection wrapped by
// examples/potrf/pmw.h:35-143, for one process).  Tile (m, n), n <= m, is
// NB x NB column-major with leading dimension NB; edge tiles hold N - m*NB
// valid rows.  The layout is the one of libttg_chol's tiled drivers.
#pragma once

#include <cstddef>
#include <cstring>
#include <memory>
#include <vector>

#include "matrixtile.h"
#include "ttg.h"

using Key1 = ttg::MultiIndex<1>;
using Key2 = ttg::MultiIndex<2>;
using Key3 = ttg::MultiIndex<3>;

template <typename T>
class TiledMatrix {
 public:
  using element_type = T;
  TiledMatrix(std::size_t N, std::size_t NB, T *storage = nullptr)
      : N_(N), NB_(NB), nt_((N + NB - 1) / NB) {
    if (!storage) {
      // pinned host memory: the tiles move between host and HBM at full
      // link speed (the push-out of every final tile, the loads)
      void *p = nullptr;
      const std::size_t bytes = num_tiles() * NB * NB * sizeof(T);
      if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) ttg::detail::fatal("hipHostMalloc");
      std::memset(p, 0, bytes);
      own_.reset(static_cast<T *>(p));
      storage = own_.get();
    }
    data_ = storage;
  }
  std::size_t num_tiles() const { return nt_ * (nt_ + 1) / 2; }
  std::size_t tile_index(std::size_t m, std::size_t n) const { return n * nt_ - n * (n - 1) / 2 + (m - n); }
  T *tile_ptr(std::size_t m, std::size_t n) const { return data_ + tile_index(m, n) * NB_ * NB_; }

  // a view of tile (row, col)
  MatrixTile<T> operator()(std::size_t row, std::size_t col) const {
    const std::size_t mb = row + 1 < nt_ ? NB_ : N_ - row * NB_;
    const std::size_t nb = col + 1 < nt_ ? NB_ : N_ - col * NB_;
    return MatrixTile<T>(mb, nb, tile_ptr(row, col), NB_);
  }
  std::size_t rows() const { return nt_; }
  std::size_t cols() const { return nt_; }
  std::size_t rows_in_tile() const { return NB_; }
  std::size_t cols_in_tile() const { return NB_; }
  std::size_t rows_in_matrix() const { return N_; }
  std::size_t cols_in_matrix() const { return N_; }
  std::size_t rank_of(std::size_t, std::size_t) const { return 0; }
  bool is_local(std::size_t, std::size_t) const { return true; }
  bool in_matrix(std::size_t row, std::size_t col) const { return col <= row; }
  int P() const { return 1; }
  int Q() const { return 1; }
  T *data() const { return data_; }

  // dense column-major copy of the lower tiles as stored, diagonal tiles
  // whole and zeros above them (pmw.h:122-139)
  std::vector<T> getLAPACKMatrix() const {
    std::vector<T> A(N_ * N_, T(0));
    for (std::size_t j = 0; j < N_; ++j)
      for (std::size_t i = (j / NB_) * NB_; i < N_; ++i)
        A[i + j * N_] = tile_ptr(i / NB_, j / NB_)[(i % NB_) + (j % NB_) * NB_];
    return A;
  }

 private:
  struct HostFree {
    void operator()(T *p) const { (void)hipHostFree(p); }
  };
  std::size_t N_, NB_, nt_;
  std::unique_ptr<T, HostFree> own_;
  T *data_ = nullptr;
};

template <typename T>
using MatrixT = TiledMatrix<T>;

// load TT: sends every lower tile, read-write, into the graph (pmw.h:145-161)
template <typename T>
auto make_load_tt(MatrixT<T> &A, ttg::Edge<Key2, MatrixTile<T>> &toop, bool defer_write) {
  auto load_tt = ttg::make_tt<void>(
      [&](std::tuple<ttg::Out<Key2, MatrixTile<T>>> &out) {
        for (std::size_t i = 0; i < A.rows(); i++)
          for (std::size_t j = 0; j <= i; j++)
            if (A.is_local(i, j)) ttg::send<0>(Key2{(int)i, (int)j}, A(i, j), out);
      },
      ttg::edges(), ttg::edges(toop), "Load Matrix", {}, {"To Op"});
  load_tt->set_defer_writer(defer_write);
  return load_tt;
}
