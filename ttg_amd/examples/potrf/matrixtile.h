// MatrixTile<T>: one column-major tile as a ttg::Buffer plus its shape
// (the tile value type of examples/matrixtile.h:41-188).  A tile either owns
// its host memory or views a tile of a larger matrix; copying a tile copies
// its data.  The runtime finds the buffer through ttg_buffers().
// DEBUG_TILES_VALUES builds carry the tile's 2-norm as its producer last set
// it (matrixtile.h:147-158); a tile that no task produced yet takes the norm
// of its host data.
#pragma once

#include <algorithm>
#include <cstddef>
#include <ostream>
#ifdef DEBUG_TILES_VALUES
#include <optional>

#include "util.h"
#endif

#include "ttg.h"

template <typename T>
class MatrixTile {
 public:
  using value_type = T;
  MatrixTile() = default;
  MatrixTile(std::size_t rows, std::size_t cols, std::size_t lda)
      : buf_(lda * cols), rows_(rows), cols_(cols), lda_(lda) {}
  // view of outside host memory (not freed by the tile)
  MatrixTile(std::size_t rows, std::size_t cols, T *data, std::size_t lda)
      : buf_(data, lda * cols), rows_(rows), cols_(cols), lda_(lda) {}
  MatrixTile(MatrixTile &&) = default;
  MatrixTile &operator=(MatrixTile &&) = default;
  MatrixTile(const MatrixTile &o) : buf_(o.lda_ * o.cols_), rows_(o.rows_), cols_(o.cols_), lda_(o.lda_) {
    std::copy_n(o.data(), lda_ * cols_, data());
#ifdef DEBUG_TILES_VALUES
    norm_ = o.norm_;
#endif
  }
  MatrixTile &operator=(const MatrixTile &o) {
    rows_ = o.rows_;
    cols_ = o.cols_;
    lda_ = o.lda_;
    buf_.reset(lda_ * cols_);
    std::copy_n(o.data(), lda_ * cols_, data());
#ifdef DEBUG_TILES_VALUES
    norm_ = o.norm_;
#endif
    return *this;
  }

  T *data() { return buf_.host_ptr(); }
  const T *data() const { return buf_.host_ptr(); }
  std::size_t size() const { return cols_ * lda_; }
  std::size_t rows() const { return rows_; }
  std::size_t cols() const { return cols_; }
  std::size_t lda() const { return lda_; }
  ttg::Buffer<T> &buffer() { return buf_; }
  const ttg::Buffer<T> &buffer() const { return buf_; }

#ifdef DEBUG_TILES_VALUES
  // the norm its producer set, else the norm of the host data (read while
  // the host copy is the tile's current one: before the task's select)
  T norm() const {
    if (!norm_) norm_ = host_nrm2(size(), data());
    return *norm_;
  }
  void set_norm(T n) { norm_ = n; }
#endif

  friend std::ostream &operator<<(std::ostream &o, const MatrixTile &t) {
    return o << "MatrixTile{rows=" << t.rows_ << " cols=" << t.cols_ << " ld=" << t.lda_ << "}";
  }

 private:
  ttg::Buffer<T> buf_;
  std::size_t rows_ = 0, cols_ = 0, lda_ = 0;
#ifdef DEBUG_TILES_VALUES
  mutable std::optional<T> norm_;
#endif
};

template <typename T>
std::vector<ttg::detail::BufferBase *> ttg_buffers(MatrixTile<T> &t) {
  return {&t.buffer()};
}
