// MatrixTile<T>: one column-major tile as a ttg::Buffer plus its shape
// (the tile value type of examples/matrixtile.h:41-188).  A tile either owns
// its host memory or views a tile of a larger matrix; copying a tile copies
// its data.  The runtime finds the buffer through ttg_buffers().
#pragma once

#include <algorithm>
#include <cstddef>
#include <ostream>

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
  }
  MatrixTile &operator=(const MatrixTile &o) {
    rows_ = o.rows_;
    cols_ = o.cols_;
    lda_ = o.lda_;
    buf_.reset(lda_ * cols_);
    std::copy_n(o.data(), lda_ * cols_, data());
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

  friend std::ostream &operator<<(std::ostream &o, const MatrixTile &t) {
    return o << "MatrixTile{rows=" << t.rows_ << " cols=" << t.cols_ << " ld=" << t.lda_ << "}";
  }

 private:
  ttg::Buffer<T> buf_;
  std::size_t rows_ = 0, cols_ = 0, lda_ = 0;
};

template <typename T>
std::vector<ttg::detail::BufferBase *> ttg_buffers(MatrixTile<T> &t) {
  return {&t.buffer()};
}
