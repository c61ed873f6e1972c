// Address-range bookkeeping of a launch group (host code only).
//
// A launch group batches the tile calls of one scheduling round by kind, so
// calls inside it run in no particular order.  Before a call is recorded,
// its footprint -- the ranges it reads and the ranges it writes -- is checked
// against the calls already in the group: a call that reads what an earlier
// call writes, or writes what an earlier call reads or writes, depends on it,
// and the group so far must run first.  The sets keep the union of the
// ranges, as disjoint half-open intervals, so each check is one ordered
// lookup.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace ttgc {

// [a, b) in bytes
struct Span {
  uintptr_t a = 0, b = 0;
};

// the bytes an m x n column-major block with leading dimension ld spans
// (from its first element to one past its last); empty for an empty block
inline Span block_span(const void *p, long m, long n, long ld, size_t elem) {
  if (!p || m <= 0 || n <= 0) return Span{};
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  return Span{a, a + (uintptr_t)(((n - 1) * ld + m) * (long)elem)};
}

class SpanSet {
 public:
  // whether [s.a, s.b) meets a range of the set
  bool hits(const Span &s) const {
    if (s.a >= s.b || iv_.empty()) return false;
    auto it = iv_.lower_bound(s.b);  // first range starting at or after s.b
    if (it == iv_.begin()) return false;
    --it;  // the last range starting before s.b
    return it->second > s.a;
  }
  void add(const Span &s) {
    if (s.a >= s.b) return;
    uintptr_t a = s.a, b = s.b;
    auto it = iv_.upper_bound(a);
    if (it != iv_.begin()) {
      auto p = std::prev(it);
      if (p->second >= a) {  // touches or overlaps the range before
        a = p->first;
        if (p->second > b) b = p->second;
        it = iv_.erase(p);
      }
    }
    while (it != iv_.end() && it->first <= b) {  // ranges that start inside
      if (it->second > b) b = it->second;
      it = iv_.erase(it);
    }
    iv_.emplace(a, b);
  }
  void clear() { iv_.clear(); }
  size_t size() const { return iv_.size(); }

 private:
  std::map<uintptr_t, uintptr_t> iv_;  // start -> end, disjoint, not touching
};

// The reads and writes of the calls recorded in one group.
struct Footprints {
  SpanSet rd, wr;
  // whether a call with these reads and writes depends on a recorded call
  template <size_t NI, size_t NO>
  bool conflicts(const Span (&ins)[NI], const Span (&outs)[NO]) const {
    for (const Span &s : ins)
      if (wr.hits(s)) return true;
    for (const Span &s : outs)
      if (wr.hits(s) || rd.hits(s)) return true;
    return false;
  }
  template <size_t NI, size_t NO>
  void add(const Span (&ins)[NI], const Span (&outs)[NO]) {
    for (const Span &s : ins) rd.add(s);
    for (const Span &s : outs) wr.add(s);
  }
  void clear() {
    rd.clear();
    wr.clear();
  }
};

}  // namespace ttgc
