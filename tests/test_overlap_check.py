"""The address-range bookkeeping a launch group uses to find calls that
depend on earlier calls of the group (ttg_amd/csrc/chol_overlap.hpp), on the
CPU: a small C++ program built with g++ against the header checks block
spans, the merging interval set and the read/write conflict rule, including
against a brute-force byte map on random blocks."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <cassert>
#include <cstdio>
#include <random>
#include <vector>
#include "chol_overlap.hpp"
using namespace ttgc;

int main() {
  alignas(64) static double buf[4096];
  // block spans: 3 x 4 block, ld 10: elements 0 .. 3*10+3 = 33
  Span s = block_span(buf, 3, 4, 10, sizeof(double));
  assert(s.b - s.a == 33 * sizeof(double));
  assert(block_span(buf, 0, 4, 10, 8).a == block_span(buf, 0, 4, 10, 8).b);
  // interval set: merging, touching, nesting
  SpanSet set;
  auto sp = [&](long a, long b) { return Span{(uintptr_t)(buf + a), (uintptr_t)(buf + b)}; };
  set.add(sp(10, 20));
  set.add(sp(30, 40));
  assert(!set.hits(sp(0, 10)) && !set.hits(sp(20, 30)) && !set.hits(sp(40, 50)));
  assert(set.hits(sp(19, 21)) && set.hits(sp(0, 11)) && set.hits(sp(35, 36)) && set.hits(sp(0, 100)));
  set.add(sp(20, 30));  // touches both: one range
  assert(set.size() == 1 && set.hits(sp(25, 26)));
  set.add(sp(5, 45));   // covers it
  assert(set.size() == 1 && set.hits(sp(44, 45)) && !set.hits(sp(45, 46)));
  set.add(sp(100, 110));
  set.add(sp(60, 70));
  assert(set.size() == 3 && !set.hits(sp(70, 100)) && set.hits(sp(69, 101)));
  // conflict rule: read after write, write after read, write after write
  Footprints fp;
  const Span none[1];
  const Span tileA[1] = {block_span(buf, 8, 8, 8, 8)};          // elements 0 .. 63
  const Span tileB[1] = {block_span(buf + 64, 8, 8, 8, 8)};     // 64 .. 127
  const Span tileC[1] = {block_span(buf + 128, 8, 8, 8, 8)};    // 128 .. 191
  fp.add(tileA, tileB);                    // reads A, writes B
  assert(!fp.conflicts(tileA, tileC));     // another reader of A, writes C: independent
  assert(fp.conflicts(tileB, tileC));      // reads B: written before
  assert(fp.conflicts(none, tileA));       // writes A: read before
  assert(fp.conflicts(none, tileB));       // writes B: written before
  fp.clear();
  assert(!fp.conflicts(tileB, tileA));
  // random blocks against a byte map
  std::mt19937 rng(7);
  for (int trial = 0; trial < 200; ++trial) {
    Footprints f;
    std::vector<int> rd(4096, 0), wr(4096, 0);
    for (int op = 0; op < 12; ++op) {
      auto rnd_block = [&](int &lo, int &hi) {
        const int ld = 1 + (int)(rng() % 40), m = 1 + (int)(rng() % ld), n = 1 + (int)(rng() % 20);
        const int off = (int)(rng() % 3000);
        lo = off;
        hi = off + (n - 1) * ld + m;
        return block_span(buf + off, m, n, ld, sizeof(double));
      };
      int ilo, ihi, olo, ohi;
      const Span in[1] = {rnd_block(ilo, ihi)}, out[1] = {rnd_block(olo, ohi)};
      bool want = false;
      for (int e = ilo; e < ihi; ++e) want = want || wr[e];
      for (int e = olo; e < ohi; ++e) want = want || wr[e] || rd[e];
      const bool got = f.conflicts(in, out);
      if (got != want) {
        std::printf("mismatch trial %d op %d: got %d want %d\n", trial, op, got, want);
        return 1;
      }
      if (got) {  // the group runs what it holds and starts over
        f.clear();
        std::fill(rd.begin(), rd.end(), 0);
        std::fill(wr.begin(), wr.end(), 0);
      }
      f.add(in, out);
      for (int e = ilo; e < ihi; ++e) rd[e] = 1;
      for (int e = olo; e < ohi; ++e) wr[e] = 1;
    }
  }
  std::printf("ok\n");
  return 0;
}
"""


def test_group_footprint_rules(tmp_path):
    src = tmp_path / "overlap_check.cc"
    src.write_text(PROG)
    exe = tmp_path / "overlap_check"
    cc = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "ttg_amd", "csrc"), str(src), "-o",
                         str(exe)], capture_output=True, text=True)
    if cc.returncode != 0:
        pytest.fail(cc.stderr)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"
