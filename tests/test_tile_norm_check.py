"""Host side of the tile-norm debug check (examples/potrf/util.h:5-10 of
the reference): check_norm's 1e-12 relative bound, the host 2-norm a tile
takes before any task produced it, and that a mismatch aborts the run with
the task, key and tile named.  Compiled and run on the CPU (no device
call)."""
import os
import subprocess

PROG = r"""
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "util.h"
int main(int argc, char **argv) {
  if (!check_norm(2.0, 2.0 * (1 + 0.9e-12)) || check_norm(2.0, 2.0 * (1 + 1.1e-12)) ||
      !check_norm(0.0, 0.0) || check_norm(0.0, 1e-300)) return 1;
  const double v[4] = {3.0, 0.0, -4.0, 0.0};
  if (host_nrm2(4, v) != 5.0) return 2;
  ttg::MultiIndex<2> key(3, 1);
  expect_norm("TRSM", key, "tile_mk", 5.0, 5.0);
  if (argc > 1) expect_norm("TRSM", key, "tile_mk", 5.0, 5.5);
  std::puts("ok");
  return 0;
}
"""


def test_check_norm_host(tmp_path):
    from ttg_amd import build as B

    src = tmp_path / "t.cc"
    src.write_text(PROG)
    exe = str(tmp_path / "t")
    subprocess.run([B.CXX] + B.HOST_FLAGS + B.INCLUDES + ["-o", exe, str(src)] + B.HOST_LIBS, check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)
    r = subprocess.run([exe, "bad"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "TRSM({3, 1}): tile_mk has norm 5.5, its producer sent 5" in r.stderr, r.stderr
