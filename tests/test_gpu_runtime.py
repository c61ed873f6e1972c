"""The runtime's data semantics on small task graphs (bin/runtime_checks,
built from ttg_amd/examples/runtime/runtime_checks.cc): a broadcast tile
read by one device task and written by another (with and without
set_defer_writer, parsec/ttg.h:958-1060), a host router that never reads
its tile (no push-out), a host writer whose edit must reach the next device
task, kernels launched after a wait, and ttg::cancel stopping a chain and a
fan-out cleanly."""
import subprocess

import pytest

pytestmark = pytest.mark.gpu

CASES = ["reader_writer_defer", "reader_writer_copy", "host_route", "cancel_chain", "cancel_fanout", "fanout_host"]


@pytest.mark.parametrize("threads", [1, 4])
def test_runtime_checks(threads):
    """threads = 1 runs host tasks inline; 4 runs them on three workers
    whose sends the scheduling thread delivers."""
    from ttg_amd.build import RUNTIME_CHECKS

    print(RUNTIME_CHECKS, "-c", threads, flush=True)
    r = subprocess.run([RUNTIME_CHECKS, "-c", str(threads)], capture_output=True, text=True, timeout=100)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert f"threads {threads}" in r.stdout
    for c in CASES:
        assert f"{c}: PASS" in r.stdout, c
    assert r.returncode == 0
