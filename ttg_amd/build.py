"""Build libttg_chol.so (HIP, gfx950) in-tree with hipcc.

Usage:  python -m ttg_amd.build   (or ttg_amd.build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libttg_chol.so")

SOURCES = [
    "chol_capi.hip",
    "chol_kernels.hip",
    "chol_gen.hip",
    "chol_tiled.hip",
    "chol_solve.hip",
    "chol_profile.hip",
    "chol_mixed.hip",
    "chol_f16.hip",
    "chol_batched.hip",
    "chol_poinv.hip",
    "chol_sweep.hip",
    "chol_check.hip", "chol_tileops.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    # packed FP32 forms are kept off (wrong results seen with some of them)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


EXAMPLES = os.path.join(HERE, "examples", "potrf")
INCLUDES = ["-I" + os.path.join(HERE, "include"), "-I" + os.path.join(os.path.dirname(HERE), "include"),
            "-I" + EXAMPLES]
GRAPH_LIB = os.path.join(HERE, "libttg_graph.so")
DRIVER = os.path.join(HERE, "bin", "testing_dpotrf")
# the same driver with the tile-norm consistency checks (DEBUG_TILES_VALUES)
DEBUG_DRIVER = os.path.join(HERE, "bin", "testing_dpotrf_debug")
POINV_DRIVER = os.path.join(HERE, "bin", "testing_dpoinv")
# device-task overhead benchmark (a chain of dependent device tasks)
CHAIN_BENCH = os.path.join(HERE, "bin", "chain_ttg_dev")
RUNTIME_CHECKS = os.path.join(HERE, "bin", "runtime_checks")
# host-only C++ (no device code): clang++ with the HIP runtime headers
CXX = os.environ.get("TTG_CXX", "/opt/rocm/llvm/bin/clang++")
HOST_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
HOST_LIBS = ["-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("compiler failed building " + what)


def build_graph(verbose=False):
    """The TTG-API runtime examples: libttg_graph.so, bin/testing_dpotrf and
    its DEBUG_TILES_VALUES build bin/testing_dpotrf_debug, the POINV driver
    bin/testing_dpoinv and the device-task chain benchmark bin/chain_ttg_dev
    (host C++ over libttg_chol.so)."""
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    link = ["-L" + HERE, "-lttg_chol", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/.."]
    _run([CXX] + HOST_FLAGS + ["-shared"] + INCLUDES + ["-o", GRAPH_LIB + ".tmp",
         os.path.join(EXAMPLES, "graph_capi.cc")] + link + HOST_LIBS, "libttg_graph.so", verbose)
    os.replace(GRAPH_LIB + ".tmp", GRAPH_LIB)
    _run([CXX] + HOST_FLAGS + INCLUDES + ["-o", DRIVER + ".tmp", os.path.join(EXAMPLES, "testing_dpotrf.cc")] + link +
         HOST_LIBS, "testing_dpotrf", verbose)
    os.replace(DRIVER + ".tmp", DRIVER)
    _run([CXX] + HOST_FLAGS + ["-DDEBUG_TILES_VALUES"] + INCLUDES +
         ["-o", DEBUG_DRIVER + ".tmp", os.path.join(EXAMPLES, "testing_dpotrf.cc")] + link + HOST_LIBS,
         "testing_dpotrf_debug", verbose)
    os.replace(DEBUG_DRIVER + ".tmp", DEBUG_DRIVER)
    _run([CXX] + HOST_FLAGS + INCLUDES + ["-o", POINV_DRIVER + ".tmp", os.path.join(EXAMPLES, "testing_dpoinv.cc")] +
         link + HOST_LIBS, "testing_dpoinv", verbose)
    os.replace(POINV_DRIVER + ".tmp", POINV_DRIVER)
    _run([CXX] + HOST_FLAGS + INCLUDES + ["-o", CHAIN_BENCH + ".tmp",
         os.path.join(HERE, "examples", "task-benchmarks", "chain_dev.cc")] + link + HOST_LIBS, "chain_ttg_dev",
         verbose)
    os.replace(CHAIN_BENCH + ".tmp", CHAIN_BENCH)
    _run([CXX] + HOST_FLAGS + INCLUDES + ["-o", RUNTIME_CHECKS + ".tmp",
         os.path.join(HERE, "examples", "runtime", "runtime_checks.cc")] + link + HOST_LIBS, "runtime_checks", verbose)
    os.replace(RUNTIME_CHECKS + ".tmp", RUNTIME_CHECKS)


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libttg_chol.so")
    os.replace(tmp, LIB)
    build_graph(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
