"""Benchmark: Cholesky solve TFLOP/s to 1e-12 backward error at N=65536 on
one MI355X (tiled POTRF/TRSM/SYRK/GEMM graph of TTG examples/potrf on the
gfx950 HIP kernels of libttg_chol).

A step is one full solve of A X = B (16 right-hand sides) from the FP64
matrix resident in HBM: working copy or low-precision conversion, tiled
factorization, triangular solves and (mixed precision) FP64 iterative
refinement, until the normwise backward error max_j ||b_j - A x_j||_inf /
(||A||_inf ||x_j||_inf + ||b_j||_inf) is <= 1e-12.  Credited work per step:
N^3/3 + N^2/2 + N/6 + 2 N^2 nrhs (refinement iterations are not credited).

Prints one JSON line (see the keys below).  `--gpus N` with N > 1 exits: the
kernels run on one GPU.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # B/s
FP64_PEAK = 78.6e12        # dense FP64 MFMA, flop/s
FP32_PEAK = 157.3e12
FP16_PEAK = 2.5e15         # dense FP16 MFMA


def flops_potrf(n):
    # examples/potrf/potrf.h:31-33 (DPLASMA)
    return n * (((1.0 / 6.0) * n + 0.5) * n + (1.0 / 3.0)) + n * (((1.0 / 6.0) * n) * n - (1.0 / 6.0))


def sigma2_bint():
    # variance of k/127 for k uniform on {-127..127}
    return sum(k * k for k in range(-127, 128)) / 255.0 / (127.0 * 127.0)


SIGMA2_UNIFORM = 1.0 / 3.0  # variance of B uniform on [-1, 1)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--preset", choices=["metric", "low-precision", "fp64"], default="metric",
                   help="metric: N=65536, 16 RHS, kappa~1e4 (the headline); low-precision: N=98304, 8 RHS, "
                        "kappa~1e6; fp64: plain FP64 factor + solve at N=32768 (explicit flags override)")
    p.add_argument("--n", type=int, default=None)
    p.add_argument("--nb", type=int, default=512)
    p.add_argument("--nrhs", type=int, default=None)
    p.add_argument("--kappa", type=float, default=None)
    p.add_argument("--mode", choices=["auto", "f16x2", "fp32", "fp64", "f16", "bf16"], default=None,
                   help="auto: the autotuner picks the fastest of f16x2 + refinement, f16 + refinement and f16 + "
                        "CG that reaches --tol (the metric's default); f16x2: FP32 factor with split-FP16 MFMA "
                        "trailing updates; f16 / bf16: one FP16 / BF16 product per update; fp32: FP32 MFMA "
                        "updates; fp64: plain FP64 factor")
    p.add_argument("--refine", choices=["ir", "cg"], default=None,
                   help="FP64 refinement of a fixed --mode: ir (iterative refinement, the default) or cg "
                        "(conjugate gradients preconditioned by the low-precision factor)")
    p.add_argument("--shift", type=float, default=0.0,
                   help="cg: factor A with its diagonal scaled by 1 + shift (lets an ill-conditioned A's "
                        "low-precision factorization finish)")
    p.add_argument("--max-iter", type=int, default=30)
    p.add_argument("--tol", type=float, default=1e-12)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--b-dist", choices=["uniform", "int"], default="uniform",
                   help="B of A = B B^T/N + alpha I: FP64 uniform on [-1,1) (the configurations' construction) or "
                        "k/127 with k a uniform integer in [-127,127] (exact products, the test suite's)")
    p.add_argument("--cpu-n", type=int, default=4096, help="order of the CPU baseline (N=4096, NB=256: cfg1)")
    p.add_argument("--cpu-nb", type=int, default=256)
    p.add_argument("--cpu-plgsy-n", type=int, default=16384,
                   help="order of the secondary CPU sample on the reference's PLGSY matrix (0: skip)")
    p.add_argument("--no-check", action="store_true",
                   help="skip the factor backward errors (||A - L L^T|| of the mixed and FP64 factors)")
    p.add_argument("--no-tune", action="store_true",
                   help="skip the schedule autotuner (the factorization then runs the library's default knobs)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-profile", action="store_true")
    p.add_argument("--no-fp64", action="store_true", help="skip the plain-FP64 comparison solve")
    p.add_argument("--no-graph", action="store_true",
                   help="skip the task-graph driver run (testing_dpotrf at the workload's N, FP64)")
    p.add_argument("--batched", action="store_true",
                   help="the batched workload instead: --batch-count FP32 SPD systems B B^T/n + I, n uniform in "
                        "[32, 1024], one right-hand side each, factor + solve in one launch (autotuned)")
    p.add_argument("--batch-count", type=int, default=4096)
    p.add_argument("--poinv", action="store_true",
                   help="POTRI (TRTRI + LAUUM) of an FP64 Cholesky factor at the preset's N (65536) unless --n")
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="only run a plain device copy of MIB MiB three times (calibrates the HBM byte counters) "
                        "and exit")
    a = p.parse_args()
    preset = {"metric": (65536, 16, 1e4, "auto"), "low-precision": (98304, 8, 1e6, "f16x2"),
              "fp64": (32768, 16, 1e4, "fp64")}[a.preset]
    a.n = preset[0] if a.n is None else a.n
    a.nrhs = preset[1] if a.nrhs is None else a.nrhs
    a.kappa = preset[2] if a.kappa is None else a.kappa
    a.mode = preset[3] if a.mode is None else a.mode
    return a


METRIC = "Cholesky solve TFLOP/s to 1e-12 backward err, N=65536: mixed-prec vs FP64"


def metric_name(a):
    if (a.n, a.nrhs, a.kappa) == (65536, 16, 1e4) and a.mode != "fp64":
        return METRIC
    what = {"f16x2": "mixed-prec (split-FP16 updates) vs FP64", "fp32": "mixed-prec (FP32 updates) vs FP64",
            "fp64": "plain FP64", "f16": "mixed-prec (FP16 updates) vs FP64",
            "bf16": "mixed-prec (BF16 updates) vs FP64"}[a.mode]
    return f"Cholesky solve TFLOP/s to {a.tol:g} backward err, N={a.n}, {a.nrhs} RHS, kappa~{a.kappa:g}: {what}"


def copy_probe(mib):
    """A plain copy kernel of known size: what FETCH_SIZE / WRITE_SIZE read for
    it gives the factor that turns those counters into HBM bytes."""
    import torch

    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 3}), flush=True)


def batched_flops(sizes):
    return sum(flops_potrf(float(n)) + 2.0 * float(n) * float(n) for n in sizes)


def cpu_baseline_batched(sizes, seed, budget_s=10.0):
    """The oracle's FP64 tile factorization plus a triangular solve per
    matrix, one thread, over a prefix of the batch bounded to ~budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    rng = np.random.default_rng(seed)
    done, work, t_all = 0, 0.0, 0.0
    for n in sizes:
        n = int(n)
        B = rng.uniform(-1, 1, (n, n))
        A = np.asfortranarray(B @ B.T / n + np.eye(n))
        b = rng.uniform(-1, 1, n)
        t0 = time.perf_counter()
        if O.lib().oracle_dpotrf_tile(n, O.ptr(A), n) != 0:
            raise RuntimeError("oracle batched baseline: not SPD")
        y = b.copy()
        L = np.tril(A)
        for i in range(n):  # forward and backward substitution, as the oracle's tile TRSM does
            y[i] = (y[i] - L[i, :i] @ y[:i]) / L[i, i]
        for i in range(n - 1, -1, -1):
            y[i] = (y[i] - L[i + 1:, i] @ y[i + 1:]) / L[i, i]
        t_all += time.perf_counter() - t0
        work += flops_potrf(float(n)) + 2.0 * n * n
        done += 1
        if t_all > budget_s:
            break
    return {"value": work / t_all / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP64 tile factor + substitution solves of the first {done} matrices of the batch "
                      f"(sizes as drawn), one thread, {t_all:.2f} s"}


def run_batched(a):
    import numpy as np
    import torch

    import ttg_amd.autotune as T
    import ttg_amd.chol as C

    rng = np.random.default_rng(a.seed)
    sizes = [int(n) for n in rng.integers(32, 1025, a.batch_count)]
    A, b = T.synthetic_batch(sizes, seed=a.seed)
    cfg, tuned = T.tune_batched(sizes, A, b, reps=2, warmup=1)
    A0, b0 = A.clone(), b.clone()
    info = torch.zeros(len(sizes), dtype=torch.int32, device="cuda")
    batch = C.Batch.packed(sizes, A, b)
    ts = []
    for i in range(a.warmup + a.steps):
        A.copy_(A0)
        b.copy_(b0)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        C.sposv_batched(batch, info, cfg=cfg)
        e1.record()
        torch.cuda.synchronize()
        if i >= a.warmup:
            ts.append(e0.elapsed_time(e1))
    if int((info != 0).sum().item()) != 0:
        raise RuntimeError("batched: a matrix was not SPD")
    # solution check on a sample: backward error of every 64th system (FP64 on the host)
    worst, oa, ob = 0.0, 0, 0
    A0h, b0h, xh = A0.cpu().numpy(), b0.cpu().numpy(), b.cpu().numpy()
    for i, n in enumerate(sizes):
        if i % 64 == 0:
            M = A0h[oa:oa + n * n].reshape(n, n).astype(np.float64)
            x = xh[ob:ob + n].astype(np.float64)
            r = b0h[ob:ob + n].astype(np.float64) - M @ x
            worst = max(worst, np.abs(r).max() / (np.abs(M).sum(1).max() * np.abs(x).max() + np.abs(b0h[ob:ob + n]).max()))
        oa += n * n
        ob += n
    ms = sum(ts) / len(ts)
    F = batched_flops(sizes)
    value = F / (ms * 1e-3) / 1e12
    waves, wgs = C.batched_configs()[cfg]
    out = {
        "metric": f"Batched FP32 SPD factor+solve TFLOP/s ({len(sizes)} matrices, n uniform in [32,1024], 1 RHS)",
        "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: B B^T/n + I, B uniform[-1,1] (seeded), 1 RHS uniform[-1,1] per matrix",
        "config": {"workload": f"{len(sizes)} SPD systems, n ~ U[32,1024], one launch", "count": len(sizes),
                   "sum_n3_over_3": sum(n ** 3 for n in sizes) / 3.0, "launch_config": cfg,
                   "waves_per_workgroup": waves, "workgroups_per_cu": wgs, "parallelism": "single GPU"},
        "autotune_ms": {str(k): round(v, 3) for k, v in tuned.items()},
        "backward_error_sampled": worst,
        "roofline": {"bound": "mfma", "achieved": round(value, 3), "peak": 157.3, "unit": "TFLOP/s",
                     "frac": round(value / 157.3, 4), "traffic": None,
                     "kernel": "posv_batched_kernel (one persistent launch, FP32 MFMA)",
                     "launches": 1, "avg_launch_ms": round(ms, 4), "flops_per_launch": F},
        "cpu_baseline": None if a.no_cpu else cpu_baseline_batched(sizes, a.seed),
    }
    print(json.dumps(out), flush=True)


def flops_dpotri(n):
    """FLOPS_DPOTRI of examples/potrf/potri.h:13-15 (DPLASMA)."""
    return n * ((2. / 3.) + n * ((1. / 3.) * n + 1.)) + n * ((1. / 6.) + n * ((1. / 3.) * n - 0.5))


def flops_dtrtri(n):
    """FLOPS_DTRTRI of examples/potrf/trtri_L.h:13-15 (DPLASMA)."""
    return n * (n * (1. / 6. * n + 0.5) + 1. / 3.) + n * (n * (1. / 6. * n - 0.5) + 1. / 3.)


def cpu_baseline_poinv(n=2048, nb=256):
    """The oracle's POTRI (TRTRI + LAUUM over tiles in the reference graphs'
    task order), one thread, on a bounded SPD sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    t = O.spd_tiles(n, nb, 1.0, 7)
    if O.potrf_tiled(n, nb, t) != 0:
        raise RuntimeError("oracle poinv baseline: not SPD")
    t0 = time.perf_counter()
    if O.potri_tiled(n, nb, t) != 0:
        raise RuntimeError("oracle poinv baseline: singular factor")
    dt = time.perf_counter() - t0
    return {"value": flops_dpotri(float(n)) / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": f"oracle POTRI of an SPD N={n}, NB={nb} factor, one thread, {dt:.2f} s"}


def run_poinv(a):
    import torch

    import ttg_amd.chol as C

    N, NB = a.n, a.nb
    t = C.alloc_tiles(N, NB)
    work = torch.empty(N * N, dtype=torch.float64, device="cuda")
    C.spd_tiled(t, N, NB, 1.0, a.seed, work)
    del work
    A0 = t.clone()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    C.dpotrf_tiled(t, N, NB, info)
    torch.cuda.synchronize()
    if int(info.item()) != 0:
        raise RuntimeError("poinv bench: matrix not SPD")
    L0 = t.clone()

    def timed(fn):
        t.copy_(L0)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    ts = []
    for i in range(a.warmup + a.steps):
        ms = timed(lambda: C.dpotri_tiled(t, N, NB, info))
        if i >= a.warmup:
            ts.append(ms)
    if int(info.item()) != 0:
        raise RuntimeError("poinv bench: singular factor")
    # A^{-1} (A v) = v for a random v, through the tiled symmetric products
    v = torch.rand(N, 1, dtype=torch.float64, device="cuda") * 2 - 1
    zero = torch.zeros_like(v)
    y, z = torch.empty_like(v), torch.empty_like(v)
    C.dresidual_tiled(A0, N, NB, v, zero, y)   # y = -A v
    C.dresidual_tiled(t, N, NB, y, zero, z)    # z = A^{-1} A v
    err = ((z - v).abs().max() / v.abs().max()).item()
    ms_trtri = timed(lambda: C.dtrtri_tiled(t, N, NB, info))
    ms_lauum = timed(lambda: C.dlauum_tiled(t, N, NB))
    ms = sum(ts) / len(ts)
    F = flops_dpotri(float(N))
    value = F / (ms * 1e-3) / 1e12
    out = {
        "metric": f"POTRI FP64 TFLOP/s (FLOPS_DPOTRI, N={N}, NB={NB})",
        "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic: Cholesky factor of B B^T/N + I (B from the seeded integer generator)",
        "config": {"workload": f"POTRI (TRTRI + LAUUM) of an FP64 factor, N={N}, NB={NB}, resident in HBM",
                   "N": N, "NB": NB, "parallelism": "single GPU"},
        "trtri_ms": round(ms_trtri, 3), "lauum_ms": round(ms_lauum, 3),
        "trtri_tflops": round(flops_dtrtri(float(N)) / (ms_trtri * 1e-3) / 1e12, 3),
        "lauum_tflops": round((F - flops_dtrtri(float(N))) / (ms_lauum * 1e-3) / 1e12, 3),
        "inverse_check_rel_err": err,
        "roofline": {"bound": "mfma", "achieved": round(value, 3), "peak": 78.6, "unit": "TFLOP/s",
                     "frac": round(value / 78.6, 4), "traffic": None,
                     "kernel": "whole POTRI (update_kernel<double> GEMMs, TRSMs, tile passes)"},
        "cpu_baseline": None if a.no_cpu else cpu_baseline_poinv(),
    }
    print(json.dumps(out), flush=True)


def read_traffic(mode, N, NB, nrhs):
    """HBM bytes per launch of the dominant kernel from the committed PMC
    summary for this workload (profiles/pmc_<kernel>_N<N>.txt), or None."""
    kern = {"f16x2": "update_f16x2", "fp32": "update_fp32", "fp64": "update_fp64", "f16": "update_f16",
            "bf16": "update_bf16"}[mode]
    path = os.path.join(ROOT, "profiles", f"pmc_{kern}_N{N}_NB{NB}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def graph_driver(n, nb):
    """The FP64 factorization through the task graph and the reference's own
    driver line (ttg_amd/bin/testing_dpotrf -N n -t nb -x -n 2: device PLGSY,
    factorization and write-backs timed, the second run reported), in a child
    process.  Reported beside the metric, not part of it."""
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ttg_amd", "bin", "testing_dpotrf")
    if not os.path.exists(exe):
        return {"error": "ttg_amd/bin/testing_dpotrf is not built"}
    args = [exe, "-N", str(n), "-t", str(nb), "-x", "-n", "2"]
    try:
        r = subprocess.run(args, capture_output=True, text=True, timeout=240)
    except subprocess.TimeoutExpired:
        return {"error": "timed out"}
    lines = [l for l in r.stdout.splitlines() if l.startswith("TTG Execution Time (milliseconds) : ")]
    if r.returncode != 0 or not lines:
        return {"error": f"exit {r.returncode}"}
    f = lines[-1].split(":")
    ms, gf = float(f[1]), float(f[2].split()[-2])
    return {"command": " ".join(["ttg_amd/bin/testing_dpotrf"] + args[1:]), "ms": ms, "tflops": round(gf / 1e3, 3),
            "line": lines[-1]}


def cpu_baseline(n, nb, nrhs, plgsy_n, seed, niter=3):
    """The oracle (C restatement of the examples/potrf graph: the same keys,
    priorities and tile updates, on a priority thread pool over every host
    core, tile kernels in plain C) on cfg1: N=4096, NB=256, A = B B^T/N + I
    with B FP64 uniform on [-1, 1), 16 right-hand sides, niter runs, the
    median reported.  Per run: generation of A (the reference driver times
    its generator with the factorization, testing_dpotrf.cc:145-161), factor,
    solve.  Plus a secondary sample on the reference's own PLGSY matrix."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import statistics

    import numpy as np
    import oracle as O

    cores = len(os.sched_getaffinity(0))
    F = flops_potrf(n)
    rng = np.random.default_rng(seed)
    B = np.asfortranarray(rng.uniform(-1, 1, (n, nrhs)))
    runs, lines = [], []
    for _ in range(niter):
        tg = time.perf_counter()
        tiles = O.spd_uniform_tiles(n, nb, 1.0, seed)
        t0 = time.perf_counter()
        info = O.lib().oracle_dpotrf_tiled(n, nb, O.ptr(tiles), cores)
        t1 = time.perf_counter()
        X = B.copy(order="F")
        O.lib().oracle_dpotrs_tiled(n, nb, O.ptr(tiles), nrhs, O.ptr(X), n)
        t2 = time.perf_counter()
        if info != 0:
            raise RuntimeError(f"oracle factorization failed, info={info}")
        runs.append((t0 - tg, t1 - t0, t2 - t1))
        ms = (t1 - tg) * 1e3
        # testing_dpotrf.cc:160-161 prints this line per iteration
        lines.append(f"TTG Execution Time (milliseconds) : {ms:g} : Flops {F:g} {F / 1e9 / (ms / 1e3):g} GF/s")
    gen = statistics.median(r[0] for r in runs)
    fac = statistics.median(r[1] for r in runs)
    sol = statistics.median(r[1] + r[2] for r in runs)
    work = F + 2.0 * n * n * nrhs
    out = {"value": work / sol / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
           "sample": f"cfg1: oracle FP64 tiled factor + {nrhs}-RHS solve of A = B B^T/N + I, N={n}, NB={nb}, "
                     f"B FP64 uniform[-1,1) (seed {seed}), {cores} threads, median of {niter} runs",
           "backend": "oracle C tile kernels (register-blocked C loops, gcc -O3 -march=x86-64-v3) on a "
                      "priority-ordered pthread pool; no vendor BLAS",
           "niter": niter, "N": n, "NB": nb, "nrhs": nrhs,
           "factor_ms_median": round(fac * 1e3, 3), "factor_gflops": round(F / fac / 1e9, 2),
           "generation_ms_median": round(gen * 1e3, 3),
           "factor_gflops_with_generation": round(F / (fac + gen) / 1e9, 2),
           "reference_lines": lines,
           "backward_error": O.solve_backward_error(n, nb, O.spd_uniform_tiles(n, nb, 1.0, seed), X, B)}
    for l in lines:
        print(l, file=sys.stderr, flush=True)
    if plgsy_n:
        tg = time.perf_counter()
        t = O.plgsy_tiles(plgsy_n, nb, plgsy_n, 3872)
        t0 = time.perf_counter()
        info = O.lib().oracle_dpotrf_tiled(plgsy_n, nb, O.ptr(t), cores)
        t1 = time.perf_counter()
        if info != 0:
            raise RuntimeError(f"oracle PLGSY factorization failed, info={info}")
        Fp = flops_potrf(plgsy_n)
        out["plgsy_sample"] = {
            "workload": f"PLGSY N={plgsy_n} NB={nb} (bump=N, seed=3872, the reference driver's matrix), factor",
            "factor_s": round(t1 - t0, 3), "generation_s": round(t0 - tg, 3),
            "factor_gflops": round(Fp / (t1 - t0) / 1e9, 2),
            "factor_gflops_with_generation": round(Fp / (t1 - tg) / 1e9, 2)}
    return out


def factor_check(C, A0, L, N, NB):
    """||A - L L^T||_inf / (||A||_inf N eps) of testing_dpotrf.cc:277-335,
    on the device in tiles (L L^T formed in FP64)."""
    a, r, q = C.dcheck_dpotrf_tiled(A0, L, N, NB)
    return {"anorm_inf": a, "residual_inf": r, "rel": r / a, "scaled": q, "passes_reference_bound_60": q <= 60.0}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the kernels and drivers here are "
              f"single-GPU", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.batched:
        run_batched(a)
        return
    if a.poinv:
        run_poinv(a)
        return
    import torch

    import ttg_amd.chol as C

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    N, NB, nrhs = a.n, a.nb, a.nrhs
    stream = torch.cuda.current_stream()

    # ---- inputs, resident in HBM before the timed region
    s2 = SIGMA2_UNIFORM if a.b_dist == "uniform" else sigma2_bint()
    lam_max = 4.0 * s2  # Marchenko-Pastur edge of B B^T / N for square B
    alpha = lam_max / (a.kappa - 1.0)
    A0 = C.alloc_tiles(N, NB)
    work = torch.empty(N * N, dtype=torch.float64, device=dev)
    (C.spd_uniform_tiled if a.b_dist == "uniform" else C.spd_tiled)(A0, N, NB, alpha, a.seed, work)
    del work
    torch.cuda.empty_cache()
    g = torch.Generator(device="cpu").manual_seed(a.seed)
    B = (torch.rand(nrhs, N, generator=g, dtype=torch.float64) * 2 - 1).to(dev).t()  # N x nrhs column-major
    X = torch.empty_like(B.t()).t()
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    R = torch.empty_like(X)

    prec = {"f16x2": C.PREC_F16X2, "fp32": C.PREC_FP32, "fp64": C.PREC_FP64, "f16": C.PREC_F16, "bf16": C.PREC_BF16}
    iters = {}

    refine = {}  # mode -> "ir" / "cg"

    def make_step(mode):
        def step():
            if refine.get(mode) == "cg":
                it, be = C.dposv_krylov(A0, N, NB, B, X, prec=prec[mode], shift=a.shift, tol=a.tol,
                                        max_iter=max(a.max_iter, 200), info=info)
            else:
                it, be = C.dposv_mixed(A0, N, NB, B, X, prec=prec[mode], tol=a.tol, max_iter=a.max_iter,
                                       info=info)
            iters[mode] = (it, be)
            if not (be <= a.tol) or int(info.item()) != 0:
                raise RuntimeError(f"{mode}/{refine.get(mode, 'ir')}: backward error {be}, info {int(info.item())}")
        return step

    # ---- method: the autotuner's choice of factor precision and refinement
    # (each candidate run once to build its plans, then timed once; the
    # fastest that reaches the tolerance wins)
    method = None
    if a.mode == "auto":
        import ttg_amd.autotune as T

        def candidate(mode, how):
            refine[mode] = how
            return make_step(mode)

        choice, mtrials = T.tune_method([("f16x2", "ir"), ("f16", "ir"), ("f16", "cg")], candidate)
        a.mode = choice[0]
        refine[a.mode] = choice[1]
        method = {"choice": {"mode": choice[0], "refine": choice[1]}, "trials": mtrials}
    else:
        refine[a.mode] = a.refine or "ir"

    def backward_error():
        return C.backward_error(A0, N, NB, X, B)

    def timed(mode, steps, warmup):
        step = make_step(mode)
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        if int(info.item()) != 0:
            raise RuntimeError(f"{mode}: factorization failed: info={int(info.item())}")
        t0 = time.perf_counter()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        berr = backward_error()
        if int(info.item()) != 0 or not (berr <= a.tol):
            raise RuntimeError(f"{mode}: solve failed: info={int(info.item())} backward error={berr}")
        return ev0.elapsed_time(ev1) / steps, wall, berr

    # ---- schedule: the autotuner's choice of the factorization knobs for
    # this workload (coordinate descent from the defaults, before the timed
    # region; cached per device and workload)
    schedule = None
    if not a.no_tune and a.mode != "fp64":
        import ttg_amd.autotune as T

        default = {k: C.get_knob(k) for k in C.KNOBS}
        choice, trials = T.tune_schedule(make_step(a.mode), N, NB, prec[a.mode], nrhs, use_cache=False)
        schedule = {"choice": choice, "default": default, "trials": trials}

    ms_step, wall, berr = timed(a.mode, a.steps, a.warmup)
    it_main = iters[a.mode][0]
    checks = {}
    L64 = None
    if not a.no_check:
        L64 = C.alloc_tiles(N, NB)
        C.dposv_mixed_factor(L64, N, NB, nrhs, prec[a.mode])
        checks[a.mode] = factor_check(C, A0, L64, N, NB)
    X_main = X.clone()
    other = "fp64" if a.mode != "fp64" and not a.no_fp64 else None
    ms_fp64 = berr_fp64 = None
    x_diff = factor_diff = None
    if other:
        ms_fp64, _, berr_fp64 = timed(other, 1, 1)
        # the refined solution against the plain-FP64 one, per right-hand side
        x_diff = float(((X_main - X).abs().amax(dim=0) / X.abs().amax(dim=0)).max().item())
        if L64 is not None:
            L64d = C.alloc_tiles(N, NB)
            C.dposv_mixed_factor(L64d, N, NB, nrhs, prec[other])
            checks[other] = factor_check(C, A0, L64d, N, NB)
            # the strict upper parts of the diagonal tiles hold the same input
            # values in both factors and add nothing to the difference
            factor_diff = float(((L64 - L64d).abs().max() / L64d.abs().max()).item())
            del L64d
    del L64
    torch.cuda.empty_cache()
    F = flops_potrf(N) + 2.0 * N * N * nrhs
    value = F / (ms_step * 1e-3) / 1e12

    # ---- roofline of the dominant kernel (the batched SYRK+GEMM update of the
    # factorization run by this mode), from one extra solve after the timed
    # region with every launch timed by HIP events on its own stream
    roof = None
    if not a.no_profile:
        for k in (C.PROF_UPDATE, C.PROF_TRSM, C.PROF_POTRF):
            C.profile_read(k)
        C.profile_enable(True)
        try:
            make_step(a.mode)()
            torch.cuda.synchronize()
        finally:
            C.profile_enable(False)
        ms_u, n_u, fl_u = C.profile_read(C.PROF_UPDATE)
        ms_t, n_t, fl_t = C.profile_read(C.PROF_TRSM)
        ms_p, n_p, fl_p = C.profile_read(C.PROF_POTRF)
        # the published dense peak of the unit the kernel runs on; the split
        # FP16 kernel issues three MFMA products per credited flop, so its
        # MFMA pipe utilisation is 3x its fraction of the FP16 peak
        peak = {"f16x2": FP16_PEAK, "fp32": FP32_PEAK, "fp64": FP64_PEAK, "f16": FP16_PEAK,
                "bf16": FP16_PEAK}[a.mode]
        products = 3.0 if a.mode == "f16x2" else 1.0
        ach = fl_u / (ms_u * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": round(peak / 1e12, 1), "unit": "TFLOP/s",
                "frac": round(ach / (peak / 1e12), 4), "traffic": read_traffic(a.mode, N, NB, nrhs),
                "mfma_pipe_util": round(products * ach / (peak / 1e12), 4),
                "kernel": {"f16x2": "update_f16x2_kernel (batched SYRK+GEMM, 3 FP16 MFMA products per output; "
                                    "achieved counts each output flop once against the FP16 dense peak)",
                           "fp32": "update_kernel<float> (batched SYRK+GEMM trailing update)",
                           "fp64": "update_kernel<double> (batched SYRK+GEMM trailing update)",
                           "f16": "update_f16x2_kernel<FP16, one product> (batched SYRK+GEMM)",
                           "bf16": "update_f16x2_kernel<BF16, one product> (batched SYRK+GEMM)"}[a.mode],
                "launches": n_u, "avg_launch_ms": round(ms_u / max(n_u, 1), 4),
                "flops_per_launch": fl_u / max(n_u, 1),
                "others": {"trsm_ms": round(ms_t, 2), "trsm_tflops": round(fl_t / max(ms_t, 1e-9) / 1e9, 2),
                           "potrf_ms": round(ms_p, 2), "potrf_avg_ms": round(ms_p / max(n_p, 1), 4)}}

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(a.cpu_n, a.cpu_nb, nrhs, a.cpu_plgsy_n, a.seed)
    graph = None if a.no_graph else graph_driver(N, NB)

    out = {
        "metric": metric_name(a),
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f16x2": "f16x2+f32+f64", "fp32": "f32+f64", "fp64": "f64", "f16": "f16+f32+f64",
                  "bf16": "bf16+f32+f64"}[a.mode],
        "data": ("synthetic: A = B B^T/N + alpha I, B FP64 uniform[-1,1) (seeded LCG) "
                 if a.b_dist == "uniform" else
                 "synthetic: A = B B^T/N + alpha I, B(i,j) = k/127 with k uniform integer in [-127,127] (seeded LCG) ")
                + f"alpha = (4 sigma^2)/(kappa-1) for kappa ~ {a.kappa:g}; {nrhs} RHS uniform[-1,1]",
        "config": {"workload": f"SPD solve N={N} NB={NB} nrhs={nrhs} kappa~{a.kappa:g} (tiled POTRF graph)",
                   "N": N, "NB": NB, "nrhs": nrhs, "mode": a.mode, "parallelism": "single GPU"},
        "backward_error": berr,
        "refinement_steps": it_main,
        "solve_tflops_frac": {"fp64_peak": round(value / (FP64_PEAK / 1e12), 4),
                              "fp32_peak": round(value / (FP32_PEAK / 1e12), 4),
                              "fp16_dense_peak": round(value / (FP16_PEAK / 1e12), 4)},
        "factor_backward_error": checks or None,
        "x_rel_diff_vs_fp64": x_diff,
        "factor_rel_diff_vs_fp64": factor_diff,
        "fp64_ms_per_solve": None if ms_fp64 is None else round(ms_fp64, 3),
        "fp64_backward_error": berr_fp64,
        "speedup_vs_fp64": None if ms_fp64 is None else round(ms_fp64 / ms_step, 3),
        "wall_s": round(wall, 3),
        "method": method or {"choice": {"mode": a.mode, "refine": refine.get(a.mode, "ir")}},
        "schedule": schedule,
        "roofline": roof,
        "cpu_baseline": cpu,
        "graph_driver": graph,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
