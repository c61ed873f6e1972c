"""Autotuner for the launch parameters of libttg_chol.

Three tunables, all measured on the GPU that will run the work and cached
per device and workload signature:

* the batched solver's launch configuration (waves per workgroup and
  persistent workgroups per CU, `ttg_chol_batched_config`) for a batch's
  size distribution;
* the tile size NB of the tiled mixed-precision solve for an order N;
* the factorization's schedule knobs (`ttg_chol_set_knob`): panels per
  bulk update, CUs reserved for the diagonal chain, the update's tile-group
  side and the grouped/one-panel schedule, searched by coordinate descent
  from the defaults;
* the solve's method: the precision of the factor's trailing updates and
  the FP64 refinement (iterative refinement or preconditioned CG), the
  fastest that reaches the tolerance on the caller's data.

Each candidate runs `warmup` untimed and `reps` timed times (HIP events on the
calling stream); the median wins.  Results go to a JSON cache
($TTG_AMD_TUNING, default ~/.cache/ttg_amd/tuning.json) keyed by the device
name and the workload signature, so later runs reuse them without measuring.
"""
import ctypes
import json
import os
import statistics

_CACHE_ENV = "TTG_AMD_TUNING"


def _cache_path():
    return os.environ.get(_CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "ttg_amd", "tuning.json")


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    data = _load()
    data[key] = value
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: tuning still returns its result


def _device_name():
    import torch

    return torch.cuda.get_device_name(0)


def batch_signature(sizes):
    """Size distribution of a batch: count and a histogram in 64-wide bins."""
    hist = [0] * 16
    for n in sizes:
        hist[min((int(n) - 1) // 64, 15)] += 1
    return f"batched:count={len(sizes)}:hist={','.join(map(str, hist))}"


def _time_events(fn, warmup, reps, reset=None):
    import torch

    ts = []
    for i in range(warmup + reps):
        if reset is not None:
            reset()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        if i >= warmup:
            ts.append(e0.elapsed_time(e1))
    return statistics.median(ts)


def synthetic_batch(sizes, seed=0, device="cuda"):
    """FP32 SPD matrices B B^T / n + I (B uniform in [-1, 1]) packed back to
    back (column-major, lda = n) and uniform right-hand sides, on the device.
    The products run on the library's own FP32 update kernel, all matrices in
    one launch group (no vendor BLAS)."""
    import torch

    import ttg_amd.chol as C

    g = torch.Generator(device=device).manual_seed(seed)
    sizes = [int(n) for n in sizes]
    total = sum(n * n for n in sizes)
    Bs = torch.rand(total, generator=g, device=device, dtype=torch.float32) * 2 - 1
    b = torch.rand(sum(sizes), generator=g, device=device, dtype=torch.float32) * 2 - 1
    A = torch.zeros(total, dtype=torch.float32, device=device)
    diag, oa = [], 0
    for n in sizes:  # A := I
        diag.append(torch.arange(n, dtype=torch.int64) * (n + 1) + oa)
        oa += n * n
    A[torch.cat(diag).to(device)] = 1.0
    fp = ctypes.c_float
    lib = C.lib()
    st = torch.cuda.current_stream(device)
    sp = ctypes.c_void_p(st.cuda_stream)
    with C.group(st):
        oa = 0
        for n in sizes:  # A += (1/n) B B^T
            pa, pb = A.data_ptr() + 4 * oa, Bs.data_ptr() + 4 * oa
            C._check(lib.ttg_chol_sgemm_nt(sp, n, n, n, fp(1.0 / n), pb, n, pb, n, fp(1.0), pa, n), "sgemm_nt")
            oa += n * n
    del Bs
    return A, b


def tune_batched(sizes, A=None, b=None, reps=3, warmup=1, use_cache=True):
    """Best launch configuration for a batch of these sizes: (cfg, {cfg: ms}).
    A and b (packed, as synthetic_batch makes them) are used as given or made
    synthetic; their contents are restored before every timed run."""
    import torch

    import ttg_amd.chol as C

    key = f"{_device_name()}|{batch_signature(sizes)}"
    if use_cache:
        hit = _load().get(key)
        if hit is not None:
            return int(hit["cfg"]), {int(k): v for k, v in hit["ms"].items()}
    if A is None:
        A, b = synthetic_batch(sizes)
    A0, b0 = A.clone(), b.clone()
    info = torch.zeros(len(sizes), dtype=torch.int32, device=A.device)
    batch = C.Batch.packed([int(n) for n in sizes], A, b)

    def reset():
        A.copy_(A0)
        b.copy_(b0)

    ms = {}
    for cfg in range(len(C.batched_configs())):
        ms[cfg] = _time_events(lambda: C.sposv_batched(batch, info, cfg=cfg), warmup, reps, reset)
    reset()
    best = min(ms, key=ms.get)
    _store(key, {"cfg": best, "ms": ms, "configs": C.batched_configs()})
    return best, ms


def tune_tiled_nb(N, candidates=(256, 512), prec=None, nrhs=16, reps=1, warmup=1, use_cache=True, seed=7):
    """Best tile size for the mixed-precision solve of order N: (NB, {NB: ms})."""
    import torch

    import ttg_amd.chol as C

    prec = C.PREC_F16X2 if prec is None else prec
    key = f"{_device_name()}|tiled:N={N}:prec={prec}:nrhs={nrhs}"
    if use_cache:
        hit = _load().get(key)
        if hit is not None:
            return int(hit["NB"]), {int(k): v for k, v in hit["ms"].items()}
    ms = {}
    for NB in candidates:
        A = C.alloc_tiles(N, NB)
        work = torch.empty(N * N, dtype=torch.float64, device="cuda")
        C.spd_tiled(A, N, NB, 1e-3, seed, work)
        del work
        B = torch.rand(nrhs, N, dtype=torch.float64, device="cuda").t()
        X = torch.empty_like(B.t()).t()
        ms[NB] = _time_events(lambda: C.dposv_mixed(A, N, NB, B, X, prec=prec), warmup, reps)
        del A, B, X
        torch.cuda.empty_cache()
    best = min(ms, key=ms.get)
    _store(key, {"NB": best, "ms": ms})
    return best, ms


# candidate values per schedule knob, searched in this order
SCHEDULE_SPACE = (
    ("paired", (1, 0)),
    ("group", (3, 2, 4)),
    ("potrf_cus", (8, 0, 16)),
    ("tile_group", (2, 1, 4)),
    # workgroups per CU of the persistent sweeps (each value gives bitwise
    # the same solution; 0, the per-step launches, sums in another order)
    ("sweep", (2, 1, 4)),
)


def schedule_signature(N, NB, prec, nrhs):
    return f"schedule:N={N}:NB={NB}:prec={prec}:nrhs={nrhs}"


def apply_schedule(choice):
    """Set the knobs of `choice` (a dict knob -> value; None clears) and drop
    the cached plans so that the next solve builds its plan with them."""
    import ttg_amd.chol as C

    for k, v in choice.items():
        C.set_knob(k, v)
    C.release_plans()


def tune_schedule(run, N, NB, prec, nrhs, reps=2, warmup=1, min_gain=0.01, use_cache=True):
    """Schedule knobs for the planned solve `run()` (one call = one solve on
    the caller's resident data, of order N, tile NB, precision prec, nrhs
    right-hand sides).  Coordinate descent from the library's defaults: each
    knob in turn takes the candidate with the lowest median time, provided it
    beats the current best by `min_gain` (relative), so a noisy win never
    displaces the default.  The knobs are left set to the choice.  Returns
    (choice, trials) with trials = [{"knobs": ..., "ms": ...}, ...]."""
    import ttg_amd.chol as C

    key = f"{_device_name()}|{schedule_signature(N, NB, prec, nrhs)}"
    if use_cache:
        hit = _load().get(key)
        if hit is not None:
            apply_schedule(hit["choice"])
            return dict(hit["choice"]), hit["trials"]
    apply_schedule({k: None for k in C.KNOBS})
    current = C.knobs()
    trials = []
    seen = {}

    def measure(knobs):
        sig = tuple(sorted(knobs.items()))
        if sig not in seen:
            apply_schedule(knobs)
            seen[sig] = _time_events(run, warmup, reps)
            trials.append({"knobs": dict(knobs), "ms": round(seen[sig], 3)})
        return seen[sig]

    best_ms = measure(current)
    for name, values in SCHEDULE_SPACE:
        if name == "group" and not current["paired"]:
            continue  # one panel per update: the group size is unused
        for v in values:
            if v == current[name]:
                continue
            cand = dict(current, **{name: v})
            ms = measure(cand)
            if ms < best_ms * (1.0 - min_gain):
                current, best_ms = cand, ms
    apply_schedule(current)
    _store(key, {"choice": current, "trials": trials})
    return current, trials


def tune_method(candidates, make_run, warmup=1, reps=1):
    """The fastest solve method: candidates is a list of (mode, refine)
    pairs, make_run(mode, refine) returns a callable that runs one solve
    and raises when it does not reach the tolerance (or the factorization
    fails).  Returns ((mode, refine), trials) with trials = [{"mode",
    "refine", "ms" or "error"}, ...]."""
    trials = []
    best, best_ms = None, None
    for mode, how in candidates:
        run = make_run(mode, how)
        try:
            ms = _time_events(run, warmup, reps)
        except Exception as e:  # a method that fails to converge is no choice
            trials.append({"mode": mode, "refine": how, "error": str(e)[:200]})
            continue
        trials.append({"mode": mode, "refine": how, "ms": round(ms, 3)})
        if best_ms is None or ms < best_ms:
            best, best_ms = (mode, how), ms
    if best is None:
        raise RuntimeError(f"no solve method reached the tolerance: {trials}")
    return best, trials
