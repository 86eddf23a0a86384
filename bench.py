"""bench.py -- FP32 GEMM M=N=K=8192 on one MI355X through numc_matmul_ex.

Workload (the headline metric's configuration): A, B ~ N(0,1) float32 from the
documented seeded generator, resident in HBM before timing.  One step = one
complete GEMM call (pack kernels + MFMA kernel) of the split-BF16 method;
the plain FP32 and the block-scaled FP8 (e4m3, a power-of-two scale per 32 k
of each row of A and column of B) methods are timed the same way and
reported beside it.  value = 2*M*N*K / ms_per_step (FP32-equivalent
TFLOP/s; extra MFMA products are not counted as work).

Configuration: the one numc_matmul runs for the shape (the committed
exact-shape table numc_amd/data/tuned.csv, from exhaustive autotune).

Errors: max componentwise |C - R| / (|A||B|) over all 67M entries (also the
normwise max|C - R| / max|R| and the plain relative |C - R| / |R| over the
entries with |R| > 1e-3 max|R|), with R and |A||B| from the library's own FP64 MFMA GEMM (validated against the CPU oracle
by the test suite), for the N(0,1) inputs and an ill-scaled set (rows of A and
columns of B scaled by 2^u, u integer uniform in [-20, 20]).

Warm-up: each method first runs for 0.3 s untimed (the clocks ramp up under
a heavy MFMA load), then the W warmup steps, then the K timed steps.

roofline: the MFMA kernel of the split-BF16 method; its average duration comes
from HIP events the library records on its own stream around each launch.
cpu_baseline: numc's own numc_matmul (oracle/_ref, built from numc's sources,
AVX2 + OpenMP) on the same bytes, min of 2 after a warmup, on this host.

configs: the other workloads, each with its own roofline and cpu_baseline --
float32 512^3, float64 2048^3, int32 2048^3 and int8 4096^3 through
numc_matmul (whole call, plus the library's per-launch kernel times), one
GPU's share of an 8-way row split of 8192^3 (1024 x 8192 x 8192, each
method), and the uneven batch of 4096 problems (M, N, K log-uniform in
[16, 1024]) through
numc_matmul_batched in float32, split-BF16 and int8 -> int32, after
numc_matmul_batched_tune picks its schedule.  host_cpu names the host's CPU
model, cores and ISA flags.  --no-configs leaves them out; --only configs or
--only batched runs just those.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GEMM TFLOPS/%peak: split-BF16 vs plain FP32 at ≤2× FP32 err vs FP64 oracle"
PEAK = {"f32": 157.3, "split_bf16": 2516.0 / 3.0, "fp8": 5030.0, "f64": 78.6}   # TFLOP/s, MI355X dense
HBM_PEAK = 8000.0  # GB/s
WARM_S = 0.3       # untimed clock warm-up per method, before the W warmup steps


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_method(nc, ctx, A, B, C, method, family, steps, warmup, n):
    # the configuration numc_matmul runs for this shape: the committed
    # exact-shape table (numc_amd/data/tuned.csv, made by exhaustive autotune
    # with `python -m numc_amd.autotune pin`), else the learned selector
    cfg = nc.lib().numc_selector_choice(family, n, n, n, 0)
    if cfg < 0:
        nc.check(cfg, "numc_selector_choice")
    # bring the GPU clocks to their loaded state first: the first ~30 ms of a
    # heavy MFMA load run up to 15% slower while the clocks ramp
    t_end = time.perf_counter() + WARM_S
    while time.perf_counter() < t_end:
        nc.check(nc.matmul(A, B, C, method), "clock warm-up")
        ctx.synchronize()
    for _ in range(warmup):
        nc.check(nc.matmul(A, B, C, method), "warmup")
    ctx.synchronize()
    ctx.profile(True)
    ctx.profile_read(1, family)
    ctx.profile_read(0, family)
    ctx.mark(0)
    t0 = time.perf_counter()
    for _ in range(steps):
        nc.check(nc.matmul(A, B, C, method), "step")
    ctx.mark(1)
    ctx.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    region = ctx.elapsed_ms(0, 1)
    gemm_ms, gemm_n = ctx.profile_read(1, family)
    pack_ms, pack_n = ctx.profile_read(0, family)
    ctx.profile(False)
    return {"config": nc.config_get(family, cfg), "ms_per_step": region / steps, "wall_ms_per_step": wall / steps,
            "kernel_ms": gemm_ms / max(gemm_n, 1), "kernel_launches": gemm_n,
            "kernel_ms_per_step": gemm_ms / steps,
            "pack_ms": pack_ms / max(pack_n, 1)}


def errors_vs_fp64(nc, ctx, a, b, outs, n):
    """componentwise max error of each output against the FP64 MFMA reference"""
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    res = {}
    R = ctx.array((n, n), nc.FLOAT64)
    S = ctx.array((n, n), nc.FLOAT64)
    for x, y, dst in ((a64, b64, R), (np.abs(a64), np.abs(b64), S)):
        X, Y = ctx.from_numpy(x), ctx.from_numpy(y)
        nc.check(nc.matmul(X, Y, dst), "fp64 reference")
    r, s = R.numpy(), S.numpy()
    ar = np.abs(r)
    tau = 1e-3 * float(np.max(ar))
    big = (ar > tau) & (ar > 0)
    for name, c in outs.items():
        d = np.abs(c.astype(np.float64) - r)
        with np.errstate(divide="ignore", invalid="ignore"):
            q = np.where(s > 0, d / s, np.where(d > 0, np.inf, 0.0))
        res[name] = {"componentwise": float(np.max(q)), "normwise": float(np.max(d) / np.max(ar)),
                     # plain relative error |C - R| / |R| over the entries with |R| > 1e-3 max|R|
                     "elementwise_rel": float(np.max(d[big] / ar[big]))}
    return res


def cpu_baseline(n, seed, cores):
    """numc's own AVX2 numc_matmul on the same bytes (oracle/_ref), bounded sample."""
    import oracle
    if not os.path.exists(oracle.REF_BENCH):
        return None, "oracle/_ref not built"
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([oracle.REF_BENCH, "8", str(n), str(n), str(n), "n", str(seed), "1", "2"],
                       capture_output=True, text=True, env=env, timeout=600)
    if r.returncode != 0:
        return None, "numc_ref_bench failed: %s" % r.stderr[-200:]
    return json.loads(r.stdout.strip().splitlines()[-1]), None


def batch_shapes(count, seed):
    """(M, N, K) of the uneven batch: each a log-uniform integer in [16, 1024]"""
    rng = np.random.default_rng(seed)
    return np.clip(np.round(np.exp(rng.uniform(np.log(16), np.log(1024), (count, 3)))), 16, 1024).astype(int)


def host_cpu():
    """model, ISA and core counts of this host (where cpu_baseline runs)"""
    info = {"model": None, "logical_cpus": os.cpu_count(), "physical_cores": None,
            "isa_flags": None, "numc_build": "x86-64-v3 (AVX2 + FMA), -mno-avx512f, OpenMP"}
    flags = set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and info["model"] is None:
                info["model"] = line.split(":", 1)[1].strip()
            elif line.startswith("flags") and not flags:
                flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    info["isa_flags"] = sorted(f for f in flags if f in ("avx", "avx2", "fma", "avx512f", "avx512bw", "avx512vnni",
                                                          "avx_vnni", "amx_tile"))
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = dict(l.split(":", 1) for l in out.splitlines() if ":" in l)
        cps, sockets = kv.get("Core(s) per socket"), kv.get("Socket(s)")
        if cps and sockets:
            info["physical_cores"] = int(cps) * int(sockets)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def _work_unit(unit):
    return "op" if unit.startswith("TOP") else "flop"


def roofline_entry(work, nbytes, kernel_ms, peak, unit, kernel):
    """roofline of one kernel: bound = whichever of the MFMA peak and HBM takes
    longer for the algorithmic work and bytes; achieved = that quantity / the
    kernel's average duration"""
    t_mfma, t_hbm = work / (peak * 1e12), nbytes / (HBM_PEAK * 1e9)
    if t_mfma >= t_hbm:
        ach = work / kernel_ms * 1e-9
        return {"bound": "mfma", "achieved": round(ach, 3), "peak": peak, "unit": unit,
                "frac": round(ach / peak, 4), "traffic": None, "kernel": kernel,
                "algorithmic": "%.4g %s per launch (%.4g B compulsory)" % (work, _work_unit(unit), nbytes)}
    ach = nbytes / kernel_ms * 1e-6
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK, 4), "traffic": None, "kernel": kernel,
            "algorithmic": "%.4g B compulsory per launch (%.4g %s)" % (nbytes, work, _work_unit(unit))}


def run_ref_bench(args, cores, stdin=None, timeout=300):
    import oracle
    if not os.path.exists(oracle.REF_BENCH):
        return None, "oracle/_ref not built"
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    try:
        r = subprocess.run([oracle.REF_BENCH] + [str(x) for x in args], input=stdin, capture_output=True, text=True,
                           env=env, timeout=timeout)
    except subprocess.TimeoutExpired:
        return None, "numc_ref_bench timed out"
    if r.returncode != 0:
        return None, "numc_ref_bench failed (%d): %s" % (r.returncode, r.stderr[-200:])
    return json.loads(r.stdout.strip().splitlines()[-1]), None


def time_whole_and_kernels(nc, ctx, fn, family, steps, warmup, batched=False):
    """whole-call ms (HIP events on the context stream, no per-kernel events),
    then the same calls with per-kernel events: average pack / GEMM / split-K
    reduce (or batched GEMM) durations"""
    t_end = time.perf_counter() + WARM_S
    while time.perf_counter() < t_end:
        nc.check(fn(), "clock warm-up")
        ctx.synchronize()
    for _ in range(warmup):
        nc.check(fn(), "warmup")
    ctx.synchronize()
    ctx.mark(0)
    for _ in range(steps):
        nc.check(fn(), "step")
    ctx.mark(1)
    ctx.synchronize()
    whole = ctx.elapsed_ms(0, 1) / steps
    ctx.profile(True)
    for cat in (0, 1, 2, 3):
        ctx.profile_read(cat, family)
    for _ in range(steps):
        nc.check(fn(), "profiled step")
    ctx.synchronize()
    k = {}
    for cat, name in ((0, "pack"), (1, "gemm"), (2, "reduce"), (3, "batched_gemm")):
        ms, n = ctx.profile_read(cat, family)
        k[name] = (ms / n if n else 0.0, n / steps)
    ctx.profile(False)
    main = k["batched_gemm"] if batched else k["gemm"]
    kernel_ms = main[0]
    if not batched and main[1] == 1 and k["pack"][1] == 0 and k["reduce"][1] == 0:
        # a one-launch call: its kernel takes no longer than the whole call timed
        # without per-kernel events (the events add their own gap to a short kernel)
        kernel_ms = min(kernel_ms, whole)
    return {"ms_per_call": whole, "kernel_ms": kernel_ms, "kernel_launches_per_call": main[1],
            "pack_ms": k["pack"][0], "pack_launches_per_call": k["pack"][1],
            "reduce_ms": k["reduce"][0], "reduce_launches_per_call": k["reduce"][1]}


# the single-problem configurations (numc_matmul on HBM-resident operands):
# (key, dtype, numc dtype code, n, dist, ref-bench dist, lo, hi, seed, family, peak, unit, bytes/elem in, out)
SINGLE_CONFIGS = [
    ("c1_f32_512", "FLOAT32", 8, 512, "DIST_UNIFORM", "u", -1, 1, 1, "FAMILY_F32", 157.3, "TFLOP/s", 4, 4),
    ("c2_f64_2048", "FLOAT64", 9, 2048, "DIST_NORMAL", "n", 0, 0, 2, "FAMILY_F64", 78.6, "TFLOP/s", 8, 8),
    ("c2_i32_2048", "INT32", 2, 2048, "DIST_INT", "i", -100, 100, 2, "FAMILY_I8", 5030.0, "TOP/s", 4, 4),
    ("c2_i8_4096", "INT8", 0, 4096, "DIST_INT", "i", -128, 127, 2, "FAMILY_I8", 5030.0, "TOP/s", 1, 1),
]


def bench_configs(nc, steps, warmup, cores, with_cpu):
    """The other configurations of the metric, each with its whole-call rate,
    its GEMM kernel's roofline and numc's own CPU numc_matmul on the same bytes."""
    res = {}
    for key, dtn, code, n, dist, rdist, lo, hi, seed, famn, peak, unit, ein, eout in SINGLE_CONFIGS:
        dt, fam = getattr(nc, dtn), getattr(nc, famn)
        ctx = nc.Ctx()
        ctx.set_sync(False)
        A, B, Cm = ctx.array((n, n), dt), ctx.array((n, n), dt), ctx.array((n, n), dt, zeros=True)
        A.fill_random(getattr(nc, dist), seed, 0, lo, hi)   # host fill: the bytes numc's CPU run gets
        B.fill_random(getattr(nc, dist), seed, 1, lo, hi)
        sums = (str(wsum(A.numpy())), str(wsum(B.numpy())))
        for x in (A, B, Cm):
            x.to_device()
        small = n <= 1024
        st = max(steps, 200) if small else max(steps, 20)
        t = time_whole_and_kernels(nc, ctx, lambda: nc.matmul(A, B, Cm), fam, st, warmup)
        cfg = nc.lib().numc_selector_choice(fam, n, n, n, 0)
        work, nbytes = 2.0 * n ** 3, float(n * n * (2 * ein + eout))
        e = {"workload": "%s %d^3, numc_matmul, operands resident in HBM" % (dtn.lower(), n),
             "steps": st, "ms_per_call": round(t["ms_per_call"], 5),
             "rate_whole_call": round(work / t["ms_per_call"] * 1e-9, 3), "unit": unit,
             "kernel_ms": round(t["kernel_ms"], 5), "kernel_launches_per_call": t["kernel_launches_per_call"],
             "pack_ms": round(t["pack_ms"], 5), "pack_launches_per_call": t["pack_launches_per_call"],
             "reduce_ms": round(t["reduce_ms"], 5), "reduce_launches_per_call": t["reduce_launches_per_call"],
             "config": nc.config_get(fam, cfg) if cfg >= 0 else None,
             "roofline": roofline_entry(work, nbytes, t["kernel_ms"], peak, unit,
                                        "%s GEMM kernel (avg of %d launches)" % (famn[7:].lower(), st))}
        log("bench: %s %.4f ms/call (kernel %.4f ms, roofline frac %.3f)" %
            (key, t["ms_per_call"], t["kernel_ms"], e["roofline"]["frac"]))
        ctx.close()
        if with_cpu:
            reps = 50 if small else (3 if n >= 4096 else 5)
            r, why = run_ref_bench([code, n, n, n, rdist, seed, 1, reps] + ([lo, hi] if rdist != "n" else []), cores)
            if r:
                same = (r["a_wsum"], r["b_wsum"]) == sums
                e["cpu_baseline"] = {"value": round(r["gflops"] * 1e-3, 5), "unit": unit, "cores": r["threads"],
                                     "kind": "reference", "ms": r["ms_min"],
                                     "sample": "numc_matmul %s %d^3 (numc AVX2+OpenMP build), min of %d after 1 "
                                               "warmup, same input bytes: %s%s" %
                                               (dtn.lower(), n, reps, same,
                                                "; numc's packed int8 path saturates per 512-k block, so its "
                                                "outputs differ (times comparable)" if code == 0 else "")}
            else:
                e["cpu_baseline"] = {"value": None, "unit": unit, "cores": cores, "kind": "reference", "sample": why}
        res[key] = e
    return res


def bench_shard(nc, steps, warmup, m=1024, n=8192):
    """One GPU's share of the 8192^3 FP32 GEMM split 8 ways by rows of C:
    A's 1024-row panel times the whole B, N(0,1) inputs, each method; whole
    call and kernel, with the kernel's roofline (the multi-GPU exchange is not
    part of it)."""
    ctx = nc.Ctx()
    ctx.set_sync(False)
    A, B = ctx.array((m, n), nc.FLOAT32), ctx.array((n, n), nc.FLOAT32)
    A.fill_random(nc.DIST_NORMAL, 3, 0)
    B.fill_random(nc.DIST_NORMAL, 3, 1)
    Cm = ctx.array((m, n), nc.FLOAT32, zeros=True)
    for x in (A, B, Cm):
        x.to_device()
    work = 2.0 * m * n * n
    res = {"workload": "FP32 GEMM %dx%dx%d (one GPU's rows of C in an 8-way row split of %d^3), N(0,1), "
                       "operands resident in HBM" % (m, n, n, n)}
    for name, method, fam, peak in (("split_bf16", nc.METHOD_SPLIT_BF16, nc.FAMILY_SPLIT_BF16, PEAK["split_bf16"]),
                                    ("f32", nc.METHOD_PLAIN, nc.FAMILY_F32, PEAK["f32"]),
                                    ("fp8", nc.METHOD_FP8_SCALED, nc.FAMILY_FP8, PEAK["fp8"])):
        t = time_whole_and_kernels(nc, ctx, lambda: nc.matmul(A, B, Cm, method), fam, steps, warmup)
        cfg = nc.lib().numc_selector_choice(fam, m, n, n, 0)
        res[name] = {"ms_per_call": round(t["ms_per_call"], 4),
                     "tflops_whole_call": round(work / t["ms_per_call"] * 1e-9, 2),
                     "kernel_ms": round(t["kernel_ms"], 4), "pack_ms": round(t["pack_ms"], 4),
                     "config": nc.config_get(fam, cfg) if cfg >= 0 else None,
                     "roofline": roofline_entry(work, float(4 * (m * n + n * n + m * n)), t["kernel_ms"], peak,
                                                "TFLOP/s", "%s GEMM kernel (avg of %d launches)" % (name, steps))}
        log("bench: shard %dx%dx%d %s %.4f ms/call (kernel %.4f ms, %.1f TF whole call)" %
            (m, n, n, name, t["ms_per_call"], t["kernel_ms"], work / t["ms_per_call"] * 1e-9))
    ctx.close()
    return res


def bench_batched(nc, count, seed, reps, steps=10, warmup=2, cores=16, with_cpu=True, schedule=None, separate=True):
    """4096 problems, M, N, K log-uniform integers in [16, 1024], U[-1,1) values
    (int8: integers in [-128, 127]): one persistent graph-replayed launch vs one
    numc_matmul_ex call per problem, the batched GEMM kernel's roofline, and
    numc's CPU numc_matmul over the same problems (numc has no batched GEMM)."""
    dims = batch_shapes(count, seed)
    ctx = nc.Ctx()
    ctx.set_sync(False)
    M_, N_, K_ = dims[:, 0].astype(float), dims[:, 1].astype(float), dims[:, 2].astype(float)
    flop = float(np.sum(2.0 * M_ * N_ * K_))
    res = {"problems": count, "flop": flop}
    cpu = {}
    if with_cpu:
        shapes = "\n".join("%d %d %d" % tuple(d) for d in dims)
        for code, lo, hi in ((8, -1, 1), (0, -128, 127)):
            r, why = run_ref_bench(["batch", code, seed, 1, 2, lo, hi], cores, stdin=shapes)
            cpu[code] = (r, why)
    for label, dt, odt, method, flags, fam, peak, unit, ein, eout in (
            ("f32", nc.FLOAT32, nc.FLOAT32, nc.METHOD_PLAIN, 0, nc.FAMILY_F32, 157.3, "TFLOP/s", 4, 4),
            ("split_bf16", nc.FLOAT32, nc.FLOAT32, nc.METHOD_SPLIT_BF16, 0, nc.FAMILY_SPLIT_BF16, 2516.0 / 3,
             "TFLOP/s", 4, 4),
            ("int8_to_int32", nc.INT8, nc.INT32, nc.METHOD_AUTO, nc.OPT_WIDEN, nc.FAMILY_I8, 5030.0, "TOP/s", 1, 4)):
        As, Bs, Os = [], [], []
        in_sum = 0
        for i, (M, N, K) in enumerate(dims):
            A = ctx.array((M, K), dt)
            B = ctx.array((K, N), dt)
            if dt == nc.INT8:
                A.fill_random(nc.DIST_INT, seed, 2 * i, -128, 127)
                B.fill_random(nc.DIST_INT, seed, 2 * i + 1, -128, 127)
            else:
                A.fill_random(nc.DIST_UNIFORM, seed, 2 * i, -1, 1)
                B.fill_random(nc.DIST_UNIFORM, seed, 2 * i + 1, -1, 1)
            if with_cpu:
                in_sum += int(A.numpy().view(np.uint8).sum(dtype=np.uint64)) + \
                    int(B.numpy().view(np.uint8).sum(dtype=np.uint64))
            O = ctx.array((M, N), odt, zeros=True)
            for x in (A, B, O):
                x.to_device()
            As.append(A)
            Bs.append(B)
            Os.append(O)
        # autotuned schedule: the kernel (direct or packed) and the large-tile
        # threshold, measured on this batch
        if schedule is None:
            thresh = nc.matmul_batched_tune(ctx, As, Bs, Os, method, flags, 3)
            if thresh < 0:
                nc.check(thresh, "numc_matmul_batched_tune")
        else:   # a fixed schedule (profiling runs): the direct kernel where the family has one
            want_direct = schedule[0] and fam in (nc.FAMILY_F32, nc.FAMILY_SPLIT_BF16, nc.FAMILY_I8)
            nc.check(nc.lib().numc_matmul_batched_set_schedule(fam, int(want_direct), schedule[1]), "schedule")
            thresh = schedule[1]
        H = nc.BatchHandles(As, Bs, Os)   # the handle arrays, built once as a C caller keeps them
        nc.check(nc.matmul_batched(ctx, H, method=method, flags=flags), "batched warmup")  # builds the graph
        ctx.synchronize()
        # whole calls (graph replays) after the clock warm-up, then the same
        # calls with per-kernel events
        t = time_whole_and_kernels(nc, ctx, lambda: nc.matmul_batched(ctx, H, method=method, flags=flags), fam,
                                   max(reps, steps), 2, batched=True)
        t_batched = t["ms_per_call"]
        # one call per problem, same stream, no graph
        t_sep = float("nan")
        if separate:
            for i in range(count):
                nc.check(nc.matmul(As[i], Bs[i], Os[i], method, -1, flags), "separate warmup")
            ctx.synchronize()
            ctx.mark(4)
            for i in range(count):
                nc.check(nc.matmul(As[i], Bs[i], Os[i], method, -1, flags), "separate")
            ctx.mark(5)
            t_sep = ctx.elapsed_ms(4, 5)
        nbytes = float(np.sum(ein * (M_ * K_ + K_ * N_) + eout * M_ * N_))
        direct = t["pack_launches_per_call"] == 0   # what ran: the direct kernel packs nothing
        e = {"ms_batched": round(t_batched, 4), "tflops_batched": round(flop / t_batched * 1e-9, 2),
             "tile_threshold": thresh, "kernel": "direct" if direct else "packed",
             "ms_separate_launches": round(t_sep, 3),
             "tflops_separate": round(flop / t_sep * 1e-9, 2),
             "speedup": round(t_sep / t_batched, 2),
             "kernel_ms": round(t["kernel_ms"], 4), "pack_ms": round(t["pack_ms"], 4),
             "pack_launches_per_call": t["pack_launches_per_call"],
             "roofline": roofline_entry(flop, nbytes, t["kernel_ms"], peak, unit,
                                        "persistent batched GEMM kernel, %s operands (avg of %d launches)" %
                                        ("direct" if direct else "packed", max(reps, steps)))}
        if with_cpu:
            r, why = cpu[0 if dt == nc.INT8 else 8]
            if r:
                e["cpu_baseline"] = {
                    "value": round(r["gflops"] * 1e-3, 5), "unit": unit, "cores": r["threads"], "kind": "reference",
                    "ms": r["ms_min"],
                    "sample": "numc_matmul per problem over the %d problems (%s; numc AVX2+OpenMP build), min of 2 "
                              "passes after 1, same input bytes: %s%s" %
                              (count, "float32" if dt != nc.INT8 else "int8 in, int8 out",
                               r["in_byte_sum"] == str(in_sum),
                               "; numc has no split-BF16, its float32 GEMM is the CPU method" if label == "split_bf16"
                               else "; numc has no int8->int32 GEMM and its packed int8 path saturates, so times are "
                                    "comparable, outputs are not" if dt == nc.INT8 else "")}
            else:
                e["cpu_baseline"] = {"value": None, "unit": unit, "cores": cores, "kind": "reference", "sample": why}
        res[label] = e
        log("bench: batched %s: %.3f ms (%.1f T/s), kernel %.3f ms, pack %.3f ms, vs %d separate calls %.2f ms" %
            (label, t_batched, flop / t_batched * 1e-9, t["kernel_ms"], t["pack_ms"], count, t_sep))
        del As, Bs, Os
    ctx.close()
    return res


def bench_suite(nc, steps, warmup, seed):
    """The other reference configurations timed the same way (one JSON line
    each): float32 512^3, float64 2048^3, int32 2048^3 in [-100,100], int8
    4096^3 full range; operands resident in HBM, one step = one numc_matmul."""
    ctx = nc.Ctx()
    ctx.set_sync(False)
    peaks = {nc.FLOAT32: ("f32", 157.3, nc.FAMILY_F32), nc.FLOAT64: ("f64", 78.6, nc.FAMILY_F64),
             nc.INT32: ("i32", 5030.0, nc.FAMILY_I8), nc.INT8: ("i8", 5030.0, nc.FAMILY_I8)}
    for dt, n, dist, lo, hi in ((nc.FLOAT32, 512, nc.DIST_UNIFORM, -1, 1), (nc.FLOAT64, 2048, nc.DIST_NORMAL, 0, 0),
                                (nc.INT32, 2048, nc.DIST_INT, -100, 100), (nc.INT8, 4096, nc.DIST_INT, -128, 127)):
        A, B, Cm = ctx.array((n, n), dt), ctx.array((n, n), dt), ctx.array((n, n), dt, zeros=True)
        for x in (A, B, Cm):
            x.to_device()
        A.fill_random(dist, seed, 0, lo, hi)   # device-current arrays fill on the GPU
        B.fill_random(dist, seed, 1, lo, hi)
        name, peak, fam = peaks[dt]
        t = time_method(nc, ctx, A, B, Cm, nc.METHOD_AUTO, fam, steps, warmup, n)
        flops = 2.0 * n ** 3
        print(json.dumps({"suite": name, "n": n, "ms_per_step": round(t["ms_per_step"], 5),
                          "t_per_s": round(flops / t["ms_per_step"] * 1e-9, 2),
                          "gemm_kernel_ms_per_step": round(t["kernel_ms_per_step"], 5),
                          "gemm_launches_per_step": t["kernel_launches"] / steps,
                          "kernel_t_per_s": round(flops / t["kernel_ms_per_step"] * 1e-9, 2)
                          if t["kernel_ms_per_step"] else None,
                          "peak": peak, "config": t["config"]}), flush=True)
    bench_halves(nc, ctx, steps, warmup, seed)
    bench_consumers(nc, ctx, steps, warmup, seed)
    ctx.close()


def bench_halves(nc, ctx, steps, warmup, seed):
    """bfloat16 and float16 arrays, 8192^3 N(0,1), resident in HBM; one step =
    one numc_matmul (pack + one-product MFMA kernel + 16-bit store)"""
    n = 8192
    for dt, name, fam in ((nc.BFLOAT16, "bf16", nc.FAMILY_BF16), (nc.FLOAT16, "f16", nc.FAMILY_F16)):
        A, B, Cm = ctx.array((n, n), dt), ctx.array((n, n), dt), ctx.array((n, n), dt, zeros=True)
        for x in (A, B, Cm):
            x.to_device()
        A.fill_random(nc.DIST_NORMAL, seed, 30, 0, 0)
        B.fill_random(nc.DIST_NORMAL, seed, 31, 0, 0)
        t = time_method(nc, ctx, A, B, Cm, nc.METHOD_AUTO, fam, steps, warmup, n)
        flops = 2.0 * n ** 3
        print(json.dumps({"suite": name, "n": n, "ms_per_step": round(t["ms_per_step"], 5),
                          "t_per_s": round(flops / t["ms_per_step"] * 1e-9, 2),
                          "gemm_kernel_ms_per_step": round(t["kernel_ms_per_step"], 5),
                          "kernel_t_per_s": round(flops / t["kernel_ms_per_step"] * 1e-9, 2)
                          if t["kernel_ms_per_step"] else None,
                          "pack_ms": round(t["pack_ms"], 5), "peak": 2516.0,
                          "config": t["config"]}), flush=True)
        del A, B, Cm


def time_calls(nc, ctx, fn, steps, warmup):
    """ms per call of fn() on the context stream (HIP events), after warm-up"""
    for _ in range(warmup):
        nc.check(fn(), "warmup")
    ctx.synchronize()
    ctx.mark(0)
    for _ in range(steps):
        nc.check(fn(), "step")
    ctx.mark(1)
    ctx.synchronize()
    return ctx.elapsed_ms(0, 1) / steps


def bench_consumers(nc, ctx, steps, warmup, seed):
    """numc_linear (a dense layer, x 8192x4096 @ w 4096x4096 + bias, float32)
    and numc_bmm (attention scores, (8,16,1024,64) @ (8,16,64,1024), float32):
    one step = one call, operands resident in HBM."""
    M, K, N = 8192, 4096, 4096
    x, w, bias, y = ctx.array((M, K), nc.FLOAT32), ctx.array((K, N), nc.FLOAT32), ctx.array((N,), nc.FLOAT32), \
        ctx.array((M, N), nc.FLOAT32, zeros=True)
    for a in (x, w, bias, y):
        a.to_device()
    for i, a in enumerate((x, w, bias)):
        a.fill_random(nc.DIST_NORMAL, seed, 10 + i, 0, 0)
    ms = time_calls(nc, ctx, lambda: nc.linear(x, w, bias, y), steps, warmup)
    flops = 2.0 * M * N * K
    print(json.dumps({"suite": "linear_f32", "shape": [M, K, N], "ms_per_step": round(ms, 5),
                      "t_per_s": round(flops / ms * 1e-9, 2), "peak": 157.3,
                      "frac": round(flops / ms * 1e-9 / 157.3, 4)}), flush=True)
    B, H, S, D = 8, 16, 1024, 64
    q, kt, s = ctx.array((B, H, S, D), nc.FLOAT32), ctx.array((B, H, D, S), nc.FLOAT32), \
        ctx.array((B, H, S, S), nc.FLOAT32, zeros=True)
    for a in (q, kt, s):
        a.to_device()
    q.fill_random(nc.DIST_NORMAL, seed, 20, 0, 0)
    kt.fill_random(nc.DIST_NORMAL, seed, 21, 0, 0)
    ms = time_calls(nc, ctx, lambda: nc.bmm(q, kt, s), steps, warmup)
    flops = 2.0 * B * H * S * S * D
    out_bytes = 4.0 * B * H * S * S
    print(json.dumps({"suite": "bmm_f32", "shape": [B, H, S, D], "ms_per_step": round(ms, 5),
                      "t_per_s": round(flops / ms * 1e-9, 2), "peak": 157.3,
                      "out_GBps": round(out_bytes / ms * 1e-6, 1),
                      "note": "K = 64: each 128x128 tile is two k-chunks, so copy latency and the tile store are not hidden"}), flush=True)


def read_traffic():
    """HBM bytes per launch of the split-BF16 kernel from the committed rocprof
    summary (FETCH_SIZE/WRITE_SIZE runs scaled by a plain copy of known size)."""
    path = os.path.join(ROOT, "profiles", "gemm_split_bf16_8192_traffic.txt")
    try:
        txt = open(path).read()
        d = json.loads(txt[txt.index("{"):])
        return {"hbm_bytes_per_launch": round(d["hbm_bytes_per_launch"]),
                "source": "profiles/gemm_split_bf16_8192_traffic.txt (%s)" % d["kernel"]}
    except (OSError, ValueError, KeyError):
        return {}


def wsum(x):
    w = np.ascontiguousarray(x).view(np.uint32).ravel().astype(np.uint64)
    return int(np.sum(w * np.arange(1, w.size + 1, dtype=np.uint64)))


def bench_vendor(steps, warmup, seed):
    """The vendor GEMMs that torch dispatches to (hipBLASLt, rocBLAS) on the shapes timed here, so the
    roofline fractions can be read against what the vendor library itself reaches on this GPU: bf16,
    fp16 and fp32 8192^3 (torch.matmul), int8 -> int32 4096^3 (torch._int_mm) and fp8 e4m3 8192^3
    with per-tensor scales (torch._scaled_mm).  Not part of the library; torch is only the caller."""
    import torch
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    res = {}

    def timed(name, fn, flop):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        res[name] = {"ms": round(ms, 5), "tflops": round(flop / ms * 1e-9, 2)}
        log("vendor: %s %.4f ms, %.1f TFLOP/s" % (name, ms, flop / ms * 1e-9))

    n = 8192
    for name, dt in (("bf16_8192", torch.bfloat16), ("f16_8192", torch.float16), ("f32_8192", torch.float32)):
        a = torch.randn(n, n, device=dev, dtype=dt, generator=g)
        b = torch.randn(n, n, device=dev, dtype=dt, generator=g)
        timed(name, lambda: torch.matmul(a, b), 2.0 * n ** 3)
        del a, b
    m = 4096
    try:
        a = torch.randint(-128, 128, (m, m), device=dev, dtype=torch.int8, generator=g)
        b = torch.randint(-128, 128, (m, m), device=dev, dtype=torch.int8, generator=g)
        timed("int8_4096", lambda: torch._int_mm(a, b), 2.0 * m ** 3)
    except (RuntimeError, AttributeError) as e:
        res["int8_4096"] = {"unsupported": str(e).splitlines()[0][:200]}
    try:
        f8 = torch.float8_e4m3fn
        a = (torch.randn(n, n, device=dev, generator=g) * 0.5).to(f8)
        b = (torch.randn(n, n, device=dev, generator=g) * 0.5).to(f8).t().contiguous().t()  # column-major
        one = torch.ones((), device=dev)
        timed("fp8_8192", lambda: torch._scaled_mm(a, b, one, one, out_dtype=torch.float32), 2.0 * n ** 3)
    except (RuntimeError, AttributeError, TypeError) as e:
        res["fp8_8192"] = {"unsupported": str(e).splitlines()[0][:200]}
    print(json.dumps({"vendor": "torch %s (hipBLASLt/rocBLAS)" % torch.__version__, "steps": steps,
                      "warmup": warmup, "results": res}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configurations (f32 512^3, f64/int32 2048^3, int8 4096^3)")
    ap.add_argument("--no-pcie", action="store_true")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--only", default=None,
                    help="time one method: split_bf16, f32 or fp8; 'batched' for the uneven batch alone; "
                         "'configs' for the other configurations alone")
    ap.add_argument("--suite", action="store_true",
                    help="time the other configurations (f32 512^3, f64 2048^3, int32 2048^3, int8 4096^3, "
                         "bf16/f16 8192^3, numc_linear, numc_bmm)")
    ap.add_argument("--no-separate", action="store_true",
                    help="batched runs: skip the one-call-per-problem comparison (profiling runs)")
    ap.add_argument("--batch-schedule", default=None,
                    help="batched runs: a fixed schedule instead of the tuner's, 'direct:T' or 'packed:T' "
                         "(T = the large-tile threshold; split-BF16 always runs packed)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run the reference copy kernel over this many MiB (HBM counter calibration)")
    ap.add_argument("--vendor", action="store_true",
                    help="only time torch.matmul / torch._int_mm / torch._scaled_mm (hipBLASLt, rocBLAS) at the "
                         "same shapes, as a calibration of what the vendor library reaches on this GPU")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (one context drives "
              "one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.vendor:
        bench_vendor(args.steps, args.warmup, args.seed)
        return
    import numc_amd as nc
    if args.copy_probe:
        ctx = nc.Ctx()
        nbytes = args.copy_probe << 20
        X = ctx.array((nbytes,), nc.UINT8)
        Y = ctx.array((nbytes,), nc.UINT8)
        X.to_device()
        Y.to_device()
        src, dst = nc.lib().numc_array_device_data(X.h), nc.lib().numc_array_device_data(Y.h)
        for _ in range(3):
            nc.check(nc.lib().numc_device_copy(ctx.h, dst, src, nbytes), "numc_device_copy")
        print(json.dumps({"copy_probe_bytes": nbytes, "launches": 3}))
        ctx.close()
        return
    if args.suite:
        bench_suite(nc, args.steps, args.warmup, args.seed)
        return
    cores = min(16, os.cpu_count() or 1)
    sched = None
    if args.batch_schedule:
        kind, t = args.batch_schedule.split(":")
        sched = (kind == "direct", int(t))
    if args.only == "batched":
        log("bench: batched uneven GEMM, %d problems" % args.batch)
        print(json.dumps(bench_batched(nc, args.batch, args.seed + 1, max(1, args.steps // 2), args.steps,
                                       args.warmup, cores, not args.no_cpu, sched, not args.no_separate)),
              flush=True)
        return
    if args.only == "configs":
        configs = bench_configs(nc, args.steps, args.warmup, cores, not args.no_cpu)
        configs["c3_shard8_rows"] = bench_shard(nc, args.steps, args.warmup, m=args.n // 8, n=args.n)
        print(json.dumps({"configs": configs, "host_cpu": host_cpu()}), flush=True)
        return
    n = args.n
    flops = 2.0 * n * n * n
    ctx = nc.Ctx()
    ctx.set_sync(False)
    log("bench: FP32 GEMM %d^3, N(0,1) inputs, seed %d" % (n, args.seed))
    A = ctx.array((n, n), nc.FLOAT32)
    B = ctx.array((n, n), nc.FLOAT32)
    A.fill_random(nc.DIST_NORMAL, args.seed, 0)
    B.fill_random(nc.DIST_NORMAL, args.seed, 1)
    a_host, b_host = A.numpy(), B.numpy()
    A.to_device()
    B.to_device()
    outs = {}
    C = {}
    for name in ("split_bf16", "f32", "fp8"):
        C[name] = ctx.array((n, n), nc.FLOAT32, zeros=True)
        C[name].to_device()

    variants = {}
    plan = [("split_bf16", nc.METHOD_SPLIT_BF16, nc.FAMILY_SPLIT_BF16),
            ("f32", nc.METHOD_PLAIN, nc.FAMILY_F32),
            ("fp8", nc.METHOD_FP8_SCALED, nc.FAMILY_FP8)]
    if args.only:
        plan = [x for x in plan if x[0] == args.only]
    for name, method, fam in plan:
        log("bench: timing %s (%d warmup, %d steps)" % (name, args.warmup, args.steps))
        t = time_method(nc, ctx, A, B, C[name], method, fam, args.steps, args.warmup, n)
        t["tflops"] = flops / t["ms_per_step"] * 1e-9
        t["kernel_tflops"] = flops / t["kernel_ms"] * 1e-9
        t["peak_tflops"] = PEAK[name]
        t["frac_of_peak"] = t["kernel_tflops"] / PEAK[name]
        variants[name] = t
        log("bench: %s %.3f ms/step, %.1f TFLOP/s (kernel %.3f ms, %.1f TFLOP/s)" %
            (name, t["ms_per_step"], t["tflops"], t["kernel_ms"], t["kernel_tflops"]))

    if args.only:
        print(json.dumps({"only": args.only, "n": n, **{k: (float("%.5g" % v) if isinstance(v, float) else v)
                                                          for k, v in variants[args.only].items()}}), flush=True)
        ctx.close()
        return

    errs = None
    if not args.no_errors:
        log("bench: errors against the FP64 reference (N(0,1) inputs)")
        for name in C:
            outs[name] = C[name].numpy()
        errs = {"normal": errors_vs_fp64(nc, ctx, a_host, b_host, outs, n)}
        rng = np.random.default_rng(args.seed)
        u = rng.integers(-20, 20, n, endpoint=True)
        v = rng.integers(-20, 20, n, endpoint=True)
        ai = np.ldexp(a_host, u[:, None]).astype(np.float32)
        bi = np.ldexp(b_host, v[None, :]).astype(np.float32)
        AI, BI = ctx.from_numpy(ai), ctx.from_numpy(bi)
        outs_i = {}
        for name, method, fam in plan:
            O = ctx.array((n, n), nc.FLOAT32, zeros=True)
            nc.check(nc.matmul(AI, BI, O, method), "ill-scaled")
            outs_i[name] = O.numpy()
        log("bench: errors against the FP64 reference (ill-scaled inputs)")
        errs["ill_scaled"] = errors_vs_fp64(nc, ctx, ai, bi, outs_i, n)
        for name in variants:
            for kind in ("componentwise", "normwise", "elementwise_rel"):
                variants[name]["err_" + kind] = errs["normal"][name][kind]
                variants[name]["err_" + kind + "_ill_scaled"] = errs["ill_scaled"][name][kind]

    cpu = None
    if not args.no_cpu:
        log("bench: cpu baseline (numc AVX2 numc_matmul, %d threads)" % cores)
        res, why = cpu_baseline(n, args.seed, cores)
        if res is not None:
            same = res["a_wsum"] == str(wsum(a_host)) and res["b_wsum"] == str(wsum(b_host))
            cpu = {"value": res["gflops"] * 1e-3, "unit": "TFLOP/s", "cores": res["threads"],
                   "kind": "reference",
                   "sample": "numc_matmul float32 %d^3 (numc AVX2+OpenMP build), min of %d after %d warmup, "
                             "same input bytes: %s" % (n, res["iters"], res["warmup"], same),
                   "ms": res["ms_min"]}
        else:
            cpu = {"value": None, "unit": "TFLOP/s", "cores": cores, "kind": "reference", "sample": why}

    # numc's default residency: host arrays, each call uploads A and B and
    # downloads C (PCIe-inclusive; reported beside, never as the value)
    pcie = None
    if not args.no_pcie:
        ctx_h = nc.Ctx()
        AH, BH = ctx_h.from_numpy(a_host), ctx_h.from_numpy(b_host)
        CH = ctx_h.array((n, n), nc.FLOAT32, zeros=True)
        nc.check(nc.matmul(AH, BH, CH, nc.METHOD_SPLIT_BF16), "host-resident warmup")
        t0 = time.perf_counter()
        for _ in range(3):
            nc.check(nc.matmul(AH, BH, CH, nc.METHOD_SPLIT_BF16), "host-resident")
        ms = (time.perf_counter() - t0) * 1e3 / 3
        pcie = {"ms_per_call": round(ms, 3), "tflops": round(flops / ms * 1e-9, 2),
                "note": "split-BF16, host-resident arrays (H2D of A and B, D2H of C each call), synchronous"}
        log("bench: host-resident split-BF16 call %.2f ms (%.1f TFLOP/s incl. PCIe)" % (ms, flops / ms * 1e-9))
        ctx_h.close()

    configs = {}
    if not args.no_configs:
        log("bench: the other configurations")
        configs = bench_configs(nc, args.steps, args.warmup, cores, not args.no_cpu)
        log("bench: one GPU's share of an 8-way row split of %d^3" % n)
        configs["c3_shard8_rows"] = bench_shard(nc, args.steps, args.warmup, m=n // 8, n=n)
    if not args.no_batched:
        log("bench: batched uneven GEMM, %d problems" % args.batch)
        configs["c4_batch_%d" % args.batch] = bench_batched(nc, args.batch, args.seed + 1, max(1, args.steps // 2),
                                                            args.steps, args.warmup, cores, not args.no_cpu)

    s = variants["split_bf16"]
    p = variants["f32"]
    traffic = read_traffic()
    out = {
        "metric": METRIC,
        "value": round(s["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "FP32 GEMM M=N=K=%d, A,B ~ N(0,1) (seeded), split-BF16 method "
                               "(hi*hi + hi*lo + lo*hi bf16 MFMA, FP32 accumulate), operands resident in HBM"
                               % n, "M": n, "N": n, "K": n, "method": "split_bf16"},
        "roofline": {"bound": "mfma", "achieved": round(s["kernel_tflops"], 3), "peak": round(PEAK["split_bf16"], 1),
                     "unit": "TFLOP/s", "frac": round(s["kernel_tflops"] / PEAK["split_bf16"], 4),
                     "traffic": traffic.get("hbm_bytes_per_launch"),
                     "traffic_source": traffic.get("source"),
                     "kernel": "split-BF16 MFMA kernel, tile %(bm)dx%(bn)d k%(kc)d waves %(waves_m)dx%(waves_n)d"
                               % s["config"] + " (avg %.4f ms over %d launches)" % (s["kernel_ms"], s["kernel_launches"]),
                     "algorithmic": "2*M*N*K = %.4g FP32-equivalent flop per launch; peak = 2516/3 TFLOP/s "
                                    "(three bf16 products per FP32 product)" % flops},
        "cpu_baseline": cpu,
        "split_vs_plain_speedup": round(s["tflops"] / p["tflops"], 3),
        # the target compares the MFMA kernels: split-BF16 kernel TFLOPS / plain FP32 kernel TFLOPS
        "split_vs_plain_kernel_speedup": round(s["kernel_tflops"] / p["kernel_tflops"], 3),
        "variants": {k: {kk: (float("%.5g" % vv) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                     for k, v in variants.items()},
    }
    if configs:
        out["configs"] = configs
    out["host_cpu"] = host_cpu()
    if pcie:
        out["host_resident_pcie_inclusive"] = pcie
    if errs:
        out["err_ratio_split_vs_plain"] = round(errs["normal"]["split_bf16"]["componentwise"] /
                                                errs["normal"]["f32"]["componentwise"], 4)
        out["err_ratio_split_vs_plain_ill_scaled"] = round(
            errs["ill_scaled"]["split_bf16"]["componentwise"] / errs["ill_scaled"]["f32"]["componentwise"], 4)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
