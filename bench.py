"""Benchmark: FP32 complex FFT, N = 2^24, forward + inverse on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one forward and one inverse transform of the whole batch
(synthetic inputs uniform in [-1, 1), seed 0, already resident in HBM).
Prints one JSON line: value = GFLOP/s by the 5 N log2 N convention (both
directions counted), the roofline (the transform's algorithmic bytes, 16 N
per direction, over the HIP-event-timed step; every launch of a step is one
pass of the pass kernel over the whole array, the launches back to back on
the plan's stream, and the per-launch rate is a secondary field), and the
CPU baseline (the oracle's FP32 port of FFTJ's threaded four-step, timed on
this host on a bounded sample).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# name -> (n, batch, description); the other configurations are parity cases
WORKLOADS = {
    "c2c_f32_2pow24": (1 << 24, 1, "C2C FP32 N=2^24 batch 1, three 256-point passes, forward+inverse"),
    "c2c_f32_4096x16384": (4096, 16384, "C2C FP32 N=4096 batch 16384, forward+inverse"),
    "c2c_f32_1e6x8": (10**6, 8, "C2C FP32 N=10^6 batch 8, forward+inverse"),
    "c2c_f32_4x3pow12x64": (4 * 3**12, 64, "C2C FP32 N=4*3^12 batch 64, two 81-point column passes and a "
                                            "324-point row pass, forward+inverse"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2c_f32_2pow24", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged-batch leg")
    ap.add_argument("--launch-breakdown", action="store_true",
                    help="also time each launch of each plan run alone (HIP events between launches)")
    ap.add_argument("--no-extras", action="store_true", help="skip the split-precision and learned-planner legs")
    ap.add_argument("--copy-probe", action="store_true",
                    help="measure the device's streaming copy rate at the 2^24 footprint and exit")
    ap.add_argument("--split-only", action="store_true",
                    help="run only the split-precision leg (one JSON line; for kernel traces)")
    ap.add_argument("--planner-only", action="store_true",
                    help="run only the learned-planner leg (one JSON line)")
    ap.add_argument("--configs-only", action="store_true",
                    help="run only the configurations leg (one JSON line; for kernel traces)")
    ap.add_argument("--cpu-configs", action="store_true",
                    help="time every configuration on the host's cores through the oracle's restatements "
                         "of the reference's routes (one JSON line; no GPU)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="counter runs: 3 forward plans of the workload, then 3 copies of the same bytes "
                         "(the calibration kernel of known traffic); no timing, no JSON")
    ap.add_argument("--pmc-ragged", action="store_true",
                    help="counter runs of the ragged batch: 3 launches of its default plan, then 3 copies of "
                         "the same bytes; no timing, no JSON")
    return ap.parse_args()


def pmc_probe(args):
    import numpy as np
    import torch
    import __graft_entry__

    F = __graft_entry__.load_package()
    n, batch, _ = WORKLOADS[args.workload]
    x = torch.from_numpy(F.rng.uniform_complex(0, n * batch).astype(np.complex64)).to("cuda:0")
    y = torch.empty_like(x)
    p = F.plan_fft(torch.float32, n, batch=batch)
    print(p.describe(), flush=True)
    for _ in range(3):
        p(y, x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("pmc probe done", flush=True)


def pmc_ragged():
    import numpy as np
    import torch
    import __graft_entry__

    F = __graft_entry__.load_package()
    sizes = F.ragged.benchmark_sizes()
    x = torch.from_numpy(F.rng.uniform_complex(1, int(sizes.sum())).astype(np.complex64)).to("cuda:0")
    y = torch.empty_like(x)
    p = F.plan_ragged(torch.float32, sizes)
    for _ in range(3):
        p(y, x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("pmc ragged probe done", flush=True)


def copy_probe():
    """Streaming ceilings on this device at the 2^24 footprint: torch copies
    of 128 MiB among three rotating buffers, between two fixed buffers, and
    an in-place pass (read and write) over one 128 MiB array -- the three
    footprints a pass of the 2^24 plans can have (the in-place one fits the
    256 MB Infinity Cache)."""
    import torch
    n = 1 << 24
    b = [torch.zeros(n, dtype=torch.complex64, device="cuda:0") for _ in range(3)]
    out = {}
    cases = (("rotating_3x128MiB", [(0, 1), (1, 2), (2, 0)]), ("fixed_2x128MiB", [(0, 1)]), ("inplace_1x128MiB", [(0, 0)]))
    for name, pairs in cases:
        def one(i):
            s, d = pairs[i % len(pairs)]
            if s == d:
                b[d].mul_(1.0)
            else:
                b[d].copy_(b[s])
        for i in range(6):
            one(i)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        it = 60
        e0.record()
        for i in range(it):
            one(i)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / it
        out[name] = {"us_per_copy": round(us, 2), "gbs": round(16.0 * n / (us * 1e-6) / 1e9, 1)}
    print(json.dumps({"copy_probe": out}), flush=True)


def host_cpu():
    """What the CPU baseline ran on: model, logical CPUs, the CPUs this
    process may use, AVX-512 presence, and how the oracle was compiled."""
    model, avx512 = "unknown", False
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags"):
                    avx512 = avx512 or " avx512f" in line
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(), "cpus_allowed": allowed, "avx512f": avx512,
            "oracle_build": "gcc -O3 -march=x86-64-v3 (AVX2; the compiler's auto-vectorisation of scalar C, "
                            "no explicit SIMD lanes)"}


def cpu_threads():
    """Threads of the CPU baseline: the CPUs this process may use, at most 16
    (the host share of one GPU)."""
    return min(host_cpu()["cpus_allowed"], 16)


def cpu_baseline(n, batch, target_s):
    """The oracle's FP32 restatement of FFTJ's threaded path, timed here."""
    import numpy as np
    from oracle import oracle as O
    import __graft_entry__

    F = __graft_entry__.load_package()
    threads = cpu_threads()
    x = F.rng.uniform_complex(0, n * batch).astype(np.complex64)
    flops_step = 2 * 5.0 * n * math.log2(n) * batch

    def one_step():
        if batch == 1:
            # FourStepPlan with the reference's AVX-512 split (fourstep.jl:166-180)
            N2 = 2048 if n == 1 << 24 else 1 << (int(math.log2(n)) // 2)
            N1 = n // N2
            y = O.fourstep(x, N1, N2, sign=-1, prec="f32", nthreads=threads)
            O.fourstep(y, N1, N2, sign=1, scale=1.0 / n, prec="f32", nthreads=threads)
        else:
            xb = x.reshape(batch, n)
            y = O.baseline_fft_batch(xb, sign=-1, prec="f32", nthreads=threads)
            O.baseline_fft_batch(y, sign=1, scale=1.0 / n, prec="f32", nthreads=threads)

    t0 = time.perf_counter()
    one_step()
    first = time.perf_counter() - t0
    reps = max(1, min(50, int(target_s / max(first, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(reps):
        one_step()
    dt = (time.perf_counter() - t0) / reps
    return {
        "value": round(flops_step / dt / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{reps} forward+inverse steps of the same workload on {threads} threads "
                  f"(oracle FP32 port of FFTJ's threaded {'FourStepPlan' if batch == 1 else 'per-transform FFTPlan loop'})",
        "host": host_cpu(),
    }


def cpu_configs_leg(F, per_config_s=3.0):
    """Every configuration of the configurations leg, forward, through the
    oracle's restatements of the reference's routes on this host's cores
    (bounded samples: about per_config_s seconds each).  FP32 arithmetic
    except the FP64 configuration; threads as the reference threads them
    (batches across cores, the four-step threaded, the N-d plans serial in
    this restatement)."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    from oracle import oracle as O

    threads = cpu_threads()
    pool = ThreadPoolExecutor(threads)
    rng = F.rng

    def rows(fn, X):
        # one transform per task: the C calls release the GIL
        return list(pool.map(fn, list(X)))

    c64 = lambda seed, count: rng.uniform_complex(seed, count).astype(np.complex64)  # noqa: E731
    r32 = lambda seed, count: rng.normal(seed, count).astype(np.float32)  # noqa: E731
    cases = []
    x1024 = rng.uniform_complex(0, 1024)
    cases.append(("c2c_f64_1024", lambda: O.fft(x1024, prec="f64"), 1, "FFTPlan chain, serial"))
    xb = c64(0, 4096 * 16384).reshape(16384, 4096)
    cases.append(("c2c_f32_4096x16384", lambda: O.baseline_fft_batch(xb, prec="f32", nthreads=threads), threads,
                  "loop of 1-D FFTPlans across cores"))
    x24 = c64(0, 1 << 24)
    cases.append(("c2c_f32_2pow24", lambda: O.fourstep(x24, 8192, 2048, prec="f32", nthreads=threads), threads,
                  "FourStepPlan 8192x2048, threaded"))
    x6 = c64(0, 8 * 10 ** 6).reshape(8, 10 ** 6)
    cases.append(("c2c_f32_1e6x8", lambda: O.baseline_fft_batch(x6, prec="f32", nthreads=min(8, threads)),
                  min(8, threads), "1-D plan per transform across cores"))
    for n in (65537, 1000003):
        xp = c64(0, 4 * n).reshape(4, n)
        cases.append((f"c2c_f32_{n}x4", (lambda xp=xp: rows(lambda r: O.hvdh(r, prec="f32"), xp)), 4,
                      f"HvdhPlan (t = {O.hvdh_t(n)}) per transform across cores"))
    xr = r32(0, 4 << 22).reshape(4, 1 << 22)
    cases.append(("r2c_f32_2pow22x4", lambda: O.baseline_rfft_batch(xr, prec="f32", nthreads=4), 4,
                  "RfftPlan per transform across cores, inner 2^21 as one chain (the reference runs it "
                  "as a threaded FourStepPlan)"))
    Xr = O.baseline_rfft_batch(xr, prec="f32", nthreads=4)
    cases.append(("c2r_f32_2pow22x4", lambda: O.baseline_irfft_batch(Xr, 1 << 22, prec="f32", nthreads=4), 4,
                  "IrfftPlan per transform across cores, inner 2^21 as one chain"))
    x3 = c64(0, 1 << 24)
    cases.append(("c2c_f32_256cubed", lambda: O.baseline_ndfft(x3, [256] * 3, prec="f32", nthreads=threads),
                  threads, "NdPlan axis passes, pencils across cores"))
    xr3 = r32(0, 1 << 24).reshape(65536, 256)

    def r2c3():
        Y = O.baseline_rfft_batch(xr3, prec="f32", nthreads=threads)
        return O.baseline_ndfft(Y, [129, 256, 256], prec="f32", nthreads=threads, first=1)

    cases.append(("r2c_f32_256cubed", r2c3, threads, "RfftNdPlan: real rows, then axis passes, across cores"))
    X3 = r2c3()

    def c2r3():
        Z = O.baseline_ndfft(X3, [129, 256, 256], sign=1, scale=1.0 / 65536, prec="f32", nthreads=threads, first=1)
        return O.baseline_irfft_batch(Z.reshape(65536, 129), 256, prec="f32", nthreads=threads)

    cases.append(("c2r_f32_256cubed", c2r3, threads, "IrfftNdPlan: axis passes, then real rows, across cores"))
    out = []
    for name, fn, cores, route in cases:
        t0 = time.perf_counter()
        fn()
        first = time.perf_counter() - t0
        reps = max(1, min(20, int(per_config_s / max(first, 1e-4))))
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ms = (time.perf_counter() - t0) / reps * 1e3
        out.append({"config": name, "ms": round(ms, 3), "cores": cores, "reps": reps, "route": route})
        print(json.dumps(out[-1]), file=sys.stderr, flush=True)
    pool.shutdown()
    return out


def ragged_leg(F, quick=False):
    """The ragged configuration: 10,000 FP32 transforms, one persistent launch
    (autotuned), against one batched plan per distinct length on the data laid
    out by length (the per-size baseline) and against the reference's loop of
    1-D plans, one call per transform; every side also captured in a HIP
    graph."""
    import numpy as np
    import torch

    sizes = F.ragged.benchmark_sizes()
    x = torch.from_numpy(F.rng.uniform_complex(1, int(sizes.sum())).astype(np.complex64)).to("cuda:0")
    y = torch.empty_like(x)
    t0 = time.perf_counter()
    plan, cfg, results = F.tune.autotune_ragged(torch.float32, sizes, x=x, y=y, iters=5,
                                               budget=2 if quick else None, seed_model=True)
    tune_s = time.perf_counter() - t0
    tuned_ms = F.tune.time_plan(plan, y, x, iters=20)
    default_ms = F.tune.time_plan(F.plan_ragged(torch.float32, sizes), y, x, iters=20)
    loop = F.ragged.PlainLoop(torch.float32, sizes)
    plain_ms = F.tune.time_plan(loop, y, x, iters=3, warmup=1)
    # the per-size baseline: one batched plan per distinct length on the
    # same data laid out by length (no gather counted)
    per = F.ragged.PerSizeLoop(torch.float32, sizes)
    xs = per.to_layout(x)
    ys = torch.empty_like(xs)
    per_size_ms = F.tune.time_plan(per, ys, xs, iters=5, warmup=1)
    # every side captured in a HIP graph (no host launch cost): the ragged
    # plan is one node, the per-size loop one node per distinct length, the
    # plain loop one node per transform
    graph_ms = {}
    for name, fn in (("ragged", lambda: plan(y, x)), ("per_size", lambda: per(ys, xs)),
                     ("plain_loop", lambda: loop(y, x))):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            fn()
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            fn()
        graph_ms[name] = F.tune.time_plan(lambda _y, _x: g.replay(), y, x, iters=5 if name == "plain_loop" else 20)
        del g
    del per, xs, ys
    bytes_ = 16.0 * float(sizes.sum())
    flops = float((5.0 * sizes * np.log2(np.maximum(sizes, 2))).sum())
    return {"transforms": int(len(sizes)), "distinct_sizes": int(len(np.unique(sizes))),
            "sum_n": int(sizes.sum()), "tuned_ms": round(tuned_ms, 4), "default_ms": round(default_ms, 4),
            "plain_per_size_ms": round(graph_ms["per_size"], 4), "plain_per_size_eager_ms": round(per_size_ms, 4),
            "speedup_vs_per_size": round(graph_ms["per_size"] / graph_ms["ragged"], 2),
            "plain_loop_ms": round(plain_ms, 3), "speedup_tuned_vs_plain": round(plain_ms / tuned_ms, 2),
            "graph_ragged_ms": round(graph_ms["ragged"], 4), "graph_plain_loop_ms": round(graph_ms["plain_loop"], 3),
            "speedup_in_graphs": round(graph_ms["plain_loop"] / graph_ms["ragged"], 2),
            "tuned_config": cfg, "tuning_s": round(tune_s, 2),
            "tuned_gbs": round(bytes_ / (tuned_ms * 1e-3) / 1e9, 1),
            "tuned_gflops": round(flops / (tuned_ms * 1e-3) / 1e9, 1),
            "candidates": [[c, round(ms, 4)] for c, ms in results]}


def split_precision_leg(F):
    """N = 2^20 and N = 3^12*4, batch 64, forward: the plain FP32 chain next to
    the split-precision MFMA chains (FP16 hi/lo, BF16 hi/lo, BF16 hi/mid/lo).
    Accuracy against the FP64 oracle is measured by the tests; here the
    round trip forward -> inverse of each mode is reported."""
    import numpy as np
    import torch

    out = []
    for n in (1 << 20, 3 ** 12 * 4):
        batch = 64
        x = torch.from_numpy(F.rng.uniform_complex(2, n * batch).astype(np.complex64)).to("cuda:0")
        y = torch.empty_like(x)
        z = torch.empty_like(x)
        for mode in (None, "f16", "bf16", "bf16x3"):
            pf = F.plan_fft(torch.float32, n, batch=batch, precision=mode)
            pi = F.plan_ifft(torch.float32, n, batch=batch, precision=mode)
            ms = F.tune.time_plan(pf, y, x, iters=20)
            pi(z, y)
            rt = (torch.linalg.vector_norm(z - x) / torch.linalg.vector_norm(x)).item()
            launches = pf.profile_launches(y, x, iters=10)
            out.append({"n": n, "batch": batch, "precision": mode or "fp32", "ms": round(ms, 4),
                        "gflops": round(5.0 * n * math.log2(n) * batch / (ms * 1e-3) / 1e9, 1),
                        "gbs_transform": round(16.0 * n * batch / (ms * 1e-3) / 1e9, 1),
                        "launch_ms": [round(v, 5) for v in launches], "roundtrip_rel_err": rt})
            del pf, pi
        del x, y, z
    return out


def configs_leg(F):
    """Every configuration of the parity matrix, forward, timed once each
    (HIP events, inputs resident): ms, GFLOP/s (5 N log2 N per complex
    transform, 2.5 N log2 N per real one) and GB/s on the algorithmic bytes
    (one read + one write of the transform's input and output)."""
    import numpy as np
    import torch

    cases = [("c2c_f64_1024", "c2c", torch.float64, (1024,), 1), ("c2c_f32_4096x16384", "c2c", torch.float32, (4096,), 16384),
             ("c2c_f32_2pow24", "c2c", torch.float32, (1 << 24,), 1), ("c2c_f32_1e6x8", "c2c", torch.float32, (10 ** 6,), 8),
             ("c2c_f32_65537x4", "c2c", torch.float32, (65537,), 4), ("c2c_f32_1000003x4", "c2c", torch.float32, (1000003,), 4),
             ("r2c_f32_2pow22x4", "r2c", torch.float32, (1 << 22,), 4), ("c2r_f32_2pow22x4", "c2r", torch.float32, (1 << 22,), 4),
             ("c2c_f32_256cubed", "c2c", torch.float32, (256, 256, 256), 1),
             ("r2c_f32_256cubed", "r2c", torch.float32, (256, 256, 256), 1),
             ("c2r_f32_256cubed", "c2r", torch.float32, (256, 256, 256), 1)]
    out = []
    for name, kind, T, dims, batch in cases:
        n = int(np.prod(dims))
        cdt = torch.complex64 if T == torch.float32 else torch.complex128
        es = 8 if T == torch.float32 else 16
        if kind == "c2c":
            x = torch.randn(batch * n, dtype=cdt, device="cuda:0")
            if len(dims) > 1:
                x = x.reshape(dims)
                p = F.plan_fft(x)
            else:
                p = F.plan_fft(T, n, batch=batch)
            y = torch.empty_like(x)
            bytes_ = 2 * es * n * batch
            flops = 5.0 * n * math.log2(n) * batch
        elif kind == "r2c":
            # N-d: the real array of torch shape dims, halved along the last axis
            h = dims[-1] // 2 + 1
            hn = n // dims[-1] * h
            x = torch.randn(batch * n, dtype=T, device="cuda:0")
            if len(dims) > 1:
                x = x.reshape(dims)
                p = F.plan_rfft(x)
                y = torch.empty(dims[:-1] + (h,), dtype=cdt, device="cuda:0")
            else:
                p = F.plan_rfft(T, n, batch=batch)
                y = torch.empty(batch * hn, dtype=cdt, device="cuda:0")
            bytes_ = (es // 2) * n * batch + es * hn * batch
            flops = 2.5 * n * math.log2(n) * batch
        else:
            h = dims[-1] // 2 + 1
            hn = n // dims[-1] * h
            x = torch.randn(batch * hn, dtype=cdt, device="cuda:0")
            if len(dims) > 1:
                x = x.reshape(dims[:-1] + (h,))
                p = F.plan_irfft(x, dims[-1])
                y = torch.empty(dims, dtype=T, device="cuda:0")
            else:
                p = F.plan_irfft(T, n, batch=batch)
                y = torch.empty(batch * n, dtype=T, device="cuda:0")
            bytes_ = (es // 2) * n * batch + es * hn * batch
            flops = 2.5 * n * math.log2(n) * batch
        ms = F.tune.time_plan(p, y, x, iters=10)
        out.append({"config": name, "ms": round(ms, 4), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
                    "gbs": round(bytes_ / (ms * 1e-3) / 1e9, 1), "launches": p.describe().count("\n  ")})
        del p, x, y
    return out


def planner_leg(F):
    """Held-out lengths of the learned planner's training set, at both
    training footprints (32 MiB and 256 MiB of complex FP32 per plan): the
    default schedule vs the learned pick, timed now (a pick that is the
    default counts as equal)."""
    import numpy as np
    import torch

    P = F.planner
    data = P.load_dataset()
    lengths = sorted({int(n) for n in data["n"][P.held_out_mask(data)]})[::8]
    lp = P.LearnedPlanner()
    rows = []
    for n in lengths:
        for fp in P.FOOTPRINTS:
            batch = P.batch_for(n, fp)
            x = torch.randn(n * batch, dtype=torch.complex64, device="cuda:0")
            y = torch.empty_like(x)
            t_def = P.time_plan(P.make_plan(n, batch, {}), y, x) * 1e3
            p, cand, _ = lp.choose(n, batch)
            t_l = P.time_plan(p, y, x) * 1e3 if cand else t_def
            rows.append((n, batch, t_def, t_l))
            del p, x, y
    r = np.array(rows, dtype=np.float64)[:, 1:]
    return {"held_out_cases": len(rows), "margin": lp.model.margin,
            "geomean_speedup_learned_over_default": round(float(np.exp(np.mean(np.log(r[:, 1] / r[:, 2])))), 4),
            "worst_learned_over_default": round(float(np.max(r[:, 2] / r[:, 1])), 4),
            "cases_learned_5pct_slower": int(np.sum(r[:, 2] > 1.05 * r[:, 1])),
            "cases": [[int(a), int(b), round(c, 2), round(e, 2)] for a, b, c, e in rows]}


def main():
    args = parse()
    if args.pmc_probe:
        pmc_probe(args)
        return
    if args.pmc_ragged:
        pmc_ragged()
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.cpu_configs:
        import __graft_entry__
        print(json.dumps({"cpu_configs": cpu_configs_leg(__graft_entry__.load_package())}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "a design, not built)"}))
        sys.exit(2)
    import numpy as np
    import torch
    import __graft_entry__

    F = __graft_entry__.load_package()
    if args.split_only:
        print(json.dumps({"split_precision": split_precision_leg(F)}), flush=True)
        return
    if args.configs_only:
        print(json.dumps({"configs": configs_leg(F)}), flush=True)
        return
    if args.planner_only:
        print(json.dumps({"learned_planner": planner_leg(F)}), flush=True)
        return
    n, batch, desc = WORKLOADS[args.workload]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    x = torch.from_numpy(F.rng.uniform_complex(0, n * batch).astype(np.complex64)).to(dev)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    pf = F.plan_fft(torch.float32, n, batch=batch)
    pi = F.plan_ifft(torch.float32, n, batch=batch)

    def step():
        pf(y, x)
        pi(z, y)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = ev0.elapsed_time(ev1) / args.steps
    flops_step = 2 * 5.0 * n * math.log2(n) * batch
    gflops = flops_step / (ms_step * 1e-3) / 1e9

    # accuracy of this run's round trip (the forward is checked against the
    # oracle by the tests; here only a cheap property)
    rt_err = (torch.linalg.vector_norm(z - x) / torch.linalg.vector_norm(x)).item()

    # Roofline on the transform's algorithmic bytes: one read and one write of
    # the complex FP32 array per direction (16*N*batch bytes), two directions
    # per step, over the HIP-event-timed step.  The plan's launches run back
    # to back on torch's current stream, which the events are recorded on (a
    # kernel trace of this command shows no gaps between them), so the step
    # time is the kernels' time.  The per-launch figure (the same bytes over
    # the mean launch duration, one launch being one pass over the whole
    # array) is kept as a secondary field: it is what the kernel trace's
    # per-launch average is compared with.
    import re as _re
    desc_f, desc_i = pf.describe(), pi.describe()
    nl_f = int(_re.search(r"launches=(\d+)", desc_f).group(1))
    nl_i = int(_re.search(r"launches=(\d+)", desc_i).group(1))
    kernel = _re.search(r"n=(\d+) radices=\S+ kernel=(\w+)", desc_f)
    t_launch = ms_step * 1e-3 / (nl_f + nl_i)
    bytes_per_direction = 16.0 * n * batch  # one read + one write of the complex FP32 array
    achieved = 2 * bytes_per_direction / (ms_step * 1e-3) / 1e9
    launch_gbs = bytes_per_direction / t_launch / 1e9
    # per-launch durations of each plan run alone (events between launches:
    # each includes the gap before it), on request
    launches = None
    if args.launch_breakdown:
        launches = pf.profile_launches(y, x, iters=20) + pi.profile_launches(z, y, iters=20)
    # HBM bytes per direction (the sum over one direction's launches) from the
    # committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs of this workload's plan
    traffic = None
    tp = os.path.join(ROOT, "profiles", f"pmc_traffic_{args.workload}.md")
    if os.path.exists(tp):
        with open(tp) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_direction:"):
                    traffic = int(line.split(":")[1])

    result = {
        "metric": "FP32 C2C GFLOP/s & %HBM peak at N=2^24; ragged-batch tuned vs plain speedup",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "wall_s": round(wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: uniform [-1,1) complex, splitmix64 seed 0, resident in HBM",
        "config": {"workload": args.workload, "description": desc, "n": n, "batch": batch,
                   "launches_per_direction": nl_f, "parallelism": "single GPU"},
        "roundtrip_rel_err": rt_err,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "basis": "16*N*batch algorithmic bytes per direction, 2 directions per step, over ms_per_step; "
                              "traffic = counter bytes of one direction's launches",
                     "us_per_direction": round(ms_step * 1e3 / 2, 2),
                     "kernel": (f"{kernel.group(2)} pass kernel, {kernel.group(1)}-point transforms, "
                                f"{nl_f} launches per direction" if kernel else "pass kernel"),
                     "per_launch": {"launch_us_mean": round(t_launch * 1e6, 2), "launch_gbs": round(launch_gbs, 1),
                                    "launch_frac": round(launch_gbs / HBM_PEAK_GBS, 4)}},
    }
    if launches is not None:
        result["roofline"]["launch_ms_isolated"] = [round(v, 5) for v in launches]
    if not args.no_ragged:
        result["ragged"] = ragged_leg(F)
    if not args.no_extras:
        result["configs"] = configs_leg(F)
        result["split_precision"] = split_precision_leg(F)
        result["learned_planner"] = planner_leg(F)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(n, batch, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
