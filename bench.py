"""Benchmark: FP64 sparse-expression evaluation throughput (output nonzeros per second).

Default workload (c4): 256 Baraff-Witkin cloth Hessians (stretch + shear by symbolic
differentiation, plus bending) of log-uniform sides 32..256, 345.8 M output nonzeros, batched
into one program in the fused assembly form (owner-node tiles, element Hessians staged in LDS,
one launch replayed from a HIP graph), against the plain single-instance form launched once per
instance (one waves-form program per cloth, two launches each) on the same stream.  The tile
shape, block size and schedule are chosen in the run by the learned cost model: every
configuration of the tuner's space is predicted, the three best predicted are timed, the
fastest is kept (--committed: the committed selection instead, --untuned: defaults).

Single-instance workloads: c3 (one 256 x 256 cloth, 4,890,648 nonzeros), c2 (cotangent
Laplacian of a 512 x 512 grid), c1 (A^T A + A of an icosphere); --form selects waves, tiled,
fused (element tiles for c2 and c3, product tiles for c1) or direct (the waves form of the
direct element assembly, c2 and c3).

One step = one evaluation of the compiled program (every launch) over one value set whose
inputs are already resident in HBM.  Prints one JSON line with the throughput, the roofline
of the dominant kernel (algorithmic bytes per launch / its average launch time, against the
8 TB/s HBM peak and against the measured ceiling of a hand-written STREAM copy) and the CPU
baseline: the reference's CPU path restated (the CPU device: the same programs as host C++
with OpenMP, compiled with the reference's flags) on the host cores, over the whole batch
(--cpu-sample: 16 cloths of evenly spaced sizes).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4|c3|c2|c1] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TUNED = os.path.join(ROOT, "symboliclib_amd", "tuning", "c4_selected.json")
# committed fused-form selections of the single workloads (the cloth's is the batch's)
SINGLE_TUNED = {"c2": os.path.join(ROOT, "symboliclib_amd", "tuning", "c2_selected.json"), "c3": TUNED}
MODEL = os.path.join(ROOT, "symboliclib_amd", "tuning", "cost_model.txt")
FP64_PEAK_TFLOPS = 78.6
TRAFFIC = os.path.join(ROOT, "profiles", "symlib_fused_bw_hessian_batch256_traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def b_alg(counts, n_out):
    """Algorithmic bytes of one evaluation: 8 B per input and per-element constant, 8 B per
    output value plus its 4 B placement index, 4 B per connectivity index."""
    return 8 * counts["n_in"] + 8 * counts["n_const"] + 12 * n_out + 4 * counts["n_conn"]


def fused_options(args, tuned=None):
    o = {"fused": True}
    if tuned:
        o.update(tile_lds_bytes=tuned["lds"], output_groups=tuned["groups"], waves_per_eu=tuned["wpe"],
                 threads_per_block=tuned["tpb"], schedule=tuned["schedule"], split=bool(tuned.get("split", 0)))
    for k, v in (("tile_lds_bytes", args.tile_lds), ("output_groups", args.groups), ("waves_per_eu", args.wpe),
                 ("threads_per_block", args.tpb), ("schedule", args.schedule), ("split", args.split)):
        if v is not None:
            o[k] = v
    if args.output_sums:
        o["output_sums"] = True
    return o


def time_events(fn, stream, reps):
    import torch

    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(reps):
        fn()
    e.record(stream)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def copy_bandwidth(dev, stream):
    """The measured HBM ceiling: a hand-written STREAM copy (symhip_stream_copy) of 1 GiB, the
    best of its one-pass form and two grid-stride grids, HIP events on the launch stream;
    torch's copy for reference."""
    import torch

    from symboliclib_amd import _native

    L = _native.lib()
    cbuf = torch.empty(1 << 27, dtype=torch.float64, device=dev)
    dbuf = torch.empty_like(cbuf)
    sp = ctypes.c_void_p(stream.cuda_stream)
    best = 0.0
    for blocks in (0, 256, 16384):
        run = lambda b=blocks: _native.check(  # noqa: E731
            L.symhip_stream_copy(ctypes.c_void_p(dbuf.data_ptr()), ctypes.c_void_p(cbuf.data_ptr()), cbuf.numel(), b,
                                 sp), "symhip_stream_copy")
        run()
        ms = time_events(run, stream, 10)
        best = max(best, 2 * cbuf.numel() * 8 / (ms * 1e-3) / 1e9)
    dbuf.copy_(cbuf)
    torch_gbs = 2 * cbuf.numel() * 8 / (time_events(lambda: dbuf.copy_(cbuf), stream, 10) * 1e-3) / 1e9
    return best, torch_gbs


# ---------------------------------------------------------------------------------------------
# single-instance workloads


def build_single(name, opt):
    from symboliclib_amd import workloads as W
    from symboliclib_amd.unit import Unit

    import oracle as O

    if name == "c1":
        n, cp, ri, a = W.c1_matrix()
        unit = Unit.ata_plus_a(cp, ri, **opt)
        counts = {"n_in": len(ri), "n_const": 0, "n_conn": len(ri) + n + 1}
        desc = {"workload": "A^T A + A, icosphere subdivision 5", "rows": n, "nnz_A": int(len(ri))}
        return unit, [a], counts, desc, lambda: O.ata_plus_a(cp, ri, a)
    if name == "c2":
        V, F = W.c2_mesh()
        unit = Unit.cotan(len(V), F, **opt)
        counts = {"n_in": 3 * len(V), "n_const": 0, "n_conn": 3 * len(F)}
        desc = {"workload": "cotangent Laplacian, 512x512 jittered grid", "vertices": len(V), "faces": len(F)}
        return unit, [W.to_colmajor(V)], counts, desc, lambda: O.cotan(V, F)
    rest, F, X = W.c3_cloth()
    unit = Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, **opt)
    counts = W.bw_algorithmic_counts(len(rest), F)
    desc = {"workload": "Baraff-Witkin cloth Hessian (stretch+shear AD + bend), 256x256 sheet",
            "vertices": len(rest), "faces": len(F), "interior_edges": counts["interior_edges"]}
    return unit, [W.to_colmajor(X)], counts, desc, lambda: O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)


def run_single(args, dev, stream, sp):
    import torch

    import oracle as O

    if args.form == "fused":
        sel = None
        path = SINGLE_TUNED.get(args.workload)
        if path and os.path.exists(path) and not args.untuned:
            with open(path) as f:
                sel = json.load(f)["selected"]
        opt = fused_options(args, sel)
    else:
        opt = {"tiled": args.form == "tiled", "tile_lds_bytes": args.tile_lds or 0,
               "threads_per_block": args.tpb or 256}
        if args.form == "direct":
            if args.workload == "c1":
                raise SystemExit("bench.py: the direct element assembly applies to c2 and c3")
            opt = {"direct_assembly": True, "threads_per_block": args.tpb or 256}
    log(f"[bench] building workload {args.workload} ({args.form})")
    t0 = time.perf_counter()
    unit, inputs, counts, desc, oracle_fn = build_single(args.workload, opt)
    t1 = time.perf_counter()
    unit.compile()
    t2 = time.perf_counter()
    info = unit.info()
    log(f"[bench] front-end+build {t1 - t0:.1f}s, compile {t2 - t1:.1f}s, {info['tiling']}")
    unit.execute(*inputs)
    torch.cuda.synchronize()
    vals = unit.get_results()
    cp, ri, rv, sc = oracle_fn()
    outer, inner = unit.output_pattern()
    pattern_ok = bool(np.array_equal(outer, cp) and np.array_equal(inner, ri))
    err = O.rel_errors(vals, rv, sc)
    log(f"[bench] parity: pattern {'bit-exact' if pattern_ok else 'DIFFERS'}, max rel err {err.max():.3e}, "
        f"median {np.median(err):.3e}")
    if not pattern_ok or err.max() > 1e-12:
        raise RuntimeError("bench.py: result does not match the oracle")

    n_out = int(info["num_results"])
    step = (lambda: unit.run_graph(sp)) if args.graph else (lambda: unit.run(sp))
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - ts) * 1e3 / args.steps
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.4f} ms/step")

    nl = unit.num_launches()
    reps = max(10, min(args.steps, 100))
    launch_ms = [time_events(lambda li=li: unit.launch_one(li, sp), stream, reps) for li in range(nl)]
    # cold evaluations: a 1 GiB write evicts the caches before each one
    flush = torch.empty(1 << 27, dtype=torch.float64, device=dev)
    cold = []
    for _ in range(20):
        flush.zero_()
        cold.append(time_events(step, stream, 1))
    del flush
    cold_ms = float(np.median(cold))
    copy_gbs, torch_copy_gbs = copy_bandwidth(dev, stream)

    dominant = int(np.argmax(launch_ms))
    names = {"fused": "symlib_fused", "tiled": "symlib_tiled"}
    kernel = names.get(args.form, f"symlib_wave{dominant}")
    balg = b_alg(counts, n_out)
    # the roofline covers the dominant kernel; with one launch that is the whole evaluation
    achieved = balg / (sum(launch_ms) * 1e-3) / 1e9
    line = {
        "metric": "output nnz/s FP64 (single instance)",
        "value": n_out / (ms_per_step * 1e-3),
        "unit": "nnz/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(desc, output_nnz=n_out, form=args.form, tiling=info["tiling"], options=opt, launches=nl,
                       graph=bool(args.graph), parallelism="single GPU"),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "frac_of_measured_copy": achieved / copy_gbs, "traffic": None,
                     "kernel": kernel,
                     "algorithmic_bytes": balg, "kernel_ms": sum(launch_ms), "launch_ms": launch_ms,
                     "note": "achieved = algorithmic bytes of the evaluation / summed launch times"},
        "cold_ms_per_step": cold_ms, "cold_nnz_per_s": n_out / (cold_ms * 1e-3),
        "copy_bandwidth_gbs": copy_gbs, "torch_copy_gbs": torch_copy_gbs,
        "parity": {"pattern_bit_exact": pattern_ok, "max_rel_err": float(err.max()),
                   "median_rel_err": float(np.median(err))},
        "build_seconds": {"frontend": info["frontend_seconds"], "build": info["build_seconds"],
                          "compile": info["compile_seconds"]},
    }
    if args.workload == "c3":
        line["consumer"] = consumer_step(unit, dev, stream, sp, reps)
    unit.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_device_single(args)
        log(f"[bench] cpu baseline: {line['cpu_baseline']['value']:.3g} nnz/s")
    return line


def consumer_step(unit, dev, stream, sp, reps):
    """The step after the Hessian on the cloth's K in place: the implicit-Euler system
    A = M - h (D + h K), b = M v + h (f - D v) (symhip_implicit_euler_system, without damping)
    and the symmetric product y = A x a solver runs (symhip_csc_symv), each timed with HIP
    events; algorithmic bytes per launch: system 20 B per nonzero (K, row index, A) + 36 B per
    column (column pointer, m, v, f, b), product 12 B per nonzero (A, row index) + 20 B per
    column (column pointer, x, y)."""
    import torch

    from symboliclib_amd import _native

    L = _native.lib()
    outer, inner = unit.output_pattern()
    n, nnz = len(outer) - 1, len(inner)
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)  # noqa: E731
    o, i = t(outer, torch.int32), t(inner, torch.int32)
    g = torch.Generator(device="cpu").manual_seed(5)
    m = torch.full((n,), 1.0 / (n // 3), dtype=torch.float64, device=dev)
    v = (0.1 * torch.rand(n, generator=g, dtype=torch.float64)).to(dev)
    f = torch.randn(n, generator=g, dtype=torch.float64).to(dev)
    x = torch.randn(n, generator=g, dtype=torch.float64).to(dev)
    A = torch.empty(nnz, dtype=torch.float64, device=dev)
    b, y = torch.empty_like(m), torch.empty_like(m)
    P = lambda a: ctypes.c_void_p(a.data_ptr())  # noqa: E731
    K = unit.output_device_ptr(0)

    def system():
        _native.check(L.symhip_implicit_euler_system(n, P(o), P(i), K, None, P(m), P(v), P(f), 0.05, P(A), P(b), sp),
                      "symhip_implicit_euler_system")

    def symv():
        _native.check(L.symhip_csc_symv(n, P(o), P(i), P(A), P(x), P(y), sp), "symhip_csc_symv")

    system()
    symv()
    torch.cuda.synchronize()
    ms_sys = time_events(system, stream, reps)
    ms_mv = time_events(symv, stream, reps)
    b_sys, b_mv = 20 * nnz + 36 * n, 12 * nnz + 20 * n
    out = {"workload": "implicit-Euler system of the cloth (h = 0.05, no damping) and y = A x on its pattern",
           "n": n, "nnz": nnz,
           "system": {"kernel": "symhipImplicitEulerSystem", "ms": ms_sys, "algorithmic_bytes": b_sys,
                      "achieved_gbs": b_sys / (ms_sys * 1e-3) / 1e9,
                      "frac": b_sys / (ms_sys * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "symv": {"kernel": "symhipCscSymv", "ms": ms_mv, "algorithmic_bytes": b_mv,
                    "achieved_gbs": b_mv / (ms_mv * 1e-3) / 1e9, "frac": b_mv / (ms_mv * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    log(f"[bench] consumer: system {ms_sys:.4f} ms ({out['system']['achieved_gbs']:.0f} GB/s), "
        f"symv {ms_mv:.4f} ms ({out['symv']['achieved_gbs']:.0f} GB/s)")
    return out


def cpu_device_single(args):
    """The reference's CPU path restated (CPU device, symcpu.h) on the same workload: the program
    of the reference's own front-end for c1 (A^T A + A from makeSymbolic), the direct element
    assembly for c2 and c3 (the same element and sum kernels, built without per-nonzero trees),
    execute() repeated for about args.cpu_seconds after one warm-up."""
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    opt = {"cpu_threads": threads} if args.workload == "c1" else {"cpu_threads": threads, "direct_assembly": True}
    unit, inputs, _, _, _ = build_single(args.workload, opt)
    unit.compile()
    build = time.perf_counter() - t0
    unit.execute(*inputs)
    nnz = int(unit.info()["num_results"])
    n, ts = 0, time.perf_counter()
    while True:
        unit.execute(*inputs)
        n += 1
        if time.perf_counter() - ts >= args.cpu_seconds:
            break
    tc = time.perf_counter() - ts
    unit.close()
    form = "the reference front-end's program" if args.workload == "c1" else "the direct element assembly's program"
    return {"value": nnz * n / tc, "unit": "nnz/s", "cores": threads, "kind": "port",
            "sample": f"the whole workload ({nnz} output nonzeros), {form} on the CPU device (the reference's "
                      f"VecWidth/NumThreads path restated: host C++ from the same program, OpenMP over {threads} "
                      f"threads, clang with the reference's flags); {n} execute() calls in {tc:.1f} s after a warm-up "
                      f"(build and compile {build:.0f} s, untimed)"}


# ---------------------------------------------------------------------------------------------
# batched cloths (c4)


def select_batched(cloths, inputs, args, committed=None):
    """Model-driven choice of the fused form's configuration, in this run: build the tiles of
    every LDS budget of the tuner's space, predict every configuration with the committed cost
    model, time the best `args.trials` predicted (and the committed selection, when it is not
    among them) and keep the fastest.  Returns (unit compiled and executed with that
    configuration, its config, trial records, seconds, configurations)."""
    from symboliclib_amd import costmodel, tuner

    t0 = time.perf_counter()
    tun = tuner.cloth_tuner(cloths, log=log)
    model = costmodel.CostModel.load(MODEL)
    cfgs, pred, units, _ = tun.predicted(model)
    model.close()
    order = [int(i) for i in np.argsort(pred)[: args.trials]]
    if committed is not None:
        extra = [i for i, c in enumerate(cfgs) if all(c[k] == committed[k] for k in c)]
        order += [i for i in extra if i not in order]
    top = [cfgs[i] for i in order]
    for lds in {c["lds"] for c in top}:
        tuner.precompile(units[lds], [c for c in top if c["lds"] == lds])
    timer = tuner.Timer()
    trials = []
    for i in order:
        c = cfgs[i]
        u = units[c["lds"]]
        u.retune(**tuner.unit_options(c))
        if not u.runtime():
            u.compile().execute(*inputs)
        ms = timer.time(u, reps=20)
        trials.append({"config": c, "predicted_ms": float(np.exp(pred[i])), "ms": ms,
                       "predicted_rank": int(np.argsort(np.argsort(pred))[i]) + 1})
        log(f"[bench] trial {c}: predicted {trials[-1]['predicted_ms']:.3f} ms, measured {ms:.4f} ms")
    best = min(trials, key=lambda t: t["ms"])
    for lds, u in units.items():
        if lds != best["config"]["lds"]:
            u.close()
    unit = units[best["config"]["lds"]]
    unit.retune(**tuner.unit_options(best["config"]))
    (unit if unit.runtime() else unit.compile()).execute(*inputs)
    return unit, best["config"], trials, time.perf_counter() - t0, len(cfgs)


def fp64_flops(stats):
    """FP64 operations of one launch of the fused cloth form from the emitted code: every
    element evaluation's adds, multiplies, divisions and square roots, one add per element
    contribution of an output."""
    per_elem = stats["adds"] + stats["muls"] + stats["divs"] + stats["sqrts"]
    return stats["element_evaluations"] * per_elem + stats["contributions"]


def cpu_device_baseline(cloths, sides, args):
    """The reference's CPU path restated (CPU device, symcpu.h) on the host cores: one CPU batch
    program over all the cloths, or (--cpu-sample) over 16 cloths at evenly spaced size ranks,
    the smallest and the largest included; execute() (inputs copied, every kernel run) repeated
    for about args.cpu_seconds after one warm-up."""
    from symboliclib_amd import workloads as W
    from symboliclib_amd.unit import Unit

    threads = max(1, min(16, os.cpu_count() or 1))
    order = np.argsort(sides, kind="stable")
    if not args.cpu_sample:
        pick = list(range(len(cloths)))
    else:
        ranks = np.unique(np.linspace(0, len(order) - 1, min(16, len(order))).round().astype(int))
        pick = sorted(int(order[r]) for r in ranks)
    sub = [cloths[k] for k in pick]
    t0 = time.perf_counter()
    u = Unit.bw_hessian_batch([(r, F) for r, F, _ in sub], W.K_STRETCH, W.K_SHEAR, W.K_BEND, cpu_threads=threads)
    u.compile()
    build = time.perf_counter() - t0
    inputs = [W.to_colmajor(X) for _, _, X in sub]
    u.execute(*inputs)
    nnz = int(u.info()["num_results"])
    n, ts = 0, time.perf_counter()
    while True:
        u.execute(*inputs)
        n += 1
        if time.perf_counter() - ts >= args.cpu_seconds:
            break
    tc = time.perf_counter() - ts
    u.close()
    sample = (f"{'all' if not args.cpu_sample else len(pick)} of the {len(cloths)} cloths"
              f"{' at evenly spaced size ranks' if args.cpu_sample else ''} (sides {min(sides[k] for k in pick)}.."
              f"{max(sides[k] for k in pick)}, {nnz} output nonzeros) as one program of the CPU device (the "
              f"reference's VecWidth/NumThreads path restated: host C++ from the same program, OpenMP over "
              f"{threads} threads, clang with the reference's flags); {n} execute() calls in {tc:.1f} s after a "
              f"warm-up (build and compile {build:.0f} s, untimed)")
    return {"value": nnz * n / tc, "unit": "nnz/s", "cores": threads, "kind": "port", "sample": sample}


def run_batched(args, dev, stream, sp):
    import torch

    from symboliclib_amd import _native
    from symboliclib_amd import workloads as W
    from symboliclib_amd.unit import Unit

    import oracle as O

    sides = W.c4_sides(count=args.count)
    log(f"[bench] c4: {len(sides)} cloths, sides {min(sides)}..{max(sides)}")
    t0 = time.perf_counter()
    cloths = W.c4_cloths(count=args.count)
    counts = [W.bw_algorithmic_counts(len(r), F) for r, F, _ in cloths]
    inputs = [W.to_colmajor(X) for _, _, X in cloths]
    committed = None
    if os.path.exists(TUNED):
        with open(TUNED) as f:
            committed = json.load(f)["selected"]
    selection = None
    if args.untuned or args.committed or any(v is not None for v in (args.tile_lds, args.groups, args.wpe, args.tpb,
                                                                     args.schedule, args.split)):
        opt = fused_options(args, committed if args.committed and not args.untuned else None)
        batched = Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, **opt)
        batched.compile().execute(*inputs)
        source = "committed selection" if args.committed else "command line / defaults"
    else:
        batched, cfg, trials, secs, ncfg = select_batched(cloths, inputs, args, committed)
        opt = fused_options(args, {**cfg, "split": 0})
        selection = {"by": f"cost model {os.path.relpath(MODEL, ROOT)}: {ncfg} configurations predicted, the best "
                           f"{args.trials} and the committed selection timed in this run, the fastest kept",
                     "seconds": secs,
                     "trials": trials, "chosen": cfg, "committed": committed,
                     "chosen_equals_committed": committed is not None and all(committed[k] == cfg[k] for k in cfg)}
        source = "selected in this run by the cost model"
        log(f"[bench] selection: {cfg} in {secs:.0f} s")
    torch.cuda.synchronize()
    info = batched.info()
    log(f"[bench] fused batch ({source}) {opt}: {info['tiling']}, set up in {time.perf_counter() - t0:.1f}s")

    # parity on a spread of instances (smallest, largest, four more); every cloth of the batch
    # is checked by the GPU tests
    order = np.argsort(sides, kind="stable")
    check = sorted({int(order[0]), int(order[-1]), *[int(k) for k in np.linspace(0, len(sides) - 1, 4).astype(int)]})
    worst, med = 0.0, []
    for k in check:
        rest, F, X = cloths[k]
        cp, ri, rv, sc = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
        outer, inner = batched.output_pattern(k)
        if not (np.array_equal(outer, cp) and np.array_equal(inner, ri)):
            raise RuntimeError(f"bench.py: pattern of cloth {k} differs from the oracle")
        err = O.rel_errors(batched.get_results(k), rv, sc)
        worst = max(worst, float(err.max()))
        med.append(float(np.median(err)))
    log(f"[bench] parity on cloths {check}: patterns bit-exact, max rel err {worst:.3e}")
    if worst > 1e-12:
        raise RuntimeError("bench.py: batched result does not match the oracle")

    n_out = int(info["num_results"])
    step = lambda: batched.run_graph(sp)  # noqa: E731
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - ts) * 1e3 / args.steps
    log(f"[bench] batched: {args.steps} steps, {ms_per_step:.4f} ms/step")
    # each launch of the step on its own (the split form has two: elements, then sums)
    names = ["symlib_fused_elem", "symlib_fused_sum"] if opt.get("split") else ["symlib_fused"]
    launch_ms = [time_events(lambda i=i: batched.launch_one(i, sp), stream, max(5, min(args.steps, 50)))
                 for i in range(batched.num_launches())]
    kernel_ms = sum(launch_ms)
    balg = sum(b_alg(c, 0) for c in counts) + 12 * n_out
    stats = batched.fused_stats()
    flops = fp64_flops(stats)
    # the same evaluation from host arrays: execute() copies every cloth's positions over PCIe
    # (blocking copies, as the reference's setArg), then launches; synchronised per step
    in_bytes = 8 * sum(a.size for a in inputs)
    batched.execute(*inputs)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for _ in range(5):
        batched.execute(*inputs)
        torch.cuda.synchronize()
    pcie_ms = (time.perf_counter() - tp) * 1e3 / 5
    log(f"[bench] with host inputs (H2D {in_bytes / 1e6:.0f} MB + launch): {pcie_ms:.3f} ms/step")

    # plain single-instance form, one program per distinct side, launched once per instance
    plain_line = None
    if not args.no_plain:
        t0 = time.perf_counter()
        by_side = {}
        for k, s in enumerate(sides):
            by_side.setdefault(s, k)
        plain = {}
        for s, k in by_side.items():
            rest, F, X = cloths[k]
            plain[s] = Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, direct_assembly=True)
        L = _native.lib()
        with ThreadPoolExecutor(16) as ex:
            rcs = list(ex.map(lambda u: L.symhip_cache_compile(u.source().encode(), None, None), plain.values()))
        if any(rcs):
            raise RuntimeError("bench.py: compiling the plain programs failed")
        for s, u in plain.items():
            u.compile().execute(inputs[by_side[s]])
        torch.cuda.synchronize()
        log(f"[bench] plain: {len(plain)} programs for {len(sides)} cloths built in {time.perf_counter() - t0:.1f}s")
        for s, k in list(by_side.items())[:: max(1, len(by_side) // 4)]:
            rest, F, X = cloths[k]
            cp, ri, rv, sc = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
            if O.rel_errors(plain[s].get_results(), rv, sc).max() > 1e-12:
                raise RuntimeError(f"bench.py: plain program of side {s} does not match the oracle")
        seq = [plain[s] for s in sides]

        def plain_step():
            for u in seq:
                u.run(sp)

        plain_step()
        torch.cuda.synchronize()
        reps = max(3, min(args.steps, 10))
        plain_ms = float(np.median([time_events(plain_step, stream, 1) for _ in range(reps)]))
        plain_line = {"value": n_out / (plain_ms * 1e-3), "unit": "nnz/s", "ms_per_step": plain_ms,
                      "programs": len(plain), "launches": 2 * len(sides),
                      "note": "one waves-form program per cloth side; cloths of equal side share a program "
                              "(same kernels and sizes, the first such cloth's positions)"}
        for u in plain.values():
            u.close()
        log(f"[bench] plain per-instance: {plain_ms:.3f} ms/step")

    copy_gbs, torch_copy_gbs = copy_bandwidth(dev, stream)
    achieved = balg / (kernel_ms * 1e-3) / 1e9
    value = n_out / (ms_per_step * 1e-3)
    # HBM bytes per launch measured by rocprofv3 counter runs on this program's tiles (committed
    # under profiles/; the bytes are those of the tile pool and outputs, the options it was
    # measured with are reported beside it), else null
    traffic, traffic_src = None, None
    if os.path.exists(TRAFFIC):
        with open(TRAFFIC) as f:
            entries = json.load(f)
        for tr in entries if isinstance(entries, list) else [entries]:
            if tr.get("tiling") == info["tiling"]:
                traffic = tr["read_bytes"] + tr["write_bytes"]
                traffic_src = {"profile": os.path.relpath(TRAFFIC, ROOT), "options": tr.get("options"),
                               "read_bytes": tr["read_bytes"], "write_bytes": tr["write_bytes"]}
    line = {
        "metric": "output nnz/s FP64, batched+learned-tuned vs plain HIP; HBM GB/s vs peak",
        "value": value,
        "unit": "nnz/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched Baraff-Witkin cloth Hessians: {len(sides)} cloths, sides log-uniform "
                               f"32..256 (seed 4), stretch+shear AD + bend",
                   "instances": len(sides), "vertices": int(sum(len(r) for r, _, _ in cloths)),
                   "faces": int(sum(len(F) for _, F, _ in cloths)), "output_nnz": n_out,
                   "form": ("fused assembly in two launches (element images, then sums)" if opt.get("split")
                            else "fused assembly, one launch") + ", replayed from a HIP graph",
                   "options": opt, "configuration_source": source,
                   "tiling": info["tiling"], "grid": batched.launch_grid(0), "parallelism": "single GPU"},
        "selection": selection,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "frac_of_measured_copy": achieved / copy_gbs,
                     "traffic": traffic, "traffic_source": traffic_src, "kernel": "+".join(names),
                     "algorithmic_bytes": balg, "kernel_ms": kernel_ms, "launch_ms": dict(zip(names, launch_ms))},
        "fp64": {"flops_per_launch": flops, "achieved_tflops": flops / (kernel_ms * 1e-3) / 1e12,
                 "peak_tflops": FP64_PEAK_TFLOPS, "frac": flops / (kernel_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                 "note": "operations of the emitted code: element evaluations x (adds + multiplies + divisions + "
                         "square roots per element) + one add per element contribution"},
        "pcie_inclusive": {"value": n_out / (pcie_ms * 1e-3), "unit": "nnz/s", "ms_per_step": pcie_ms,
                           "h2d_bytes": in_bytes,
                           "note": "execute() from host arrays: positions copied to HBM every step, then the launch"},
        "plain_per_instance": plain_line,
        "speedup_vs_plain": (value / plain_line["value"]) if plain_line else None,
        "copy_bandwidth_gbs": copy_gbs, "torch_copy_gbs": torch_copy_gbs,
        "parity": {"pattern_bit_exact": True, "checked_instances": check, "max_rel_err": worst,
                   "median_rel_err": float(np.median(med))},
        "build_seconds": {"frontend": info["frontend_seconds"], "build": info["build_seconds"],
                          "compile": info["compile_seconds"]},
    }
    batched.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_device_baseline(cloths, sides, args)
        log(f"[bench] cpu baseline: {line['cpu_baseline']['value']:.3g} nnz/s")
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="c4", choices=["c1", "c2", "c3", "c4"])
    ap.add_argument("--count", type=int, default=256, help="c4: number of cloths")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="c4: skip the plain per-instance baseline")
    ap.add_argument("--untuned", action="store_true",
                    help="default options: no selection (c4), no committed selection (c2, c3 fused)")
    ap.add_argument("--committed", action="store_true", help="c4: the committed selection, no selection in the run")
    ap.add_argument("--trials", type=int, default=3, help="c4: configurations timed in the in-run selection")
    ap.add_argument("--cpu-sample", action="store_true",
                    help="c4: CPU baseline over 16 cloths at evenly spaced size ranks instead of the whole batch")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--form", default="waves", choices=["tiled", "waves", "fused", "direct"],
                    help="c1-c3 execution form (direct: waves form of the direct element assembly, c2/c3)")
    ap.add_argument("--graph", action="store_true", help="c1-c3: replay the launches from a HIP graph")
    ap.add_argument("--tile-lds", type=int, default=None, help="LDS bytes per tile")
    ap.add_argument("--tpb", type=int, default=None, help="threads per block")
    ap.add_argument("--groups", type=int, default=None, help="fused: work items per element")
    ap.add_argument("--wpe", type=int, default=None, help="fused: minimum waves per SIMD")
    ap.add_argument("--schedule", default=None, choices=["static", "queue", "largest", "interleaved", "strided"])
    ap.add_argument("--split", type=int, default=None, choices=[0, 1],
                    help="fused: 1 = element images in HBM, then sums (two launches)")
    ap.add_argument("--output-sums", action="store_true",
                    help="fused: one contribution list per output instead of node-block sums")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); the batched workload "
              "shards by instance without communication (see the sharding section of the design notes)",
              file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))

    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("bench.py needs a HIP device")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    from symboliclib_amd.unit import device_available

    if not device_available():
        raise RuntimeError("bench.py: the native runtime sees no HIP device")
    sp = ctypes.c_void_p(stream.cuda_stream)
    line = run_batched(args, dev, stream, sp) if args.workload == "c4" else run_single(args, dev, stream, sp)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
