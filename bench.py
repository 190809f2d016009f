"""Benchmark: FP64 sparse-expression evaluation throughput (output nonzeros per second).

One step = one evaluation of the compiled program (all launches) over one value set whose
inputs are already resident in HBM.  Default workload: the Baraff-Witkin cloth Hessian
(stretch + shear by symbolic differentiation, plus bending) of a 256 x 256 sheet, 4,890,648
output nonzeros, through the plain single-instance HIP path.

Prints one JSON line with the throughput, the roofline figure of the evaluation (algorithmic
bytes / kernel time against the 8 TB/s HBM peak), and the CPU baseline: the FP64 oracle
(OpenMP over the host cores) timed on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|c1] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(name, **opt):
    from symboliclib_amd import workloads as W
    from symboliclib_amd.unit import Unit

    if name == "c1":
        n, cp, ri, a = W.c1_matrix()
        unit = Unit.ata_plus_a(cp, ri, **opt)
        inputs = a
        counts = {"n_in": len(ri), "n_const": 0, "n_conn": len(ri) + n + 1}
        desc = {"workload": "A^T A + A, icosphere subdivision 5", "rows": n, "nnz_A": int(len(ri))}

        def oracle_fn():
            import oracle as O

            return O.ata_plus_a(cp, ri, a)

        def cpu_eval():
            import oracle as O

            O.ata_plus_a(cp, ri, a)

    elif name == "c2":
        V, F = W.c2_mesh()
        unit = Unit.cotan(len(V), F, **opt)
        inputs = W.to_colmajor(V)
        counts = {"n_in": 3 * len(V), "n_const": 0, "n_conn": 3 * len(F)}
        desc = {"workload": "cotangent Laplacian, 512x512 jittered grid", "vertices": len(V), "faces": len(F)}

        def oracle_fn():
            import oracle as O

            return O.cotan(V, F)

        plan = {}

        def cpu_eval():
            import oracle as O

            if "p" not in plan:
                plan["p"] = O.CotanPlan(len(V), F)
                plan["nnz"] = O.cotan_pattern(len(V), F)[0][-1]
            plan["p"].values(V, plan["nnz"])

    else:
        rest, F, X = W.c3_cloth()
        unit = Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, **opt)
        inputs = W.to_colmajor(X)
        import oracle as O

        ne = len(O.interior_edges(F))
        counts = {"n_in": 3 * len(rest), "n_const": 5 * len(F) + 10 * ne, "n_conn": 3 * len(F) + 4 * ne}
        desc = {"workload": "Baraff-Witkin cloth Hessian (stretch+shear AD + bend), 256x256 sheet",
                "vertices": len(rest), "faces": len(F), "interior_edges": ne}

        def oracle_fn():
            return O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)

        plan = {}

        def cpu_eval():
            if "p" not in plan:
                plan["p"] = O.BWPlan(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
            plan["p"].values(X, with_scale=False)

    return unit, inputs, counts, desc, oracle_fn, cpu_eval


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--form", default="waves", choices=["tiled", "waves"],
                    help="waves: one launch per wave, intermediates in HBM (default); tiled: LDS-staged work tiles")
    ap.add_argument("--tile-lds", type=int, default=0, help="LDS bytes per tile (0 = default)")
    ap.add_argument("--tpb", type=int, default=256, help="threads per block")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "the path shards without communication, see DESIGN.md", file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))

    import torch

    from symboliclib_amd.unit import device_available

    if not torch.cuda.is_available() or not device_available():
        raise RuntimeError("bench.py needs a HIP device")
    dev = torch.device("cuda:0")

    log(f"[bench] building workload {args.workload}")
    t0 = time.perf_counter()
    unit, inputs, counts, desc, oracle_fn, cpu_eval = build_workload(
        args.workload, tiled=args.form == "tiled", tile_lds_bytes=args.tile_lds, threads_per_block=args.tpb)
    t1 = time.perf_counter()
    unit.compile()
    t2 = time.perf_counter()
    info = unit.info()
    log(f"[bench] front-end+build {t1 - t0:.1f}s, compile {t2 - t1:.1f}s, info {info}")

    unit.execute(inputs)  # H2D of the inputs + one evaluation
    torch.cuda.synchronize()
    vals = unit.get_results()

    import oracle as O

    cp, ri, rv, sc = oracle_fn()
    outer, inner = unit.output_pattern()
    pattern_ok = bool(np.array_equal(outer, cp) and np.array_equal(inner, ri))
    err = O.rel_errors(vals, rv, sc)
    log(f"[bench] parity: pattern {'bit-exact' if pattern_ok else 'DIFFERS'}, max rel err {err.max():.3e}, "
        f"median {np.median(err):.3e}")
    if not pattern_ok or err.max() > 1e-12:
        raise RuntimeError("bench.py: result does not match the oracle")

    n_out = int(info["num_results"])
    b_alg = 8 * counts["n_in"] + 8 * counts["n_const"] + 12 * n_out + 4 * counts["n_conn"]

    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    for _ in range(args.warmup):
        unit.run(sp)
    torch.cuda.synchronize()
    log(f"[bench] warmup done ({args.warmup} steps)")

    # timed region: K evaluations back to back
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(args.steps):
        unit.run(sp)
    torch.cuda.synchronize()
    te = time.perf_counter()
    ms_per_step = (te - ts) * 1e3 / args.steps
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.4f} ms/step")

    # kernel time per launch (HIP events on the launch stream)
    nl = unit.num_launches()
    reps = max(10, min(args.steps, 100))
    launch_ms = []
    for li in range(nl):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(reps):
            unit.launch_one(li, sp)
        e.record(stream)
        torch.cuda.synchronize()
        launch_ms.append(s.elapsed_time(e) / reps)
    kernel_ms = sum(launch_ms)

    # cold evaluations: a 1 GiB write evicts caches before each timed evaluation
    flush = torch.empty(1 << 27, dtype=torch.float64, device=dev)
    cold = []
    for _ in range(20):
        flush.zero_()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        unit.run(sp)
        e.record(stream)
        torch.cuda.synchronize()
        cold.append(s.elapsed_time(e))
    cold_ms = float(np.median(cold))
    del flush

    # device copy bandwidth for reference
    cbuf = torch.empty(1 << 27, dtype=torch.float64, device=dev)
    dbuf = torch.empty_like(cbuf)
    dbuf.copy_(cbuf)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(10):
        dbuf.copy_(cbuf)
    e.record(stream)
    torch.cuda.synchronize()
    copy_gbs = 2 * cbuf.numel() * 8 * 10 / (s.elapsed_time(e) * 1e-3) / 1e9
    del cbuf, dbuf

    value = n_out / (ms_per_step * 1e-3)
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    line = {
        "metric": "output nnz/s FP64 (plain single-instance HIP)",
        "value": value,
        "unit": "nnz/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": dict(desc, output_nnz=n_out, form=f"plain single-instance, {args.form}", tiling=info["tiling"],
                       threads_per_block=args.tpb, launches=nl, parallelism="single GPU"),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "symlib_wave* (all launches of one evaluation)",
            "algorithmic_bytes": b_alg,
            "kernel_ms": kernel_ms,
            "launch_ms": launch_ms,
        },
        "cold_ms_per_step": cold_ms,
        "cold_nnz_per_s": n_out / (cold_ms * 1e-3),
        "copy_bandwidth_gbs": copy_gbs,
        "parity": {"pattern_bit_exact": pattern_ok, "max_rel_err": float(err.max()),
                   "median_rel_err": float(np.median(err))},
        "build_seconds": {"frontend": info["frontend_seconds"], "build": info["build_seconds"],
                          "compile": info["compile_seconds"]},
    }

    if not args.no_cpu:
        cpu_eval()  # warm (builds the plan)
        n = 0
        tc0 = time.perf_counter()
        while True:
            cpu_eval()
            n += 1
            if time.perf_counter() - tc0 >= args.cpu_seconds:
                break
        tc = time.perf_counter() - tc0
        line["cpu_baseline"] = {
            "value": n_out * n / tc,
            "unit": "nnz/s",
            "cores": O.num_threads(),
            "kind": "port",
            "sample": f"{n} full evaluations of the same workload by the FP64 oracle (OpenMP, "
                      f"{O.num_threads()} threads) in {tc:.1f} s",
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
