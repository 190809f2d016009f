"""Autotuner of the fused assembly form: tile shape, output groups, register target, block size
and schedule, chosen by exhaustive timing or by the learned cost model (costmodel.py).
`python -m symboliclib_amd.tuner tune-cotan` searches the cotangent Laplacian's fused form
exhaustively (its selection is committed as tuning/c2_selected.json).

Search space (SPACE): tile LDS budget (the whole workgroup's LDS, which sets the tile shape:
how many owner nodes a tile takes), output groups per element, minimum waves per SIMD, threads
per block, and the schedule (one workgroup per tile; persistent work queue in build order or
largest tiles first).  A tile budget needs its own build of the tiles; everything else is a retune of
the same unit (new code, new tile order), compiled ahead in parallel.

Features (FEATURES, 24 per program and configuration) come from the tiles (outputs,
contributions, elements, staged inputs, LDS, runs per tile), the emitted code (FP64 operations
per element, values a work item keeps live) and the launch (groups, waves per SIMD, block size,
schedule, estimated residency).  The model predicts log(milliseconds); the tuner ranks all
configurations by prediction and times only the best few.  Exhaustive search stays the
fallback and the measure of the model's regret.
"""
import ctypes
import itertools
import json
import math
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _native

SPACE = {
    "lds": [40448, 53760, 65024, 81408],  # whole-workgroup LDS: 4, 3, 2, 2 workgroups per CU
    "groups": [2, 3, 5],
    "wpe": [0, 4],
    "tpb": [256, 512, 768, 1024],
    "schedule": ["static", "queue", "largest"],
}

FEATURES = [
    "log2_outputs", "log2_tiles", "log2_outputs_per_tile", "cv_outputs_per_tile", "contributions_per_output",
    "constant_fraction", "log2_elements_per_tile", "recompute", "element_lds_per_tile", "lds_per_workgroup",
    "runs_per_output", "staged_per_output", "log2_adds", "log2_muls", "divs_sqrts", "max_live", "groups", "wpe",
    "tpb", "persistent", "counter", "largest", "blocks_per_cu", "log2_waves_of_tiles",
]
CUS = 256
LDS_PER_CU = 160 * 1024


def configs(space=SPACE):
    for lds in space["lds"]:
        for g, w, t, s in itertools.product(space["groups"], space["wpe"], space["tpb"], space["schedule"]):
            yield {"lds": lds, "groups": g, "wpe": w, "tpb": t, "schedule": s}


def code_key(c):
    return (c["groups"], c["wpe"], c["tpb"], c["schedule"] == "static", c["schedule"] == "strided")


def unit_options(c):
    return {"output_groups": c["groups"], "waves_per_eu": c["wpe"], "threads_per_block": c["tpb"],
            "schedule": c["schedule"]}


def residency(lds_bytes, tpb, max_live, wpe):
    """Workgroups per compute unit estimated without compiling: LDS, and registers from the
    live values (two VGPRs per double plus addressing), capped by the register target."""
    by_lds = LDS_PER_CU // max(1, lds_bytes)
    vgprs = min(512, 2 * max_live + 24)
    waves_per_simd = max(1, 512 // max(vgprs, 1))
    if wpe:
        waves_per_simd = max(waves_per_simd, wpe)
    waves_per_simd = min(8, waves_per_simd)
    by_regs = max(1, (4 * waves_per_simd) // max(1, tpb // 64))
    return max(1, min(by_lds, by_regs, 2048 // tpb))


PRODUCT_SPACE = {
    "lds": [32768, 49152, 65536, 98304],
    "groups": [1],
    "wpe": [0, 4],
    "tpb": [256, 512, 1024],
    "schedule": ["static", "strided", "queue", "largest"],
}


def features(fstats, tile_stats, cfg):
    t = np.asarray(tile_stats, np.float64)
    m, S, el, st, vals, words, nconst, runs = (t[:, k] for k in range(8))
    n_out = max(1.0, m.sum())
    n_tiles = max(1, len(m))
    ev = max(1.0, el.sum())
    # code statistics are summed over the element templates: per element, average (product
    # tiles have no element code: the op mix is one multiply-add per product)
    n_t = 2.0
    if fstats["num_elements"] == 0:
        fstats = dict(fstats, adds=S.sum() / n_out, muls=S.sum() / n_out, element_evaluations=0, num_elements=1)
        n_t = 1.0
    bpc = residency(int(fstats["lds_bytes"]), cfg["tpb"], int(fstats["max_live"]), cfg["wpe"])
    f = [
        math.log2(n_out), math.log2(n_tiles), math.log2(max(1.0, m.mean())), float(m.std() / max(1.0, m.mean())),
        float(S.sum() / n_out), float(nconst.sum() / n_out), math.log2(max(1.0, el.mean())),
        float(fstats["element_evaluations"] / max(1, fstats["num_elements"])), float(8 * (vals + st).mean() / 65536),
        float(fstats["lds_bytes"] / 65536), float(runs.sum() / n_out), float(st.sum() / n_out),
        math.log2(1 + fstats["adds"] / n_t), math.log2(1 + fstats["muls"] / n_t),
        float((fstats["divs"] + fstats["sqrts"]) / n_t / 10), float(fstats["max_live"] / 64), cfg["groups"] / 16,
        cfg["wpe"] / 8, cfg["tpb"] / 1024, float(cfg["schedule"] != "static"),
        float(cfg["schedule"] in ("queue", "largest")), float(cfg["schedule"] == "largest"), bpc / 8,
        math.log2(max(1.0, n_tiles / (CUS * bpc))),
    ]
    assert len(f) == len(FEATURES)
    return np.array(f, np.float32)


class Timer:
    """HIP-event timing of unit launches on torch's current stream."""

    def __init__(self):
        import torch

        self.torch = torch
        self.stream = torch.cuda.current_stream()
        self.sp = ctypes.c_void_p(self.stream.cuda_stream)

    def time(self, unit, reps=10, warmup=2):
        torch = self.torch
        for _ in range(warmup):
            unit.run(self.sp)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(self.stream)
        for _ in range(reps):
            unit.run(self.sp)
        e.record(self.stream)
        torch.cuda.synchronize()
        return s.elapsed_time(e) / reps


def precompile(unit, cfgs, threads=16):
    """Compile every code variant of `cfgs` for `unit` into the code-object cache."""
    L = _native.lib()
    seen = {}
    for c in cfgs:
        k = code_key(c)
        if k not in seen:
            seen[k] = unit.fused_source(**unit_options(c)).encode()
    with ThreadPoolExecutor(threads) as ex:
        rcs = list(ex.map(lambda src: L.symhip_cache_compile(src, None, None), seen.values()))
    if any(rcs):
        raise _native.SymlibError("precompiling fused variants failed: " + (L.symhip_last_error() or b"").decode())
    return len(seen)


def code_stats(unit, cfgs):
    """fused_stats() for every code variant (emitting only; the unit must not be compiled)."""
    out = {}
    for c in cfgs:
        k = (c["groups"], c["wpe"])
        if k not in out:
            unit.retune(**unit_options(c))
            out[k] = unit.fused_stats()
    return out


class Tuner:
    """build(lds) -> a fused-form Unit (not compiled); inputs: its input arrays."""

    def __init__(self, build, inputs, space=SPACE, log=None):
        self.build = build
        self.inputs = inputs
        self.space = space
        self.log = log or (lambda *_: None)

    def _units(self):
        for lds in self.space["lds"]:
            t0 = time.perf_counter()
            u = self.build(lds)
            self.log(f"[tune] tiles for lds={lds}: {u.info()['tiling']} ({time.perf_counter() - t0:.1f}s)")
            yield lds, u, time.perf_counter() - t0

    def _cfgs(self, lds):
        return [c for c in configs(self.space) if c["lds"] == lds]

    def exhaustive(self, timer, reps=10, record_features=True):
        """Times every configuration; returns records {config, ms, features}."""
        recs = []
        for lds, u, _ in self._units():
            cfgs = self._cfgs(lds)
            tiles = u.tile_stats()
            stats = code_stats(u, cfgs) if record_features else {}
            precompile(u, cfgs)
            u.compile().execute(*self.inputs)
            for c in cfgs:
                u.retune(**unit_options(c))
                ms = timer.time(u, reps=reps)
                rec = {"config": c, "ms": ms}
                if record_features:
                    rec["features"] = features(stats[(c["groups"], c["wpe"])], tiles, c).tolist()
                recs.append(rec)
            u.close()
            best = min(recs, key=lambda r: r["ms"])
            self.log(f"[tune] lds={lds}: {len(cfgs)} timed, best so far {best['ms']:.4f} ms {best['config']}")
        return recs

    def predicted(self, model):
        """Features and predicted log(ms) of every configuration (builds tiles, emits code, no
        timing).  Returns (configs, predictions, units by lds, build seconds)."""
        cfgs, feats, units, secs = [], [], {}, 0.0
        for lds, u, s in self._units():
            secs += s
            cs = self._cfgs(lds)
            stats = code_stats(u, cs)
            tiles = u.tile_stats()
            for c in cs:
                cfgs.append(c)
                feats.append(features(stats[(c["groups"], c["wpe"])], tiles, c))
            units[lds] = u
        pred = model.predict(np.stack(feats))
        return cfgs, pred, units, secs

    def select(self, model, timer, top_k=3, reps=10):
        """Model-driven choice: rank by prediction, time the top_k, keep the fastest.
        Returns (config, ms, trials, seconds)."""
        t0 = time.perf_counter()
        cfgs, pred, units, _ = self.predicted(model)
        order = np.argsort(pred)[:top_k]
        best = None
        for i in order:
            c = cfgs[int(i)]
            u = units[c["lds"]]
            u.retune(**unit_options(c))
            if not u.runtime():
                u.compile().execute(*self.inputs)
            ms = timer.time(u, reps=reps)
            if best is None or ms < best[1]:
                best = (c, ms)
        for u in units.values():
            u.close()
        return best[0], best[1], len(order), time.perf_counter() - t0


def regret(chosen_ms, records):
    best = min(r["ms"] for r in records)
    return chosen_ms / best - 1.0


def write_db(path, program, records):
    """Append records to a timing database (gzip-compressed when the path ends in .gz)."""
    import gzip

    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "at") as f:
        for r in records:
            f.write(json.dumps(dict(r, program=program)) + "\n")


def read_db(path):
    """Timing database: JSON lines, optionally gzip-compressed (.gz)."""
    import gzip

    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as f:
        return [json.loads(line) for line in f if line.strip()]


def db_arrays(records):
    X = np.array([r["features"] for r in records], np.float32)
    y = np.log(np.array([r["ms"] for r in records], np.float64)).astype(np.float32)
    return X, y


# ---------------------------------------------------------------------------------------------
# command line: collect timing samples, train the model, select a configuration

def cloth_program(seed, max_vertices=1_500_000):
    """A random batch of cloths (training programs of the cost model): count 2^U{0..6}, sides
    log-uniform in a random sub-range of [8, 256], total vertices bounded."""
    from . import workloads as W

    rng = np.random.default_rng(1000 + seed)
    while True:
        count = int(2 ** rng.integers(0, 7))
        lo = int(rng.integers(8, 96))
        hi = int(rng.integers(lo, 257))
        sides = W.c4_sides(count=count, lo=lo, hi=hi, seed=2000 + seed)
        if sum(s * s for s in sides) <= max_vertices:
            break
    return sides, [W.cloth(s, seed=3000 + 97 * seed + k) for k, s in enumerate(sides)]


def cloth_tuner(cloths, space=SPACE, log=None):
    from . import workloads as W
    from .unit import Unit

    def build(lds):
        return Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True,
                                     tile_lds_bytes=lds)

    return Tuner(build, [W.to_colmajor(X) for _, _, X in cloths], space, log)


def matrix_tuner(cp, ri, a, space=PRODUCT_SPACE, log=None):
    from .unit import Unit

    def build(lds):
        return Unit.ata_plus_a(cp, ri, fused=True, tile_lds_bytes=lds)

    return Tuner(build, [a], space, log)


def evaluate(model, recs):
    """Per program of `recs` (timing records of every configuration): rank the configurations by
    the model's prediction and take the fastest recorded time among the best k (k = 1, 3); the
    regret is how much slower that is than the fastest configuration.  Means and maxima per kind."""
    key = lambda r: (r["program"]["kind"], r["program"]["seed"])  # noqa: E731
    out = {"heldout": []}
    for kind, sd in sorted({key(r) for r in recs}):
        rs = [r for r in recs if key(r) == (kind, sd)]
        Xh, yh = db_arrays(rs)
        pred = model.predict(Xh)
        row = {"kind": kind, "seed": sd, "configurations": len(rs), "best_ms": min(r["ms"] for r in rs),
               "median_config_ms": float(np.median([r["ms"] for r in rs])),
               "rank_corr": float(np.corrcoef(pred, yh)[0, 1])}
        for k in (1, 3):
            top = np.argsort(pred)[:k]
            chosen = min(rs[int(i)]["ms"] for i in top)
            row[f"regret_top{k}"] = regret(chosen, rs)
        out["heldout"].append(row)
    for kind in ("cloth_batch", "ata_plus_a"):
        rows = [h for h in out["heldout"] if h["kind"] == kind]
        if rows:
            out[f"{kind}_regret_top1_mean"] = float(np.mean([h["regret_top1"] for h in rows]))
            out[f"{kind}_regret_top3_mean"] = float(np.mean([h["regret_top3"] for h in rows]))
            out[f"{kind}_regret_top3_max"] = float(np.max([h["regret_top3"] for h in rows]))
    return out


def _main():
    import argparse
    import sys

    ap = argparse.ArgumentParser(prog="python -m symboliclib_amd.tuner")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect", help="exhaustive timings of random cloth batches -> timing database")
    c.add_argument("--db", required=True)
    c.add_argument("--first", type=int, default=0)
    c.add_argument("--programs", type=int, default=48)
    c.add_argument("--reps", type=int, default=8)
    c5 = sub.add_parser("collect-patterns", help="exhaustive timings of the synthetic A^T A + A patterns")
    c5.add_argument("--db", required=True)
    c5.add_argument("--first", type=int, default=0)
    c5.add_argument("--count", type=int, default=64)
    c5.add_argument("--reps", type=int, default=8)
    t = sub.add_parser("train", help="train the cost model on a timing database")
    t.add_argument("--db", required=True, nargs="+")
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--holdout", type=int, default=4, help="cloth programs kept out of training (the last seeds)")
    t.add_argument("--holdout-patterns", type=int, default=16, help="A^T A + A patterns kept out (the last seeds)")
    e = sub.add_parser("evaluate", help="regret of a trained model on the programs of a timing database")
    e.add_argument("--model", required=True)
    e.add_argument("--db", required=True, nargs="+")
    s = sub.add_parser("select", help="tune the batched cloth workload: model choice vs exhaustive search")
    s.add_argument("--model", required=True)
    s.add_argument("--out", required=True)
    s.add_argument("--count", type=int, default=256)
    s.add_argument("--top-k", type=int, default=3)
    s.add_argument("--no-exhaustive", action="store_true")
    c2 = sub.add_parser("tune-cotan", help="exhaustive search of the cotangent Laplacian's fused form (512 x 512 grid)")
    c2.add_argument("--out", required=True)
    c2.add_argument("--side", type=int, default=512)
    c2.add_argument("--reps", type=int, default=20)
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731

    if args.cmd == "collect":
        timer = Timer()
        for p in range(args.first, args.first + args.programs):
            sides, cloths = cloth_program(p)
            t0 = time.perf_counter()
            recs = cloth_tuner(cloths, log=log).exhaustive(timer, reps=args.reps)
            write_db(args.db, {"kind": "cloth_batch", "seed": p, "sides": sides}, recs)
            best = min(recs, key=lambda r: r["ms"])
            log(f"[collect] program {p}: {len(sides)} cloths, {len(recs)} samples in {time.perf_counter() - t0:.0f}s, "
                f"best {best['ms']:.4f} ms {best['config']}")
    elif args.cmd == "collect-patterns":
        from . import workloads as W

        timer = Timer()
        sizes = W.c5_sizes()
        for idx in range(args.first, min(len(sizes), args.first + args.count)):
            t0 = time.perf_counter()
            n, cp, ri, a = W.c5_pattern(idx)
            recs = matrix_tuner(cp, ri, a, log=log).exhaustive(timer, reps=args.reps)
            write_db(args.db, {"kind": "ata_plus_a", "seed": idx, "n": n, "nnz": int(len(ri))}, recs)
            best = min(recs, key=lambda r: r["ms"])
            log(f"[collect] pattern {idx} (n={n}): {len(recs)} samples in {time.perf_counter() - t0:.0f}s, "
                f"best {best['ms']:.4f} ms {best['config']}")
    elif args.cmd == "train":
        from .costmodel import CostModel

        recs = [r for path in args.db for r in read_db(path)]
        key = lambda r: (r["program"]["kind"], r["program"]["seed"])  # noqa: E731
        held = set()
        for kind, k in (("cloth_batch", args.holdout), ("ata_plus_a", args.holdout_patterns)):
            seeds = sorted({r["program"]["seed"] for r in recs if r["program"]["kind"] == kind})
            held |= {(kind, s) for s in (seeds[-k:] if k else [])}
        tr = [r for r in recs if key(r) not in held]
        X, y = db_arrays(tr)
        m = CostModel.create(len(FEATURES), 64, seed=1)
        t0 = time.perf_counter()
        loss = m.train(X, y, epochs=args.epochs, batch=256, lr=1e-3, seed=2)
        secs = time.perf_counter() - t0
        progs = {key(r) for r in tr}
        log(f"[train] {len(tr)} samples from {len(progs)} programs, {args.epochs} epochs in {secs:.1f}s on the GPU, "
            f"loss {loss:.4f} (standardised)")
        m.save(args.out)
        out = {"train_samples": len(tr), "train_programs": len(progs), "epochs": args.epochs,
               "train_seconds": secs, "loss": loss}
        out.update(evaluate(m, [r for r in recs if key(r) in held]))
        print(json.dumps(out))
    elif args.cmd == "evaluate":
        from .costmodel import CostModel

        out = {"model": args.model, "db": args.db}
        out.update(evaluate(CostModel.load(args.model), [r for path in args.db for r in read_db(path)]))
        print(json.dumps(out))
    elif args.cmd == "select":
        from . import workloads as W
        from .costmodel import CostModel

        timer = Timer()
        cloths = W.c4_cloths(count=args.count)
        tuner = cloth_tuner(cloths, log=log)
        model = CostModel.load(args.model)
        cfg, ms, trials, secs = tuner.select(model, timer, top_k=args.top_k)
        res = {"workload": f"c4 ({args.count} cloths)", "selected": cfg, "selected_ms": ms, "model_trials": trials,
               "model_seconds": secs, "space": SPACE, "configurations": len(list(configs()))}
        log(f"[select] model choice {cfg}: {ms:.4f} ms after {trials} trials in {secs:.0f}s")
        if not args.no_exhaustive:
            t0 = time.perf_counter()
            recs = tuner.exhaustive(timer, record_features=False)
            best = min(recs, key=lambda r: r["ms"])
            # regret within the exhaustive pass (the chosen configuration's time there), and
            # across passes (its time in the model's own trials), which includes run-to-run noise
            same = next((r for r in recs if r["config"] == cfg), None)
            res.update(exhaustive_best=best["config"], exhaustive_ms=best["ms"],
                       regret=regret(same["ms"] if same else ms, recs), regret_across_passes=regret(ms, recs),
                       exhaustive_seconds=time.perf_counter() - t0, exhaustive_trials=len(recs))
            log(f"[select] exhaustive best {best['config']}: {best['ms']:.4f} ms; regret {res['regret']:.3f}")
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
        print(json.dumps(res))


    elif args.cmd == "tune-cotan":
        from . import workloads as W
        from .unit import Unit

        V, F = W.c2_mesh(side=args.side)
        # element matrices of 6 values: one output group is in the space
        space = dict(SPACE, groups=[1] + SPACE["groups"])
        t = Tuner(lambda lds: Unit.cotan(len(V), F, fused=True, tile_lds_bytes=lds), [W.to_colmajor(V)], space, log)
        t0 = time.perf_counter()
        recs = sorted(t.exhaustive(Timer(), reps=args.reps, record_features=False), key=lambda r: r["ms"])
        res = {"workload": f"c2 (cotangent Laplacian, {args.side} x {args.side} grid)", "selected": recs[0]["config"],
               "selected_ms": recs[0]["ms"], "by": "exhaustive search", "seconds": time.perf_counter() - t0,
               "space": space, "configurations": len(recs),
               "best": [{"config": r["config"], "ms": r["ms"]} for r in recs[:8]],
               "median_ms": recs[len(recs) // 2]["ms"]}
        log(f"[tune-cotan] best {recs[0]['config']}: {recs[0]['ms']:.4f} ms of {len(recs)} configurations")
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
        print(json.dumps(res))


if __name__ == "__main__":
    _main()
