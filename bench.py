"""CFPQ fixpoint benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 4] [--kernel auto]

A step is one complete CFPQ fixpoint (all rules to the least fixpoint) of
the chosen seeded workload, starting from the terminal matrices, which are
resident in HBM before the timed region (each step first restores the
matrices of rule heads, the only ones a fixpoint writes, with device-side
copies inside the timed region).  The default workload
is the 65,536-vertex ontology same-generation query (SG grammar, 7
nonterminals, ~4.5e5 edges).

value = effective Tbit-op/s = 2 N^3 x A_ref / (time per step), where A_ref
is the number of add_mul calls the reference worklist driver makes on this
workload (recorded in tests/golden/fixpoints.json by the CPU oracle).

The JSON line also carries
  roofline      the dominant product kernel, timed per launch with hipEvents
                on the engine's stream:
                  sparse / dense (HBM bound): algorithmic bytes per launch --
                    A rows 32 B, gathered B rows 32 B, written C rows 32 B,
                    B tiles staged in LDS 8 KiB -- / mean launch time, vs 8 TB/s;
                  MFMA (matrix-core bound): 2 x 256^3 int8 ops per tile-product
                    / mean launch time, vs 5.0 POP/s dense int8;
  cpu_baseline  the CPU oracle's reference worklist (dense M4RM, k = 8) on the
                same workload, run for a bounded time and expressed in the
                same unit.
"""
import argparse
import importlib
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "fixpoints.json")
TRAFFIC = os.path.join(ROOT, "profiles", "k_product_traffic.json")

HBM_PEAK_GBS = 8000.0
MFMA_I8_PEAK_TOPS = 5000.0
WORKLOADS = {
    0: ("config0", "dyck_er_n1024"),
    1: ("config1", "sg_forest_n16384"),
    2: ("config2", "dense_stress_n32768"),
    4: ("config4", "sg_forest_n65536_types8"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(q, a_ref, budget_s):
    import oracle  # the checker; timed here only as the reported baseline

    g = oracle.ParsedGrammar(q.grammar)
    mats = oracle.initial_matrices(g, q.n, q.src, q.dst, q.label, q.labels)
    cores = oracle.num_threads()
    calls, secs = oracle.fixpoint_worklist_budget(q.n, g.rules_array(), mats, budget_s, "m4rm")
    per_call = secs / max(calls, 1)
    value = 2.0 * q.n ** 3 * calls / secs / 1e12
    # the same first add_mul on one thread (dense M4RM costs the same per call
    # whatever the bits, so one call stands for the rest)
    mats1 = oracle.initial_matrices(g, q.n, q.src, q.dst, q.label, q.labels)
    oracle.set_num_threads(1)
    try:
        c1, s1 = oracle.fixpoint_worklist_budget(q.n, g.rules_array(), mats1, 0.0, "m4rm")
    finally:
        oracle.set_num_threads(cores)
    whole = calls >= a_ref
    return {
        "value": round(value, 4),
        "unit": "Tbit-op/s",
        "cores": cores,
        "cpu_model": cpu_model(),
        "kind": "port",
        "complete_fixpoint": whole,
        "sample": (f"the whole reference worklist ({calls} add_mul calls, dense M4RM, k=8) in {secs:.1f} s"
                   if whole else
                   f"first {calls} of {a_ref} add_mul calls of the reference worklist "
                   f"(dense M4RM, k=8), {secs:.1f} s, {per_call:.2f} s per call; "
                   f"extrapolated full fixpoint {per_call * a_ref:.0f} s") +
                  f"; 1 thread: {s1 / max(c1, 1):.2f} s for the first call",
        "value_1thread": round(2.0 * q.n ** 3 * c1 / s1 / 1e12, 4) if s1 > 0 else None,
    }


def product_roofline(stats, wname):
    """Roofline of the product kernel that took the most device time."""
    launches = max(sum(s["product_launches"] for s in stats), 1)
    t_sparse = sum(s["sparse_ms"] for s in stats)
    t_dense = sum(s["dense_ms"] for s in stats)
    t_mfma = sum(s["mfma_ms"] for s in stats)
    if t_mfma >= t_sparse + t_dense:
        ops = sum(s["tile_products_mfma"] for s in stats) * 2.0 * 256 ** 3
        executed = sum(s["mfma_instr"] for s in stats) * 2.0 * 32 ** 3
        per_launch_ms = t_mfma / launches
        achieved = ops / launches / (per_launch_ms / 1e3) / 1e12 if per_launch_ms > 0 else 0.0
        return {"bound": "mfma", "achieved": round(achieved, 3), "peak": MFMA_I8_PEAK_TOPS, "unit": "TFLOP/s",
                "frac": round(achieved / MFMA_I8_PEAK_TOPS, 5), "traffic": None, "kernel": "k_product_mfma",
                "ops": "int8 MACs x2", "launch_ms": round(per_launch_ms, 5),
                "alg_ops_per_launch": round(ops / launches), "executed_ops_per_launch": round(executed / launches)}
    if t_sparse >= t_dense:
        kern, t_k = "k_product_sparse", t_sparse
        # what the sparse kernel's algorithm reads and writes: per candidate
        # tile-product the filter's metadata (B tile row mask 32 B, A quarter
        # column union 32 B and row mask 8 B, B tile count 4 B); per gathered
        # (hit, column) item the quarter's column word (8 B) and the B row
        # (32 B), per A row of the row-major gathers 32 B; per output row the
        # head row read and the new row written (2 x 32 B)
        tp_sparse = sum(s["tile_products"] - s["tile_products_dense"] - s["tile_products_mfma"] for s in stats)
        alg_bytes = 76 * tp_sparse + sum(32 * (s["a_rows"] + s["b_rows"]) + 8 * s["a_cols"] + 64 * s["out_rows"]
                                         for s in stats)
    else:
        kern, t_k = "k_product_dense", t_dense
        alg_bytes = sum(32 * (s["a_rows_dense"] + s["b_rows_dense"] + s["out_rows_dense"]) +
                        8192 * s["tile_products_dense"] for s in stats)
    per_launch_ms = t_k / launches
    per_launch_bytes = alg_bytes / launches
    achieved = per_launch_bytes / (per_launch_ms / 1e3) / 1e9 if per_launch_ms > 0 else 0.0
    traffic = None
    if os.path.exists(TRAFFIC):
        t = json.load(open(TRAFFIC))
        if t.get("workload") == wname and t.get("kernel") == kern:
            traffic = t.get("bytes_per_launch")
    # the unit-based figure of a tile-product (256^3): (256*256 + 256*256)/8
    # bytes of operand tiles per candidate tile-product, as if every
    # candidate were read whole -- a dense-equivalent rate, beside the bytes
    # the kernel actually had to touch
    tp_bytes = sum(s["tile_products"] - s["tile_products_dense"] - s["tile_products_mfma"] for s in stats) * 16384.0
    if kern == "k_product_dense":
        tp_bytes = sum(s["tile_products_dense"] for s in stats) * 16384.0
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": kern,
            "launch_ms": round(per_launch_ms, 5), "alg_bytes_per_launch": round(per_launch_bytes),
            "tile_product_bytes_per_launch": round(tp_bytes / launches),
            "tile_product_equivalent_GBps": round(tp_bytes / launches / (per_launch_ms / 1e3) / 1e9, 1)
            if per_launch_ms > 0 else None}


def measure_worklist(args, eng, gr, mats, init, want, opts):
    """The reference driver's structure on the same GPU: its worklist loop
    (cfpq-cpp/Grammar.h:58-69) with one cfpq_add_mul per popped rule, each
    returning its changed flag to the host, as the reference's GPU backends do
    (cuda/naive-gpu/Multiplication.cu:145-166).  Same kernels, no delta."""
    eng.set_opts(**opts)
    runs = max(1, min(args.steps, 3))
    times, stats = [], []
    for k in range(runs + 1):
        for m, c in zip(mats, init):
            m.copy_from(c)
        eng.sync()
        t0 = time.perf_counter()
        calls = gr._worklist(eng)
        eng.sync()
        if k:  # the first run warms up
            times.append((time.perf_counter() - t0) * 1e3)
            stats.append({"add_mul_calls": calls})
    ok = gr.results() == want
    if not ok:
        log("[bench] worklist MISMATCH vs oracle fixture")
    log(f"[bench] worklist: {statistics.median(times):.2f} ms per fixpoint, {stats[-1]['add_mul_calls']} add_mul calls")
    return statistics.median(times), stats, ok


def measure_pairs(args, eng, gr, q, want):
    """The sparse pair-list formulation (cfpq_pairs_fixpoint) on the same
    workload and context: host pair arrays in and out."""
    keys = gr.pair_keys(cf_module().Graph.from_query(q))
    rules = gr.rule_triples()
    out, _ = eng.pairs_fixpoint(rules, keys, q.n)  # warm-up, sizes the output arrays
    cap = max(len(k) for k in out) if out else None
    times, stats = [], []
    for _ in range(max(1, min(args.steps, 3))):
        t0 = time.perf_counter()
        out, st = eng.pairs_fixpoint(rules, keys, q.n, cap=cap)
        times.append((time.perf_counter() - t0) * 1e3)
        stats.append(st)
    ok = gr.pair_results(out, q.n) == want
    if not ok:
        log("[bench] pairs MISMATCH vs oracle fixture")
    log(f"[bench] pairs: {statistics.median(times):.2f} ms per call, "
        f"{statistics.median(s['fixpoint_ms'] for s in stats):.2f} ms device")
    return statistics.median(times), stats, ok, int(sum(len(k) for k in out))


def measure_mfma(eng, gr, mats, init, rules, opts, want, step):
    """The same fixpoint with every dense tile-product on the int8 matrix cores
    (v_mfma_i32_32x32x32_i8, 0/1 operands expanded from the bit tiles): its
    time and its rate against the 5.0 POP/s dense int8 peak, beside the
    tuned (Four Russians) run.  One run: it is the slow leg."""
    eng.set_opts(**dict(opts, kernel="mfma"))
    st = step()
    ok = gr.results() == want
    eng.set_opts(**opts)
    ops = st["tile_products_mfma"] * 2.0 * 256 ** 3
    executed = st["mfma_instr"] * 2.0 * 32 ** 3
    t = st["mfma_ms"] / 1e3
    log(f"[bench] mfma leg: {st['fixpoint_ms']:.1f} ms, {st['tile_products_mfma']} MFMA tile-products")
    return {"fixpoint_ms": round(st["fixpoint_ms"], 3), "mfma_kernel_ms": round(st["mfma_ms"], 3),
            "tile_products_mfma": st["tile_products_mfma"],
            "achieved_TOPs": round(ops / t / 1e12, 2) if t > 0 else None,
            "frac_of_int8_peak": round(ops / t / 1e12 / MFMA_I8_PEAK_TOPS, 4) if t > 0 else None,
            "executed_TOPs": round(executed / t / 1e12, 2) if t > 0 else None, "bit_exact": ok}


def measure_total(args, eng, gr, mats, init, rules, opts, want):
    """The reference's "total" window (cfpq-cpp/Grammar.h:53-80): upload of
    every nonterminal matrix from host buffers (environment_preprocessing), the
    fixpoint, and download of every matrix back to host buffers
    (environment_postprocessing), PCIe included.  The host buffers of the
    inputs (m4ri row layout) are made once, outside the window."""
    eng.set_opts(**opts)
    host = [c.download() for c in init]
    times, algo = [], []
    for _ in range(2):
        eng.sync()
        t0 = time.perf_counter()
        for m, h in zip(mats, host):
            m.upload(h)
        ta = time.perf_counter()
        eng.fixpoint(rules, mats)
        eng.sync()
        tb = time.perf_counter()
        out = [m.download() for m in mats]
        t1 = time.perf_counter()
        times.append((t1 - t0) * 1e3)
        algo.append((tb - ta) * 1e3)
    ok = gr.results() == want
    del out
    mb = sum(h.nbytes for h in host) / 2 ** 20
    log(f"[bench] total window: {min(times):.1f} ms (algorithm {min(algo):.2f} ms), {mb:.0f} MiB each way")
    return {"total_ms": round(min(times), 3), "algorithm_ms": round(min(algo), 4), "mib_each_way": round(mb, 1),
            "runs": len(times), "bit_exact": ok}


def cf_module():
    return importlib.import_module("cfpq-on-gpgpu_amd.cfpq")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=4, choices=sorted(WORKLOADS) + [3])
    ap.add_argument("--graph", action="store_true", help="replay captured hipGraphs per step")
    ap.add_argument("--kernel", default="auto")
    ap.add_argument("--copies", type=int, default=1,
                    help="--formulation pairs: run K disjoint copies of the workload as one graph")
    ap.add_argument("--formulation", choices=["bits", "nt", "pairs"], default="bits",
                    help="nt: the single matrix of nonterminal sets (cfpq_nt_closure; host buffers, small N)")
    ap.add_argument("--schedule", choices=["delta", "sweep", "rules"], default=None,
                    help="rule schedule of the fixpoint (default: the tuned one)")
    ap.add_argument("--iters-per-sync", type=int, default=0)
    ap.add_argument("--sparse-batch", type=int, default=0)
    ap.add_argument("--sparse-waves", type=int, default=0)
    ap.add_argument("--sparse-mode", type=int, default=-1, choices=[-1, 0, 1, 2],
                    help="sparse product: 0 quarter units, 2 tile units, 1 row-per-lane (default: tuned)")
    ap.add_argument("--tune", action="store_true", help="autotune on this workload first (cached in tuned.json)")
    ap.add_argument("--tune-repeats", type=int, default=3)
    ap.add_argument("--tune-passes", type=int, default=1)
    ap.add_argument("--tune-dims", default="",
                    help="comma-separated parameters: walk only these, from the cached choice")
    ap.add_argument("--cpu-budget", type=float, default=30.0,
                    help="seconds of CPU oracle time; the whole 65,536-vertex worklist takes about 23 s on 16 threads")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-scheduler engine")
    ap.add_argument("--no-worklist", action="store_true",
                    help="skip the reference-structured engine (worklist of add_mul calls)")
    ap.add_argument("--no-pairs", action="store_true", help="skip the pair-list formulation's leg")
    ap.add_argument("--no-mfma", action="store_true", help="--config 2: skip the forced int8 MFMA leg")
    ap.add_argument("--no-graph-leg", action="store_true", help="skip the graph-replay leg")
    ap.add_argument("--no-total", action="store_true",
                    help="skip the total window (upload + fixpoint + download from host buffers)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run one plain device copy of MIB MiB (torch), the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--opt", action="append", default=[], metavar="NAME=VALUE",
                    help="override one engine option (cfpq_opts field), e.g. --opt split_cands=128")
    ap.add_argument("--sched-threshold", type=float, default=0.0,
                    help="defer items with predicted activity below this probability (default: tuned/library)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is designed but not built", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        src = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {args.copy_probe} MiB read + {args.copy_probe} MiB written")
        del src, dst
    pkg = importlib.import_module("cfpq-on-gpgpu_amd")
    gen = importlib.import_module("cfpq-on-gpgpu_amd.gen")
    cf = importlib.import_module("cfpq-on-gpgpu_amd.cfpq")
    if args.config == 3:
        return bench_batch(args, pkg, gen, cf)
    if args.formulation == "pairs":
        return bench_pairs(args, pkg, gen, cf)
    if args.formulation == "nt":
        return bench_nt(args, pkg, gen, cf)
    key, wname = WORKLOADS[args.config]
    golden = json.load(open(GOLDEN))[key]
    q = gen.config(args.config)
    assert q.digest() == golden["digest"], "workload generator drifted from the golden fixture"
    # work unit: 2 N^3 bit-ops per add_mul of the reference worklist; where the
    # fixture has no worklist count (the dense-stress grammar, whose worklist
    # takes hours on the CPU), per M4RM product of the oracle's semi-naive run
    work_key = "add_mul_calls" if "add_mul_calls" in golden else "seminaive_m4rm_products"
    a_ref = int(golden[work_key])
    n = q.n
    log(f"[bench] workload {wname}: n={n} edges={q.n_edges} A_ref={a_ref}")

    tune = importlib.import_module("cfpq-on-gpgpu_amd.tune")
    if args.tune:
        tune.tune(pkg, cf, q, repeats=args.tune_repeats, log=log, passes=args.tune_passes,
                  dims=[d for d in args.tune_dims.split(",") if d])
        os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
        with open(tune.CACHE) as f, open(os.path.join(ROOT, "gpurun_out", "tuned.json"), "w") as g:
            g.write(f.read())
    opts = tune.tuned_opts(n, q.grammar)
    opts["kernel"] = args.kernel
    if args.schedule:
        opts["schedule"] = {"delta": 0, "sweep": 1, "rules": 2}[args.schedule]
    if args.iters_per_sync:
        opts["iters_per_sync"] = args.iters_per_sync
    if args.graph:
        opts["use_graph"] = True
    if args.sparse_batch:
        opts["sparse_batch"] = args.sparse_batch
    if args.sparse_waves:
        opts["sparse_waves_per_cu"] = args.sparse_waves
    if args.sparse_mode >= 0:
        opts["sparse_mode"] = args.sparse_mode
    for kv in args.opt:
        k, v = kv.split("=", 1)
        opts[k] = int(v, 0)
    log(f"[bench] engine options {opts}")
    eng = pkg.Engine(0, **opts)
    gr = cf.Grammar(q.grammar, text=True)
    mats = gr.load(eng, cf.Graph.from_query(q))
    init = []
    for m in mats:
        c = eng.matrix(n)
        c.copy_from(m)
        init.append(c)
    rules = gr.rule_triples()
    # only the heads of rules change during a fixpoint; the other matrices
    # (terminal-only nonterminals) are never written and need no restore
    changing = sorted({h for h, _, _ in rules})

    def step():
        for i in changing:
            mats[i].copy_from(init[i])
        return eng.fixpoint(rules, mats)

    want = {k: tuple(v) for k, v in golden["fingerprints"].items()}

    def measure(label, run_opts):
        eng.set_opts(**run_opts)
        for w in range(args.warmup):
            st = step()
            log(f"[bench] {label} warmup {w}: {st['fixpoint_ms']:.3f} ms, {st['iterations']} steps")
        eng.sync()
        stats = []
        t0 = time.perf_counter()
        for k in range(args.steps):
            stats.append(step())
        eng.sync()
        wall = time.perf_counter() - t0
        log(f"[bench] {label}: timed {args.steps} steps: {wall * 1e3:.2f} ms")
        # size-independent parity check at full size: fingerprints + counts vs the oracle fixture
        got = gr.results()
        ok = got == want
        if not ok:
            log(f"[bench] {label} MISMATCH vs oracle fixture: {got} vs {want}")
        return wall * 1e3 / args.steps, stats, ok

    runs = {"plain": measure("plain", opts)}
    # the same eager engine without the per-kernel timing events (the plain
    # leg's events give the roofline's kernel times; each costs a few
    # microseconds of gap between launches)
    runs["untimed"] = measure("untimed", dict(opts, no_kernel_timing=True))
    if not args.no_graph_leg and not opts.get("use_graph"):
        # the same engine with every step replayed from captured hipGraphs
        # (fewer launch gaps; per-kernel timing needs the eager plain leg)
        runs["graph"] = measure("graph", dict(opts, use_graph=True, iters_per_sync=8))
    sched_mod = importlib.import_module("cfpq-on-gpgpu_amd.sched")
    weights = None if args.no_learned else sched_mod.load()
    if weights is not None:
        eng.sched_set_weights(weights)
        # convergence of the learned engine is certified by one exact full
        # pass over every rule, which must find no new bit
        lopts = dict(opts, sched=True, final_full_pass=True)
        if args.sched_threshold:
            lopts["sched_threshold"] = args.sched_threshold
        runs["learned"] = measure("learned", lopts)
    if not args.no_worklist:
        runs["worklist"] = measure_worklist(args, eng, gr, mats, init, want, opts)
    pairs_run = None
    if not args.no_pairs:
        try:
            pairs_run = measure_pairs(args, eng, gr, q, want)
        except pkg.CfpqError as e:
            # a saturating closure (the dense-stress grammar: 24 nonterminals
            # close to n^2 pairs each) does not fit as 64-bit pair keys
            log(f"[bench] pairs leg skipped: {e}")
            pairs_run = str(e)
    total_window = None if args.no_total else measure_total(args, eng, gr, mats, init, rules, opts, want)
    mfma_leg = measure_mfma(eng, gr, mats, init, rules, opts, want, step) if (args.config == 2 and not args.no_mfma) else None
    sched_name = {1: "head-sweep fixpoint", 2: "rule-sweep fixpoint"}.get(opts.get("schedule"), "delta fixpoint")
    ok_runs = {k: v for k, v in runs.items() if v[2] and k != "worklist"}
    chosen = min(ok_runs, key=lambda k: ok_runs[k][0]) if ok_runs else "plain"
    ms_per_step, stats, bit_exact = runs[chosen]
    bit_exact = all(v[2] for v in runs.values()) and (total_window is None or total_window["bit_exact"]) and \
        (mfma_leg is None or mfma_leg["bit_exact"])
    fix_ms = [s["fixpoint_ms"] for s in stats]
    value = 2.0 * n ** 3 * a_ref / (ms_per_step / 1e3) / 1e12
    engines = {}
    for k, (ms, sts, ok) in runs.items():
        last = sts[-1]
        if k == "worklist":
            engines[k] = {"ms_per_step": round(ms, 4), "add_mul_calls": last["add_mul_calls"],
                          "runs": len(sts), "bit_exact": ok}
            continue
        engines[k] = {"ms_per_step": round(ms, 4), "fixpoint_ms_device_median": round(
            statistics.median(x["fixpoint_ms"] for x in sts), 4), "iterations": last["iterations"],
            "tile_products": last["tile_products"], "bit_exact": ok}
        if k == "learned":
            engines[k].update({x: last[x] for x in ("sched_scored", "sched_deferred", "sched_full", "flush_steps")})
        if last["unit_clocks"]:
            # what a scheduler that skipped exactly the sparse units adding no
            # bit could save (measured per unit with s_memtime in the kernel)
            cs = last["unit_clocks_idle"] / last["unit_clocks"]
            engines[k]["skip_bound"] = {
                "units_idle_share": round(last["units_idle"] / max(last["units_done"], 1), 4),
                "unit_clock_share": round(cs, 4),
                "b_rows_share": round(last["b_rows_idle"] / max(last["b_rows"], 1), 4),
                "sparse_speedup_bound": round(1.0 / max(1.0 - cs, 1e-9), 3)}
    if "learned" in runs:
        engines["learned_vs_plain_speedup"] = round(runs["plain"][0] / runs["learned"][0], 4)
    if "worklist" in runs:
        engines["plain_vs_worklist_speedup"] = round(runs["worklist"][0] / runs["plain"][0], 3)
    if isinstance(pairs_run, str):
        engines["pairs"] = {"error": pairs_run}
    elif pairs_run is not None:
        p_ms, p_stats, p_ok, p_pairs = pairs_run
        engines["pairs"] = {"ms_per_call": round(p_ms, 4), "fixpoint_ms_device_median": round(
            statistics.median(x["fixpoint_ms"] for x in p_stats), 4), "iterations": p_stats[-1]["iterations"],
            "pairs": p_pairs, "runs": len(p_stats), "bit_exact": p_ok,
            "note": "sparse pair-list formulation; ms_per_call includes host arrays both ways"}
        bit_exact = bit_exact and p_ok

    # the roofline needs per-kernel event timing: the eager plain leg's stats
    roofline = product_roofline(runs["plain"][1], wname)

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline (budget {args.cpu_budget:.0f} s)")
        cpu = cpu_baseline(q, a_ref, args.cpu_budget)

    s0 = runs["plain"][1][-1]  # work counts and per-kernel times: the eager plain leg
    line = {
        "metric": "CFPQ fixpoint time (s) & eff. Tbit-op/s; batched/learned vs plain, bit-exact",
        "value": round(value, 3),
        "unit": "Tbit-op/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, same inputs as the CPU oracle)",
        "config": {"workload": wname, "n": n, "edges": q.n_edges, "nonterminals": gr.nonterminals_count,
                   "rules": len(rules), "A_ref": a_ref, "A_ref_kind": work_key,
                   "engine": {"plain": sched_name,
                              "untimed": sched_name + " without per-kernel timing events",
                              "graph": sched_name + ", steps replayed from captured hipGraphs",
                              "learned": "delta fixpoint + learned tile scheduler"}[chosen],
                   "kernel": args.kernel, "parallelism": "single GPU"},
        "fixpoint_s": round(statistics.median(fix_ms) / 1e3, 6),
        "fixpoint_ms_device_median": round(statistics.median(fix_ms), 4),
        "fixpoint_ms_device_std": round(statistics.pstdev(fix_ms), 4),
        "iterations": s0["iterations"],
        "tile_products": s0["tile_products"],
        "work": {k: s0[k] for k in ("items", "tile_products", "tile_products_dense", "tile_products_mfma",
                                    "tile_products_unchanged",
                                    "a_rows", "a_cols", "b_rows", "b_rows_lds", "m4r_lookups", "mfma_instr", "out_rows",
                                    "new_bits", "product_launches", "host_syncs", "cand_hits",
                                    "max_unit_cands", "a_rows_dense", "b_rows_dense", "out_rows_dense")},
        "kernel_ms": {k: round(s0[k], 4) for k in ("sparse_ms", "merge_ms", "dense_ms", "mfma_ms")},
        "engine_opts": opts,
        "engines": engines,
        "window_total": total_window,
        "mfma_leg": mfma_leg,
        "bit_exact": bit_exact,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    gr.free()
    for c in init:
        c.free()
    eng.close()
    if not bit_exact:
        sys.exit(1)


def bench_nt(args, pkg, gen, cf):
    """The single-matrix formulation on a workload small enough for an n x n
    uint32 matrix: one step = cfpq_nt_closure from the initial nonterminal sets
    (host buffer in and out, so the time includes both PCIe copies)."""
    key, wname = WORKLOADS[args.config]
    golden = json.load(open(GOLDEN))[key]
    q = gen.config(args.config)
    assert q.digest() == golden["digest"], "workload generator drifted from the golden fixture"
    a_ref = int(golden["add_mul_calls"])
    gr = cf.Grammar(q.grammar, text=True)
    graph = cf.Graph.from_query(q)
    init = gr.nt_elements(graph)
    rules = gr.rule_triples()
    eng = pkg.Engine(0)
    for _ in range(args.warmup):
        eng.nt_closure(rules, init.copy())
    eng.sync()
    walls, stats = [], []
    for _ in range(args.steps):
        el = init.copy()
        t0 = time.perf_counter()
        stats.append(eng.nt_closure(rules, el))
        walls.append((time.perf_counter() - t0) * 1e3)
    eng.close()
    want = {k: tuple(v) for k, v in golden["fingerprints"].items()}
    bit_exact = gr.nt_results(el) == want
    ms = statistics.median(walls)
    line = {
        "metric": "CFPQ fixpoint time (s) & eff. Tbit-op/s; batched/learned vs plain, bit-exact",
        "value": round(2.0 * q.n ** 3 * a_ref / (ms / 1e3) / 1e12, 3), "unit": "Tbit-op/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded generator, same inputs as the CPU oracle)",
        "config": {"workload": wname, "n": q.n, "edges": q.n_edges, "nonterminals": gr.nonterminals_count,
                   "rules": len(rules), "A_ref": a_ref, "engine": "single matrix of nonterminal sets",
                   "parallelism": "single GPU"},
        "fixpoint_ms_device_median": round(statistics.median(s["fixpoint_ms"] for s in stats), 4),
        "iterations": stats[-1]["iterations"], "bit_exact": bit_exact,
        "note": "ms_per_step includes the host-to-device and device-to-host copies of the n x n matrix",
    }
    print(json.dumps(line), flush=True)
    if not bit_exact:
        sys.exit(1)


def bench_pairs(args, pkg, gen, cf):
    """The sparse pair-list formulation on the same workload: one step =
    cfpq_pairs_fixpoint from the initial pairs (host arrays in and out, so the
    time includes both PCIe copies and the host-side key check).  With
    --copies K the graph is K disjoint copies of the workload (vertex ids of
    copy c shifted by c * n), one fixpoint over K * n vertices; its result
    must be the K shifted copies of the single-copy result, which is checked
    against the golden fingerprints first."""
    key, wname = WORKLOADS[args.config]
    golden = json.load(open(GOLDEN))[key]
    q = gen.config(args.config)
    assert q.digest() == golden["digest"], "workload generator drifted from the golden fixture"
    a_ref = int(golden["add_mul_calls"])
    gr = cf.Grammar(q.grammar, text=True)
    keys = gr.pair_keys(cf.Graph.from_query(q))
    rules = gr.rule_triples()
    eng = pkg.Engine(0)
    single, _ = eng.pairs_fixpoint(rules, keys, q.n)
    want = {k: tuple(v) for k, v in golden["fingerprints"].items()}
    bit_exact = gr.pair_results(single, q.n) == want
    K = max(1, args.copies)
    n = q.n * K
    if n > 1 << 31:
        raise SystemExit(f"--copies {K}: {n} vertices (at most 2^31)")
    shift = [np.uint64(c * q.n * ((1 << 32) + 1)) for c in range(K)]
    if K > 1:
        keys = [np.concatenate([k + d for d in shift]) for k in keys]
    out = None
    for _ in range(args.warmup):
        out, _ = eng.pairs_fixpoint(rules, keys, n)
    cap = max(len(k) for k in out) if out else None
    eng.sync()
    walls, stats = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out, st = eng.pairs_fixpoint(rules, keys, n, cap=cap)
        walls.append((time.perf_counter() - t0) * 1e3)
        stats.append(st)
    eng.close()
    bit_exact = bit_exact and all(np.array_equal(o, np.concatenate([k + d for d in shift]))
                                  for o, k in zip(out, single))
    ms = statistics.median(walls)
    line = {
        "metric": "CFPQ fixpoint time (s) & eff. Tbit-op/s; batched/learned vs plain, bit-exact",
        "value": round(K * 2.0 * q.n ** 3 * a_ref / (ms / 1e3) / 1e12, 3), "unit": "Tbit-op/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded generator, same inputs as the CPU oracle)",
        "config": {"workload": wname if K == 1 else f"{K} disjoint copies of {wname}", "n": n,
                   "edges": q.n_edges * K, "nonterminals": gr.nonterminals_count, "rules": len(rules),
                   "A_ref": a_ref, "engine": "sorted pair lists", "parallelism": "single GPU"},
        "fixpoint_ms_device_median": round(statistics.median(s["fixpoint_ms"] for s in stats), 4),
        "iterations": stats[-1]["iterations"], "pairs": int(sum(len(k) for k in out)),
        "join_outputs": stats[-1]["items"], "host_syncs": stats[-1]["host_syncs"], "bit_exact": bit_exact,
        "note": "ms_per_step includes the copies of the pair arrays both ways and the host's key check; "
                "value counts 2 n^3 A_ref bit-ops per copy",
    }
    print(json.dumps(line), flush=True)
    if not bit_exact:
        sys.exit(1)


def bench_batch(args, pkg, gen, cf):
    """512 independent queries (log-uniform N in [256, 8192], Dyck or SG
    family): one batched fixpoint per step vs the same queries one call each."""
    qs = gen.config(3)
    allg = json.load(open(GOLDEN))
    golden = allg.get("config3") or allg["config3_first16"]
    for q, g in zip(qs, golden):
        assert q.digest() == g["digest"], "workload generator drifted from the golden fixture"
    tune = importlib.import_module("cfpq-on-gpgpu_amd.tune")
    if args.tune:
        tune.tune_batch(pkg, cf, qs, repeats=1, log=log, passes=args.tune_passes,
                        dims=[d for d in args.tune_dims.split(",") if d])
        os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
        with open(tune.CACHE) as f, open(os.path.join(ROOT, "gpurun_out", "tuned.json"), "w") as g:
            g.write(f.read())
    opts = tune.tuned_batch_opts(qs)
    opts["kernel"] = args.kernel
    if args.iters_per_sync:
        opts["iters_per_sync"] = args.iters_per_sync
    if args.sparse_mode >= 0:
        opts["sparse_mode"] = args.sparse_mode
    if args.sparse_waves:
        opts["sparse_waves_per_cu"] = args.sparse_waves
    if args.graph:
        opts["use_graph"] = True
    log(f"[bench] batch of {len(qs)} queries, N in [{min(q.n for q in qs)}, {max(q.n for q in qs)}], "
        f"engine options {opts}")
    eng = pkg.Engine(0, **opts)
    grs, init = [], []
    for q in qs:
        gr = cf.Grammar(q.grammar, text=True)
        mats = gr.load(eng, cf.Graph.from_query(q))
        cp = []
        for m in mats:
            c = eng.matrix(q.n)
            c.copy_from(m)
            cp.append(c)
        grs.append(gr)
        init.append(cp)
    work = [(gr.rule_triples(), gr.matrices) for gr in grs]

    def restore():
        for gr, cp in zip(grs, init):
            for m, c in zip(gr.matrices, cp):
                m.copy_from(c)
        eng.sync()

    def timed(fn):
        restore()
        t0 = time.perf_counter()
        st = fn()
        eng.sync()
        return (time.perf_counter() - t0) * 1e3, st

    batched = lambda: eng.batch_fixpoint(work)
    for w in range(args.warmup):
        ms, st = timed(batched)
        log(f"[bench] warmup {w}: {ms:.2f} ms, {st['iterations']} steps")
    runs = [timed(batched) for _ in range(args.steps)]
    ms_per_step = sum(r[0] for r in runs) / len(runs)
    st = runs[-1][1]
    got = [gr.results() for gr in grs]
    want = [{k: tuple(v) for k, v in g["fingerprints"].items()} for g in golden]
    bit_exact = all(a == b for a, b in zip(got, want))
    checked = min(len(got), len(want))

    def sequential():
        tot = {"fixpoint_ms": 0.0, "iterations": 0}
        for rules, mats in work:
            s1 = eng.fixpoint(rules, mats)
            tot["fixpoint_ms"] += s1["fixpoint_ms"]
            tot["iterations"] += s1["iterations"]
        return tot
    seq_ms, seq = timed(sequential)
    bit_exact = bit_exact and all(gr.results() == w for gr, w in zip(grs, want))
    log(f"[bench] batched {ms_per_step:.2f} ms/step, sequential {seq_ms:.2f} ms, bit_exact {bit_exact} "
        f"({checked} queries checked against the oracle fixture)")
    line = {
        "metric": "CFPQ fixpoint time (s) & eff. Tbit-op/s; batched/learned vs plain, bit-exact",
        "value": round(len(qs) / (ms_per_step / 1e3), 3),
        "unit": "queries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded generator, same inputs as the CPU oracle)",
        "config": {"workload": "batch512_n256_8192_dyck_sg", "queries": len(qs), "engine": "batched delta fixpoint",
                   "graph": bool(opts.get("use_graph")), "parallelism": "single GPU"},
        "engine_opts": opts,
        "batched_fixpoint_ms_device": round(st["fixpoint_ms"], 4),
        "batched_steps": st["iterations"],
        "sequential_ms": round(seq_ms, 4),
        "sequential_fixpoint_ms_device": round(seq["fixpoint_ms"], 4),
        "sequential_steps_total": seq["iterations"],
        "speedup_batched_vs_sequential": round(seq_ms / ms_per_step, 3),
        "bit_exact": bit_exact,
        "queries_checked": checked,
    }
    print(json.dumps(line), flush=True)
    for gr in grs:
        gr.free()
    for cp in init:
        for c in cp:
            c.free()
    eng.close()
    if not bit_exact:
        sys.exit(1)


if __name__ == "__main__":
    main()
