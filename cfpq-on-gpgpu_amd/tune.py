"""Autotuner for the fixpoint engine's launch and crossover parameters.

For a workload it times complete fixpoints (device time from the engine's own
hipEvents, median of a few runs) while walking the parameters one at a time
(coordinate descent, best value kept before moving to the next parameter):

  schedule             semi-naive delta steps over all rules (0), head sweeps (1) or rule sweeps (2) over full matrices
  sparse_mode          sparse product and its work-tile shape: lane-parallel candidates over
                       work units of 64 (0), 128 (3) or 256 (2) output rows of a 256-column
                       tile, or row-per-lane (1)
  sparse_batch         k-tiles whose loads the row-per-lane sparse product issues together
  sparse_waves_per_cu  persistent waves of the sparse product per CU
  sparse_chunk         work units a sparse-product wave claims per atomic
  dense_bits           A-tile popcount from which a tile goes to the LDS kernel
  m4r_bits             A-tile popcount from which the Four Russians tables are used
  mfma_bits            A-tile popcount from which a tile goes to the int8 MFMA kernel
  iters_per_sync       fixpoint steps enqueued per host read of the counters
  fuse_apply           merge new rows into the head matrices inside the products
  split_cands          candidate k-tiles above which a quarter is split into sub-units
  split_rows           split quarters shared out by candidates (merge pass) or by rows
  split_piece          candidates per sub-unit of a split quarter (0: split_cands)

Every candidate is checked against the first run's fingerprints, so a tuning
run is also a parity run.  Results are cached in ``tuned.json`` next to this
file, keyed by workload class (log2 of N, grammar family), and
:func:`tuned_opts` returns the cached choice for a query.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import statistics

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuned.json")

SPACE = [
    ("schedule", [0, 1, 2]),
    ("sparse_mode", [0, 1, 2, 3]),
    ("sparse_batch", [1, 2, 4]),
    ("sparse_waves_per_cu", [2, 4, 8, 12, 16, 32]),
    ("sparse_chunk", [1, 2, 4, 8, 16]),
    ("dense_bits", [512, 2048, 8192, 0xFFFFFFF0]),
    ("m4r_bits", [8192, 16384, 32768, 0xFFFFFFF0]),
    ("mfma_bits", [8192, 16384, 32768, 0xFFFFFFF0]),
    ("iters_per_sync", [2, 4, 8]),
    ("fuse_apply", [0, 1]),
    ("split_cands", [32, 64, 128, 0xFFFFFFFF]),
    ("split_rows", [0, 1]),
    ("split_piece", [0, 16, 32, 64]),
]
# Graph replay (use_graph) is a dimension for batches (tune_batch), where the
# steps are short and launch-bound; for single queries the caller decides, as
# per-kernel event timing needs the eager launch path.
DEFAULTS = {"schedule": 0, "sparse_mode": 0, "sparse_batch": 2, "sparse_waves_per_cu": 8, "sparse_chunk": 4, "dense_bits": 2048, "m4r_bits": 16384,
            "mfma_bits": 32768, "iters_per_sync": 4, "fuse_apply": 0, "split_cands": 64, "split_rows": 0, "split_piece": 0}


def workload_key(n: int, grammar_text: str) -> str:
    fam = hashlib.sha1(" ".join(sorted(grammar_text.split())).encode()).hexdigest()[:8]
    return f"n2^{int(round(math.log2(max(n, 2))))}:{fam}"


def load_cache() -> dict:
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            return json.load(f)
    return {}


def tuned_opts(n: int, grammar_text: str) -> dict:
    """Cached parameters for this workload class (library defaults if none)."""
    ent = load_cache().get(workload_key(n, grammar_text))
    return dict(ent["opts"]) if ent else {}


def batch_key(queries) -> str:
    h = hashlib.sha1("".join(q.digest() for q in queries).encode()).hexdigest()[:8]
    return f"batch{len(queries)}:{h}"


def _descend(run, log, passes: int, space=None, start=None):
    """Coordinate descent from DEFAULTS (or `start`) over `space` (default
    SPACE); returns (best opts, ms, trace)."""
    space = SPACE if space is None else space
    best = dict(DEFAULTS if start is None else start)
    best_ms = run(best)
    log(f"[tune] start {best} -> {best_ms:.3f} ms")
    trace = [{"opts": dict(best), "ms": best_ms}]
    for _ in range(passes):
        for name, values in space:
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best)
                cand[name] = v
                ms = run(cand)
                trace.append({"opts": dict(cand), "ms": ms})
                log(f"[tune] {name}={v} -> {ms:.3f} ms")
                if ms < best_ms * 0.98:
                    best, best_ms = cand, ms
    return best, best_ms, trace


def _store(key, name, best, best_ms, trace, log):
    cache = load_cache()
    cache[key] = {"workload": name, "opts": best, "ms": round(best_ms, 4), "trace": trace}
    with open(CACHE, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    log(f"[tune] best {best} -> {best_ms:.3f} ms (cached as {key})")


def _space_and_start(dims, cached: dict, base: dict):
    """With `dims` (parameter names), walk only those, starting from the
    cached choice completed by `base`; otherwise the whole space from `base`."""
    if not dims:
        return None, None
    space = [(n, v) for n, v in SPACE + [("use_graph", [0, 1])] if n in dims]
    return space, dict(base, **cached)


def tune(pkg, cf, query, repeats: int = 3, log=print, passes: int = 1, dims=None) -> dict:
    """Tune on `query` (a gen.Query); returns the best option dict and stores it.
    `passes` > 1 walks the parameters again from the best point found; `dims`
    restricts the walk to those parameters, from the cached choice."""
    eng = pkg.Engine(0)
    gr = cf.Grammar(query.grammar, text=True)
    mats = gr.load(eng, cf.Graph.from_query(query))
    init = []
    for m in mats:
        c = eng.matrix(query.n)
        c.copy_from(m)
        init.append(c)
    rules = gr.rule_triples()
    ref = None

    def run(opts) -> float:
        nonlocal ref
        # timed as the engine runs without per-kernel timing events
        eng.set_opts(**opts, no_kernel_timing=True)
        times = []
        for _ in range(repeats):
            for m, c in zip(mats, init):
                m.copy_from(c)
            st = eng.fixpoint(rules, mats)
            times.append(st["fixpoint_ms"])
        res = gr.results()
        if ref is None:
            ref = res
        elif res != ref:
            raise AssertionError(f"tuning candidate {opts} changed the result")
        return statistics.median(times)

    space, start = _space_and_start(dims, tuned_opts(query.n, query.grammar), DEFAULTS)
    best, best_ms, trace = _descend(run, log, passes, space=space, start=start)
    for c in init:
        c.free()
    gr.free()
    eng.close()
    _store(workload_key(query.n, query.grammar), query.name, best, best_ms, trace, log)
    return best


def tune_batch(pkg, cf, queries, repeats: int = 1, log=print, passes: int = 1, dims=None) -> dict:
    """Tune the batched fixpoint of `queries` (gen.Query list); cached under batch_key."""
    eng = pkg.Engine(0)
    grs, inits = [], []
    for q in queries:
        gr = cf.Grammar(q.grammar, text=True)
        mats = gr.load(eng, cf.Graph.from_query(q))
        cp = []
        for m in mats:
            c = eng.matrix(q.n)
            c.copy_from(m)
            cp.append(c)
        grs.append(gr)
        inits.append(cp)
    work = [(gr.rule_triples(), gr.matrices) for gr in grs]
    ref = None

    def run(opts) -> float:
        nonlocal ref
        eng.set_opts(**opts, no_kernel_timing=True)
        times = []
        for _ in range(repeats):
            for gr, cp in zip(grs, inits):
                for m, c in zip(gr.matrices, cp):
                    m.copy_from(c)
            times.append(eng.batch_fixpoint(work)["fixpoint_ms"])
        res = [gr.results() for gr in grs]
        if ref is None:
            ref = res
        elif res != ref:
            raise AssertionError(f"tuning candidate {opts} changed a result")
        return statistics.median(times)

    space, start = _space_and_start(dims, tuned_batch_opts(queries), dict(DEFAULTS, use_graph=0))
    best, best_ms, trace = _descend(run, log, passes, space=space or SPACE + [("use_graph", [0, 1])],
                                    start=start or dict(DEFAULTS, use_graph=0))
    for gr, cp in zip(grs, inits):
        for c in cp:
            c.free()
        gr.free()
    eng.close()
    _store(batch_key(queries), f"batch of {len(queries)} queries", best, best_ms, trace, log)
    return best


def tuned_batch_opts(queries) -> dict:
    ent = load_cache().get(batch_key(queries))
    return dict(ent["opts"]) if ent else {}
