"""Autotuner host logic (no GPU): the search, the cache keys, and that the
committed cache covers the benchmark workloads."""
import importlib

import pytest

tune = importlib.import_module("cfpq-on-gpgpu_amd.tune")
gen = importlib.import_module("cfpq-on-gpgpu_amd.gen")


def test_chunk_sizes_divide_a_claim_block():
    # the sparse kernel hands out claims in blocks of 64 units per cursor shard
    chunks = dict(tune.SPACE)["sparse_chunk"]
    assert all(64 % c == 0 for c in chunks)
    assert 64 % tune.DEFAULTS["sparse_chunk"] == 0


def test_descent_keeps_only_clear_gains():
    # a synthetic cost: waves matter (best 8), chunk 4 is 1% better (below the 2% bar)
    def cost(o):
        return 10.0 + abs(o["sparse_waves_per_cu"] - 8) - (0.1 if o["sparse_chunk"] == 4 else 0.0)

    best, best_ms, trace = tune._descend(cost, log=lambda *a: None, passes=2)
    assert best["sparse_waves_per_cu"] == 8
    assert best_ms == min(t["ms"] for t in trace if t["opts"] == best)
    assert len(trace) > 1


def test_cache_covers_bench_workloads():
    cache = tune.load_cache()
    q4 = gen.config(4)
    assert tune.workload_key(q4.n, q4.grammar) in cache
    assert tune.tuned_opts(q4.n, q4.grammar)["sparse_waves_per_cu"] >= 1
    qs = gen.config(3)
    key = tune.batch_key(qs)
    assert key.startswith(f"batch{len(qs)}:") and key in cache
    assert tune.tuned_batch_opts(qs) == cache[key]["opts"]


def test_keys_are_stable():
    q = gen.config(0)
    assert tune.workload_key(q.n, q.grammar) == tune.workload_key(q.n, q.grammar)
    assert tune.workload_key(1024, q.grammar).startswith("n2^10:")
    assert tune.batch_key([q]) != tune.batch_key([q, q])


@pytest.mark.parametrize("n", [1, 255, 256, 257, 65536])
def test_workload_class_by_log2(n):
    assert tune.workload_key(n, "S a\n").startswith("n2^")
