"""The learned tile scheduler on the GPU (needs a GPU), through the C ABI.

  * the HIP forward / backward kernels against the fp64 CPU statement of the
    network (oracle/mlp.py) at the same bf16 parameters;
  * training on the device lowers the loss and learns a separable rule;
  * traces recorded by the plain fixpoint are well formed;
  * with the scheduler deferring work, every fixpoint stays bit-exact.
"""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def _problem(seed, n):
    rng = np.random.default_rng(seed)
    w = mlp.f32_to_bf16(rng.normal(0, 0.3, mlp.PARAMS).astype(np.float32))
    x = rng.random((n, mlp.F)).astype(np.float32)
    rule = rng.integers(0, 70, n).astype(np.uint32)
    y = (rng.random(n) < 0.3).astype(np.float32)
    return w, x, rule, y


@pytest.mark.parametrize("n", [1, 63, 1000, 5000])
def test_forward_and_gradient_match_fp64(engine, n):
    w, x, rule, y = _problem(n, n)
    logits, loss, grad = engine.sched_eval(w, x, rule, y)
    p = mlp.bf16_to_f64(w)
    want_loss, want_grad, want_z = mlp.loss_and_grad(p, x.astype(np.float64), rule, y.astype(np.float64))
    # fp32 arithmetic against fp64: tolerance scaled by the magnitudes involved
    assert np.allclose(logits, want_z, rtol=1e-4, atol=1e-4)
    assert abs(loss - want_loss) <= 1e-5 + 1e-5 * abs(want_loss)
    scale = np.abs(want_grad).max()
    assert np.abs(grad - want_grad).max() <= 1e-4 * scale + 1e-7
    # every block of the layout receives gradient where fp64 says so
    for a, b in [(mlp.OFF_E, mlp.OFF_W1), (mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_W2, mlp.OFF_B2), (mlp.OFF_W3, mlp.OFF_B3)]:
        assert np.array_equal(grad[a:b] != 0, want_grad[a:b] != 0) or n >= 1000


def test_training_learns_a_separable_rule(engine):
    rng = np.random.default_rng(5)
    n = 20000
    x = rng.random((n, mlp.F)).astype(np.float32)
    rule = rng.integers(0, 6, n).astype(np.uint32)
    y = ((x[:, 0] + 0.5 * x[:, 3] > 0.8) ^ (rule == 2)).astype(np.float32)
    engine.sched_init(1)
    losses = engine.sched_train(x, rule, y, epochs=12, batch=1024, lr=3e-3, seed=1)
    assert losses[-1] < 0.5 * losses[0]
    logits, _, _ = engine.sched_eval(engine.sched_get_weights(), x, rule, y, want_grad=False)
    acc = np.mean((logits > 0) == (y > 0.5))
    assert acc > 0.95, acc


def test_trace_rows(pkg, gen, engine):
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    q = gen.dense_stress(900)
    engine.trace_begin(1 << 20)
    gr = cf.Grammar(q.grammar, text=True)
    gr.intersection_with_graph(cf.Graph.from_query(q), engine)
    x, rule, y, count = engine.trace_read(1 << 20)
    engine.trace_end()
    gr.free()
    assert count == x.shape[0] > 0
    assert set(np.unique(y)) <= {0.0, 1.0} and 0 < y.mean() < 1
    assert np.isfinite(x).all() and (x >= 0).all()
    assert rule.max() < len(q.grammar.splitlines())


@pytest.mark.parametrize("threshold", [0.3, 0.9, 0.999])
@pytest.mark.parametrize("which", ["dense", "dyck", "sg"])
def test_scheduled_fixpoint_is_bit_exact(pkg, gen, oracle, which, threshold):
    """Deferral on (an untrained network and thresholds that defer much of the
    work): results must not change."""
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    q = {"dense": lambda: gen.dense_stress(1200), "dyck": lambda: gen.dyck_er(1500, 5000, seed=6),
         "sg": lambda: gen.sg_forest(6000, seed=7, types_per_vertex=3)}[which]()
    g_ref, mats_ref, _ = oracle.solve(q, "seminaive")
    want = oracle.fingerprints(g_ref, mats_ref, q.n)
    for graph in (False, True):
        with pkg.Engine(0, sched=True, sched_threshold=threshold, use_graph=graph) as eng:
            eng.sched_init(11)
            gr = cf.Grammar(q.grammar, text=True)
            _, _, st = gr.intersection_with_graph(cf.Graph.from_query(q), eng)
            assert gr.results() == want, (which, threshold, graph)
            assert st["sched_scored"] > 0
            if st["sched_deferred"]:
                assert st["flush_steps"] >= 1 and st["sched_full"] > 0
            gr.free()


def test_scheduled_fixpoint_with_final_full_pass(pkg, gen, oracle):
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    q = gen.dense_stress(800)
    g_ref, mats_ref, _ = oracle.solve(q, "seminaive")
    with pkg.Engine(0, sched=True, sched_threshold=0.9, final_full_pass=True) as eng:
        eng.sched_init(3)
        gr = cf.Grammar(q.grammar, text=True)
        _, _, st = gr.intersection_with_graph(cf.Graph.from_query(q), eng)
        assert gr.results() == oracle.fingerprints(g_ref, mats_ref, q.n)
        assert st["full_passes"] == 1 and st["sched_deferred"] > 0
        gr.free()


def test_weights_roundtrip_and_errors(pkg, engine):
    w = mlp.f32_to_bf16(np.linspace(-1, 1, mlp.PARAMS).astype(np.float32))
    engine.sched_set_weights(w)
    assert np.array_equal(engine.sched_get_weights(), w)
    with pytest.raises(pkg.CfpqError):
        engine.sched_set_weights(w[:10])


def _auc(score, y):
    pos, neg = score[y > 0.5], score[y <= 0.5]
    order = np.argsort(np.concatenate([pos, neg]), kind="stable")
    ranks = np.empty(len(order))
    ranks[order] = np.arange(1, len(order) + 1)
    return (ranks[:len(pos)].sum() - len(pos) * (len(pos) + 1) / 2) / max(len(pos) * len(neg), 1)


def test_trained_network_gains_over_untrained(pkg, gen, engine):
    """What the trained network gains over the scheduler without it, on
    workloads it was not trained on (the forest generator at another seed and
    size, a dense-stress grammar of another seed).  Trace rows come from the
    plain fixpoint; each network defers its lowest-scored 10% of the items.
    The trained one must rank far better than an untrained network of the
    same shape and defer far fewer items that would have added bits (every
    such item is later re-evaluated over full operands)."""
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    sched = __import__("importlib").import_module("cfpq-on-gpgpu_amd.sched")
    trained = sched.load()
    assert trained is not None, "trained scheduler weights are missing"
    held_out = [gen.sg_forest_dense_types(16384, seed=4242), gen.dense_stress(3000, seed=77)]
    xs, rs, ys = sched.collect(pkg, cf, held_out, log=lambda *a: None)
    engine.sched_init(19)
    untrained = engine.sched_get_weights()
    res = {}
    for name, w in (("trained", trained), ("untrained", untrained)):
        logits, _, _ = engine.sched_eval(w, xs, rs, ys, want_grad=False)
        cut = np.quantile(logits, 0.10)
        deferred = logits <= cut
        res[name] = (_auc(logits, ys), float(ys[deferred].mean()))
    print("held-out rows", len(ys), "positive rate", float(ys.mean()), res)
    auc_t, waste_t = res["trained"]
    auc_u, waste_u = res["untrained"]
    assert auc_t > 0.85 and auc_t > auc_u + 0.2, res
    # share of the deferred items that were active: the base rate for a
    # scorer that knows nothing is the positive rate itself
    assert waste_t < 0.5 * float(ys.mean()) and waste_t < 0.5 * waste_u, res


def test_learned_sg_forest_65536_golden(pkg, gen):
    """The learned engine at full size: the 65,536-vertex ontology query with
    the trained network ordering and deferring (rule, tile) work, the tuned
    options of the workload class, and the exact full pass that certifies the
    fixpoint.  Every nonterminal's fingerprint and count against the oracle's
    golden fixture; the certification pass runs once and adds no bit."""
    import json
    import os
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    sched = __import__("importlib").import_module("cfpq-on-gpgpu_amd.sched")
    tune = __import__("importlib").import_module("cfpq-on-gpgpu_amd.tune")
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixpoints.json")))["config4"]
    q = gen.config(4)
    assert q.digest() == golden["digest"]
    weights = sched.load()
    assert weights is not None, "trained scheduler weights are missing"
    opts = dict(tune.tuned_opts(q.n, q.grammar), sched=True, final_full_pass=True)
    with pkg.Engine(0, **opts) as eng:
        eng.sched_set_weights(weights)
        gr = cf.Grammar(q.grammar, text=True)
        _, _, st = gr.intersection_with_graph(cf.Graph.from_query(q), eng)
        assert gr.results() == {k: tuple(v) for k, v in golden["fingerprints"].items()}
        assert st["full_passes"] == 1
        assert st["sched_scored"] > 0
        gr.free()


def test_learned_sg_forest_65536_repeatable(pkg, gen):
    """The deferral path on the 65,536-vertex forest, three times under
    different launch shapes, each against the golden fingerprints.  (It once
    lost bits from run to run: the retire launch read head row masks through
    the scalar cache, which could hold them from an earlier launch.)"""
    import json
    import os
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    sched = __import__("importlib").import_module("cfpq-on-gpgpu_amd.sched")
    tune = __import__("importlib").import_module("cfpq-on-gpgpu_amd.tune")
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixpoints.json")))["config4"]
    want = {k: tuple(v) for k, v in golden["fingerprints"].items()}
    q = gen.config(4)
    weights = sched.load()
    assert weights is not None, "trained scheduler weights are missing"
    for waves in (8, 12, 16):
        opts = dict(tune.tuned_opts(q.n, q.grammar), sched=True, sparse_waves_per_cu=waves)
        with pkg.Engine(0, **opts) as eng:
            eng.sched_set_weights(weights)
            gr = cf.Grammar(q.grammar, text=True)
            _, _, st = gr.intersection_with_graph(cf.Graph.from_query(q), eng)
            assert gr.results() == want, waves
            assert st["sched_deferred"] > 0
            gr.free()


@pytest.mark.parametrize("which", ["dense", "dyck", "sg", "closure"])
def test_ordered_fixpoint_is_bit_exact(pkg, gen, oracle, which):
    """The scheduler's ranking (its predicted-active items evaluated in a
    first phase of each delta step, the rest in a second) with trained and
    untrained weights, with and without the fused head update, alone and
    with deferral: the least fixpoint is the same."""
    import numpy as np
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    sched = __import__("importlib").import_module("cfpq-on-gpgpu_amd.sched")
    if which == "closure":
        rng = np.random.default_rng(81)
        n, m = 1500, 2200
        src, dst = rng.integers(0, n, m).astype(np.uint32), rng.integers(0, n, m).astype(np.uint32)
        q = gen.Query("closure", n, "S S S\nS a\n", src, dst, np.zeros(m, np.uint8), ("a",))
    else:
        q = {"dense": lambda: gen.dense_stress(1200), "dyck": lambda: gen.dyck_er(1500, 5000, seed=6),
             "sg": lambda: gen.sg_forest(6000, seed=7, types_per_vertex=3)}[which]()
    g_ref, mats_ref, _ = oracle.solve(q, "seminaive")
    want = oracle.fingerprints(g_ref, mats_ref, q.n)
    trained = sched.load()
    for weights in ("trained", "untrained"):
        for fuse, defer in ((True, False), (False, False), (True, True)):
            with pkg.Engine(0, sched_order=True, fuse_apply=fuse, sched=defer, sched_threshold=0.3) as eng:
                if weights == "trained":
                    eng.sched_set_weights(trained)
                else:
                    eng.sched_init(5)
                gr = cf.Grammar(q.grammar, text=True)
                _, _, st = gr.intersection_with_graph(cf.Graph.from_query(q), eng)
                assert gr.results() == want, (which, weights, fuse, defer)
                assert st["sched_scored"] > 0
                gr.free()
