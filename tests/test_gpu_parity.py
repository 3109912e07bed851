"""Parity of the HIP engine with the CPU oracle, through the C ABI (needs a GPU).

Bar: bit-exact.  Every product and fixpoint result is compared word for word
(or by the reference fingerprint plus popcount at sizes where a full dump is
large) with the oracle on the same seeded inputs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KERNELS = ["gather", "lds", "auto", "m4r", "mfma"]


def rand_mat(rng, n, density):
    return rng.random((n, n)) < density


def upload(engine, oracle, d):
    m = engine.matrix(d.shape[0])
    m.upload(oracle.from_dense(d))
    return m


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("n", [1, 63, 64, 255, 256, 257, 600])
def test_add_mul_random(pkg, oracle, n, kernel):
    eng = pkg.Engine(0, kernel=kernel)
    try:
        rng = np.random.default_rng(n * 7 + len(kernel))
        for density in (0.0, 0.003, 0.05, 0.5, 1.0):
            a, b, c = rand_mat(rng, n, density), rand_mat(rng, n, density), rand_mat(rng, n, 0.01)
            A, B, C = oracle.from_dense(a), oracle.from_dense(b), oracle.from_dense(c)
            want = C.copy()
            want_changed = oracle.addmul_m4rm(want, A, B, n)
            dA, dB, dC = upload(eng, oracle, a), upload(eng, oracle, b), upload(eng, oracle, c)
            changed = eng.add_mul(dC, dA, dB)
            got = dC.download()
            assert np.array_equal(got, want), (n, density, kernel)
            assert changed == want_changed
            # metadata stays consistent: popcount and fingerprint agree with the oracle
            assert dC.popcount() == int(np.unpackbits(want.view(np.uint8)).sum())
            assert dC.fingerprint() == oracle.fingerprint(want, n)
            for m in (dA, dB, dC):
                m.free()
    finally:
        eng.close()


@pytest.mark.parametrize("kernel", ["auto", "mfma"])
@pytest.mark.parametrize("alias", ["c=a", "c=b", "c=a=b"])
def test_add_mul_aliasing(pkg, oracle, kernel, alias):
    n = 300
    rng = np.random.default_rng(5)
    c, o = rand_mat(rng, n, 0.02), rand_mat(rng, n, 0.02)
    C, O = oracle.from_dense(c), oracle.from_dense(o)
    want = C.copy()
    if alias == "c=a":
        oracle.addmul_m4rm(want, C.copy(), O, n)
    elif alias == "c=b":
        oracle.addmul_m4rm(want, O, C.copy(), n)
    else:
        oracle.addmul_m4rm(want, C.copy(), C.copy(), n)
    with pkg.Engine(0, kernel=kernel) as engine:
        dC, dO = upload(engine, oracle, c), upload(engine, oracle, o)
        if alias == "c=a":
            engine.add_mul(dC, dC, dO)
        elif alias == "c=b":
            engine.add_mul(dC, dO, dC)
        else:
            engine.add_mul(dC, dC, dC)
        assert np.array_equal(dC.download(), want)
        dC.free()
        dO.free()


def test_upload_download_roundtrip(pkg, oracle, engine):
    rng = np.random.default_rng(1)
    for n in (1, 5, 64, 100, 256, 513, 1000):
        d = rand_mat(rng, n, 0.1)
        m = upload(engine, oracle, d)
        assert np.array_equal(oracle.to_dense(m.download(), n), d)
        m.free()


def test_set_bits(pkg, oracle, engine):
    rng = np.random.default_rng(2)
    n = 777
    r = rng.integers(0, n, 5000)
    c = rng.integers(0, n, 5000)
    m = engine.matrix(n)
    m.set_bits(r, c)
    m.set_bits(r[:10], c[:10])  # idempotent
    want = oracle.zeros(n)
    oracle.set_bits(want, r, c)
    assert np.array_equal(m.download(), want)
    assert m.fingerprint() == oracle.fingerprint(want, n)
    with pytest.raises(pkg.CfpqError):
        m.set_bits([n], [0])
    m.free()


def _fixpoint_vs_oracle(pkg, gen, oracle, q, method="seminaive", **opts):
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    g_ref, mats_ref, info = oracle.solve(q, method)
    want = oracle.fingerprints(g_ref, mats_ref, q.n)
    eng = pkg.Engine(0, **opts)
    try:
        gr = cf.Grammar(q.grammar, text=True)
        total_ms, algo_ms, stats = gr.intersection_with_graph(cf.Graph.from_query(q), eng)
        got = gr.results()
        assert got == want
        # word-exact on sizes where a dump is cheap
        if q.n <= 4096:
            for name, idx in g_ref.nt_index.items():
                assert np.array_equal(gr.matrices[gr.nonterminal_to_index[name]].download(), mats_ref[idx]), name
        gr.free()
        return stats
    finally:
        eng.close()


@pytest.mark.parametrize("kernel", ["gather", "lds", "auto"])
def test_fixpoint_dyck_er_1024(pkg, gen, oracle, kernel):
    stats = _fixpoint_vs_oracle(pkg, gen, oracle, gen.config(0), kernel=kernel)
    assert stats["iterations"] >= 2


def test_fixpoint_without_kernel_timing(pkg, gen, oracle):
    """no_kernel_timing: the same fixpoint, with no per-kernel times but the
    whole call still timed."""
    for schedule in ("delta", "rules"):
        stats = _fixpoint_vs_oracle(pkg, gen, oracle, gen.config(0), no_kernel_timing=True, schedule=schedule)
        assert stats["sparse_ms"] == 0 and stats["dense_ms"] == 0 and stats["product_ms"] == 0
        assert stats["fixpoint_ms"] > 0


def test_fixpoint_sg_forest_16384(pkg, gen, oracle):
    _fixpoint_vs_oracle(pkg, gen, oracle, gen.config(1))


@pytest.mark.parametrize("n", [512, 1500])
@pytest.mark.parametrize("kernel", ["auto", "mfma", "m4r"])
def test_fixpoint_dense_stress_scaled(pkg, gen, oracle, n, kernel):
    stats = _fixpoint_vs_oracle(pkg, gen, oracle, gen.dense_stress(n), kernel=kernel)
    if kernel == "mfma":
        assert stats["tile_products_mfma"] == stats["tile_products"] > 0
        assert stats["mfma_instr"] > 0


def test_fixpoint_mixed_paths(pkg, gen, oracle):
    # thresholds that send some items to each of the three product kernels
    stats = _fixpoint_vs_oracle(pkg, gen, oracle, gen.dense_stress(1500), kernel="auto", dense_bits=64,
                                m4r_bits=8192, mfma_bits=30000)
    assert stats["tile_products_mfma"] > 0 and stats["tile_products_dense"] > 0
    assert stats["tile_products"] > stats["tile_products_mfma"] + stats["tile_products_dense"]


def test_fixpoint_dense_stress_4096_golden(pkg, gen):
    import json
    import os
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixpoints.json")))["config2_n4096"]
    q = gen.dense_stress(4096)
    assert q.digest() == golden["digest"]
    for kernel in ("auto", "mfma"):
        with pkg.Engine(0, kernel=kernel) as eng:
            gr = cf.Grammar(q.grammar, text=True)
            gr.intersection_with_graph(cf.Graph.from_query(q), eng)
            assert gr.results() == {k: tuple(v) for k, v in golden["fingerprints"].items()}, kernel
            gr.free()


def test_fixpoint_final_full_pass(pkg, gen, oracle):
    stats = _fixpoint_vs_oracle(pkg, gen, oracle, gen.dyck_er(700, 2000, seed=3), final_full_pass=True)
    assert stats["full_passes"] == 1


@pytest.mark.parametrize("sync", [1, 3, 8])
def test_fixpoint_sync_cadence(pkg, gen, oracle, sync):
    _fixpoint_vs_oracle(pkg, gen, oracle, gen.dyck_er(400, 1200, seed=8), iters_per_sync=sync)


def test_fixpoint_windows_ahead(pkg, gen, oracle):
    """Windows of steps enqueued one ahead of the host's wait: every window
    length gives the same result and the same step count as one step per
    window, with and without per-kernel events and under graph replay (no
    window ahead), the steps ahead of the converging window are not counted,
    and the delta buffers are left empty (a second call on the converged
    matrices is one quiet step)."""
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    q = gen.dyck_er(600, 1800, seed=31)
    g_ref, mats_ref, _ = oracle.solve(q, "seminaive")
    want = oracle.fingerprints(g_ref, mats_ref, q.n)
    steps = set()
    for opts in (dict(iters_per_sync=1), dict(iters_per_sync=2), dict(iters_per_sync=5),
                 dict(iters_per_sync=2, no_kernel_timing=True), dict(iters_per_sync=3, use_graph=True)):
        with pkg.Engine(0, **opts) as eng:
            gr = cf.Grammar(q.grammar, text=True)
            mats = gr.load(eng, cf.Graph.from_query(q))
            st = eng.fixpoint(gr.rule_triples(), mats)
            assert gr.results() == want, opts
            steps.add(st["iterations"])
            again = eng.fixpoint(gr.rule_triples(), mats)
            assert again["iterations"] == 1 and again["new_bits"] == 0, opts
            assert gr.results() == want, opts
            gr.free()
    assert len(steps) == 1, steps


def test_fixpoint_iteration_cap_with_windows_ahead(pkg, gen):
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    with pkg.Engine(0, max_iterations=3, iters_per_sync=2) as eng:
        gr = cf.Grammar(gen.DYCK_GRAMMAR, text=True)
        with pytest.raises(pkg.CfpqError) as ei:
            gr.intersection_with_graph(cf.Graph.from_query(gen.config(0)), eng)
        assert ei.value.code == pkg.E_LIMIT
        _ = eng.matrix(10).popcount()


def test_fixpoint_epsilon_and_self_rules(pkg, gen, oracle):
    q = gen.dyck_er(300, 800, seed=21)
    q.grammar = "S\nS A S1\nS1 S B\nS S S\nA a\nB b\n"
    _fixpoint_vs_oracle(pkg, gen, oracle, q)


def test_fixpoint_iteration_cap(pkg, gen, oracle):
    eng = pkg.Engine(0, max_iterations=2, iters_per_sync=1)
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    try:
        gr = cf.Grammar(gen.DYCK_GRAMMAR, text=True)
        with pytest.raises(pkg.CfpqError) as ei:
            gr.intersection_with_graph(cf.Graph.from_query(gen.config(0)), eng)
        assert ei.value.code == pkg.E_LIMIT
        # the context stays usable after the error
        eng.set_opts()
        _ = eng.matrix(10).popcount()
    finally:
        eng.close()


def test_worklist_mode_matches(pkg, gen, oracle, engine):
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    q = gen.dyck_er(500, 1500, seed=12)
    g_ref, mats_ref, _ = oracle.solve(q, "m4rm")
    gr = cf.Grammar(q.grammar, text=True)
    gr.intersection_with_graph(cf.Graph.from_query(q), engine, mode="worklist")
    assert gr.results() == oracle.fingerprints(g_ref, mats_ref, q.n)
    gr.free()


def test_batch_matches_single(pkg, gen, oracle, engine):
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    qs = gen.batch_queries(12, seed=99, n_lo=256, n_hi=1500)
    grs = []
    for q in qs:
        gr = cf.Grammar(q.grammar, text=True)
        gr.load(engine, cf.Graph.from_query(q))
        grs.append(gr)
    stats = engine.batch_fixpoint([(gr.rule_triples(), gr.matrices) for gr in grs])
    assert stats["iterations"] >= 2
    for q, gr in zip(qs, grs):
        g_ref, mats_ref, _ = oracle.solve(q, "seminaive")
        assert gr.results() == oracle.fingerprints(g_ref, mats_ref, q.n), q.name
        gr.free()


def test_invalid_arguments(pkg, engine):
    a = engine.matrix(10)
    b = engine.matrix(11)
    with pytest.raises(pkg.CfpqError) as ei:
        engine.add_mul(a, a, b)
    assert ei.value.code == pkg.E_INVALID
    with pytest.raises(pkg.CfpqError):
        engine.fixpoint([(0, 0, 5)], [a])
    with pytest.raises(pkg.CfpqError):
        engine.fixpoint([(0, 0, 1)], [a, b])
    a.free()
    b.free()


def test_batch_rejects_shared_head(pkg, gen, oracle, engine):
    """A matrix one query writes (a rule head) cannot appear in another query
    of the batch: that query's delta buffers would never see its new bits.
    Read-only operands may be shared, and such a batch gives each query's own
    fixpoint."""
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    n = 300
    S1, S2, A, B = (engine.matrix(n) for _ in range(4))
    rules = [(0, 1, 2)]  # S -> A B
    with pytest.raises(pkg.CfpqError) as ei:
        engine.batch_fixpoint([(rules, [S1, A, B]), (rules, [S2, S1, B])])  # S1 written by 0, read by 1
    assert ei.value.code == pkg.E_INVALID
    with pytest.raises(pkg.CfpqError) as ei:
        engine.batch_fixpoint([(rules, [S1, A, B]), (rules, [S1, A, B])])   # S1 written by both
    assert ei.value.code == pkg.E_INVALID
    for m in (S1, S2, A, B):
        m.free()
    # shared read-only terminal matrices: both queries close as they would alone
    q = gen.dyck_er(400, 1200, seed=17)
    g_ref, mats_ref, _ = oracle.solve(q, "seminaive")
    want = oracle.fingerprints(g_ref, mats_ref, q.n)
    g1, g2 = cf.Grammar(q.grammar, text=True), cf.Grammar(q.grammar, text=True)
    m1 = g1.load(engine, cf.Graph.from_query(q))
    m2 = g2.load(engine, cf.Graph.from_query(q))
    heads = {h for h, _, _ in g1.rule_triples()}
    shared = [m1[x] if x not in heads else m2[x] for x in range(len(m2))]
    engine.batch_fixpoint([(g1.rule_triples(), m1), (g2.rule_triples(), shared)])
    assert g1.results() == want
    for x in heads:
        assert shared[x].fingerprint() == m1[x].fingerprint()
    g1.free()
    g2.free()


def test_fixpoint_sg_forest_65536_golden(pkg, gen):
    """Full-size ontology workload: fingerprints and counts of every nonterminal
    against the oracle's golden fixture (size-independent check)."""
    import json
    import os
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixpoints.json")))["config4"]
    q = gen.config(4)
    assert q.digest() == golden["digest"]
    with pkg.Engine(0, final_full_pass=True) as eng:
        gr = cf.Grammar(q.grammar, text=True)
        _, _, stats = gr.intersection_with_graph(cf.Graph.from_query(q), eng)
        assert gr.results() == {k: tuple(v) for k, v in golden["fingerprints"].items()}
        assert stats["full_passes"] == 1
        gr.free()


@pytest.mark.parametrize("variant", ["tuned", "m4r", "tuned_certified"])
def test_fixpoint_dense_stress_32768_golden(pkg, gen, variant):
    """Full-size dense-tile stress workload (24 nonterminals, 64 binary rules,
    N = 32,768, power-law out-degree 8): every nonterminal's fingerprint and
    count against the oracle's golden fixture (its semi-naive M4RM run).

    * tuned: the cached options of this workload class (rule sweeps, Four
      Russians from 8,192 bits per A tile, the LDS path below);
    * m4r: every tile-product forced onto the Four Russians path;
    * tuned_certified: the tuned options plus the certification sweep, which
      must add no bit."""
    import json
    import os
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    tune = __import__("importlib").import_module("cfpq-on-gpgpu_amd.tune")
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixpoints.json")))["config2"]
    q = gen.config(2)
    assert q.digest() == golden["digest"]
    opts = tune.tuned_opts(q.n, q.grammar)
    assert opts, "no tuned options cached for the dense-stress workload class"
    if variant == "m4r":
        opts["kernel"] = "m4r"
    if variant == "tuned_certified":
        opts["final_full_pass"] = True
    print(f"[dense32768] {variant}: {opts}", flush=True)
    with pkg.Engine(0, **opts) as eng:
        gr = cf.Grammar(q.grammar, text=True)
        _, _, stats = gr.intersection_with_graph(cf.Graph.from_query(q), eng)
        assert gr.results() == {k: tuple(v) for k, v in golden["fingerprints"].items()}, variant
        if variant == "m4r":
            assert stats["tile_products_dense"] == stats["tile_products"] > 0
            assert stats["m4r_lookups"] > 0
        if variant == "tuned_certified":
            assert stats["full_passes"] == 1
        gr.free()


@pytest.mark.parametrize("opts", [
    # sparse kernel only (no heavy-tile bitsets, no dense / MFMA launches), 8 waves per CU
    dict(dense_bits=0xFFFFFFF0, mfma_bits=0xFFFFFFF0, sparse_waves_per_cu=8, sparse_chunk=4),
    # one claim per atomic, 16 waves per CU: waves drain and steal across all cursor shards
    dict(dense_bits=0xFFFFFFF0, mfma_bits=0xFFFFFFF0, sparse_waves_per_cu=16, sparse_chunk=1),
    # 1 wave per CU: most shards are drained by stealing waves
    dict(sparse_waves_per_cu=1, sparse_chunk=16, use_graph=True),
    # tile units: one unit per tile for its quarters with candidates (heavy quarters still split)
    dict(dense_bits=0xFFFFFFF0, mfma_bits=0xFFFFFFF0, sparse_mode=2, sparse_waves_per_cu=8, sparse_chunk=4),
    # heavy quarters shared out by rows (no merge pass), and by candidates from 16 on
    dict(dense_bits=0xFFFFFFF0, mfma_bits=0xFFFFFFF0, split_rows=True, split_cands=16),
    dict(dense_bits=0xFFFFFFF0, mfma_bits=0xFFFFFFF0, split_rows=False, split_cands=16),
])
def test_fixpoint_sg_forest_65536_options(pkg, gen, opts):
    """Full-size ontology workload under other launch shapes: fingerprints
    against the golden fixture, and the step counters (summed over the sparse
    kernel's per-XCD shards) account for exactly the bits the fixpoint added."""
    import json
    import os
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixpoints.json")))["config4"]
    q = gen.config(4)
    with pkg.Engine(0, **opts) as eng:
        gr = cf.Grammar(q.grammar, text=True)
        mats = gr.load(eng, cf.Graph.from_query(q))
        before = sum(c for _, c in gr.results().values())
        stats = eng.fixpoint(gr.rule_triples(), mats)
        got = gr.results()
        assert got == {k: tuple(v) for k, v in golden["fingerprints"].items()}
        assert stats["new_bits"] == sum(c for _, c in got.values()) - before
        assert stats["cand_hits"] <= stats["tile_products"]
        gr.free()


@pytest.mark.parametrize("which", ["dyck", "dense"])
def test_fixpoint_graph_replay(pkg, gen, oracle, which):
    """Steps captured as hipGraphs and replayed: several fixpoints in a row on
    one context (graphs reused), with eager work in between, each checked."""
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    q = gen.dyck_er(900, 3000, seed=4) if which == "dyck" else gen.dense_stress(700)
    g_ref, mats_ref, _ = oracle.solve(q, "seminaive")
    want = oracle.fingerprints(g_ref, mats_ref, q.n)
    with pkg.Engine(0, use_graph=True, iters_per_sync=3) as eng:
        gr = cf.Grammar(q.grammar, text=True)
        mats = gr.load(eng, cf.Graph.from_query(q))
        init = []
        for m in mats:
            c = eng.matrix(q.n)
            c.copy_from(m)
            init.append(c)
        for rep in range(3):
            for m, c in zip(mats, init):
                m.copy_from(c)
            st = eng.fixpoint(gr.rule_triples(), mats)
            assert gr.results() == want, rep
            # eager work on the same stream between replays
            x = eng.matrix(q.n)
            eng.add_mul(x, mats[0], mats[0])
            x.free()
        assert st["iterations"] >= 2
        for c in init:
            c.free()
        gr.free()


@pytest.mark.parametrize("which", ["dense", "dyck", "sg"])
def test_fixpoint_fused_apply(pkg, gen, oracle, which):
    """Head matrices updated inside the product kernels: new rows become
    visible to later units of the same step; the least fixpoint is unchanged."""
    q = {"dense": lambda: gen.dense_stress(1500), "dyck": lambda: gen.dyck_er(1200, 4000, seed=9),
         "sg": lambda: gen.sg_forest(8000, seed=10, types_per_vertex=4)}[which]()
    for kernel in ("auto", "gather", "m4r", "mfma"):
        _fixpoint_vs_oracle(pkg, gen, oracle, q, kernel=kernel, fuse_apply=True, dense_bits=64, m4r_bits=4096,
                            mfma_bits=30000)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_512_queries_golden(pkg, gen, graph):
    """The 512-query batch (N log-uniform in [256, 8192], Dyck or SG family)
    in one batched fixpoint, with its steps launched eagerly or replayed from
    captured hipGraphs (twice, so that the second run replays the graphs the
    first captured): every query's fingerprints against the oracle fixture."""
    import json
    import os
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixpoints.json")))["config3"]
    qs = gen.config(3)
    assert [q.digest() for q in qs] == [g["digest"] for g in golden]
    with pkg.Engine(0, use_graph=graph) as eng:
        grs, init = [], []
        for q in qs:
            gr = cf.Grammar(q.grammar, text=True)
            mats = gr.load(eng, cf.Graph.from_query(q))
            grs.append(gr)
            if graph:
                cp = []
                for m in mats:
                    c = eng.matrix(q.n)
                    c.copy_from(m)
                    cp.append(c)
                init.append(cp)
        for rep in range(2 if graph else 1):
            if rep:
                for gr, cp in zip(grs, init):
                    for m, c in zip(gr.matrices, cp):
                        m.copy_from(c)
            st = eng.batch_fixpoint([(gr.rule_triples(), gr.matrices) for gr in grs])
            assert st["iterations"] >= 2
            for q, gr, g in zip(qs, grs, golden):
                assert gr.results() == {k: tuple(v) for k, v in g["fingerprints"].items()}, (q.name, rep)
        for gr in grs:
            gr.free()
        for cp in init:
            for c in cp:
                c.free()


@pytest.mark.parametrize("schedule", ["sweep", "rules"])
@pytest.mark.parametrize("which", ["dense", "dyck", "sg", "self"])
def test_fixpoint_head_sweeps(pkg, gen, oracle, which, schedule):
    """CFPQ_SCHEDULE_SWEEP / _RULES: heads (or single rules) evaluated one at a
    time over the full matrices, later ones seeing earlier ones' bits, sweeps
    until one adds no bit.  Same least fixpoint with every kernel, with and
    without the head update fused into the products."""
    q = {"dense": lambda: gen.dense_stress(1500), "dyck": lambda: gen.dyck_er(1200, 4000, seed=9),
         "sg": lambda: gen.sg_forest(8000, seed=10, types_per_vertex=4),
         "self": lambda: gen.dyck_er(300, 800, seed=21)}[which]()
    if which == "self":  # epsilon rule and a rule whose operands alias its head
        q.grammar = "S\nS A S1\nS1 S B\nS S S\nA a\nB b\n"
    for kernel, fuse in (("auto", False), ("auto", True), ("gather", False), ("m4r", True), ("mfma", True)):
        st = _fixpoint_vs_oracle(pkg, gen, oracle, q, kernel=kernel, fuse_apply=fuse, schedule=schedule, dense_bits=64,
                                 m4r_bits=4096, mfma_bits=30000)
        assert st["iterations"] >= 2 and st["full_passes"] == 0


@pytest.mark.parametrize("schedule", ["sweep", "rules"])
def test_fixpoint_head_sweeps_final_full_pass_and_cap(pkg, gen, oracle, schedule):
    q = gen.dyck_er(700, 2000, seed=3)
    st = _fixpoint_vs_oracle(pkg, gen, oracle, q, schedule=schedule, final_full_pass=True)
    assert st["full_passes"] == 1
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    with pkg.Engine(0, schedule=schedule, max_iterations=3) as eng:
        gr = cf.Grammar(q.grammar, text=True)
        with pytest.raises(pkg.CfpqError) as ei:
            gr.intersection_with_graph(cf.Graph.from_query(q), eng)
        assert ei.value.code == pkg.E_LIMIT
        gr.free()


def test_fixpoint_head_sweeps_dense_4096_golden(pkg, gen):
    import json
    import os
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixpoints.json")))["config2_n4096"]
    q = gen.dense_stress(4096)
    assert q.digest() == golden["digest"]
    for schedule in ("sweep", "rules"):
        with pkg.Engine(0, schedule=schedule, fuse_apply=True) as eng:
            gr = cf.Grammar(q.grammar, text=True)
            gr.intersection_with_graph(cf.Graph.from_query(q), eng)
            assert gr.results() == {k: tuple(v) for k, v in golden["fingerprints"].items()}, schedule
            gr.free()


@pytest.mark.parametrize("schedule", ["sweep", "rules"])
def test_batch_head_sweeps(pkg, gen, oracle, engine, schedule):
    """A batch under head or rule sweeps: every query's fixpoint as computed alone."""
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    qs = gen.batch_queries(6, seed=77, n_lo=200, n_hi=900)
    want = []
    for q in qs:
        g_ref, mats_ref, _ = oracle.solve(q, "seminaive")
        want.append(oracle.fingerprints(g_ref, mats_ref, q.n))
    with pkg.Engine(0, schedule=schedule) as eng:
        grs = []
        for q in qs:
            gr = cf.Grammar(q.grammar, text=True)
            gr.load(eng, cf.Graph.from_query(q))
            grs.append(gr)
        eng.batch_fixpoint([(gr.rule_triples(), gr.matrices) for gr in grs])
        for gr, w in zip(grs, want):
            assert gr.results() == w
            gr.free()


@pytest.mark.parametrize("which", ["dense", "dyck", "sg", "self"])
def test_fixpoint_tile_units(pkg, gen, oracle, which):
    """sparse_mode 2: a work unit covers every quarter of its tile with
    candidates (rows 64 q + r of one accumulator); same least fixpoint with
    the other kernels, fused or separate apply, and with delta or rule-sweep
    schedules."""
    q = {"dense": lambda: gen.dense_stress(1500), "dyck": lambda: gen.dyck_er(1200, 4000, seed=9),
         "sg": lambda: gen.sg_forest(8000, seed=10, types_per_vertex=4),
         "self": lambda: gen.dyck_er(300, 800, seed=21)}[which]()
    if which == "self":
        q.grammar = "S\nS A S1\nS1 S B\nS S S\nA a\nB b\n"
    for kernel, fuse, schedule, mode in (("auto", False, "delta", 2), ("auto", True, "delta", 2),
                                         ("gather", False, "rules", 2), ("auto", False, "delta", 3),
                                         ("auto", True, "rules", 3)):
        st = _fixpoint_vs_oracle(pkg, gen, oracle, q, kernel=kernel, fuse_apply=fuse, schedule=schedule,
                                 sparse_mode=mode, dense_bits=0xFFFFFFF0 if kernel == "auto" else 0,
                                 mfma_bits=0xFFFFFFF0)
        assert st["iterations"] >= 2


@pytest.mark.parametrize("which", ["dense", "dyck", "sg"])
@pytest.mark.parametrize("split_rows", [False, True])
def test_fixpoint_split_quarters(pkg, gen, oracle, which, split_rows):
    """Quarters with more than split_cands candidates are shared out over
    sub-units, by candidates (partial rows, merge pass) or by rows (each
    sub-unit writes its own rows and ORs the quarter's metadata): with
    thresholds of 1 and 2 nearly every quarter is split (and with pieces of
    one candidate, quarters of more than 4 split 16 ways), and the fixpoint
    must still equal the oracle's word for word, for quarter and tile units.
    Thresholds of 8 and 12 mix the three kinds of entry in one step: unsplit
    quarters of more than 3/4 of the threshold in the heavy list, split ones
    (of at least 8 pieces: 16 sub-units) and ordinary ones."""
    q = {"dense": lambda: gen.dense_stress(1100), "dyck": lambda: gen.dyck_er(1200, 4000, seed=9),
         "sg": lambda: gen.sg_forest(6000, seed=10, types_per_vertex=4)}[which]()
    for cands, piece, mode, schedule in ((1, 0, 0, "delta"), (2, 0, 2, "delta"), (1, 0, 0, "rules"),
                                         (4, 1, 0, "delta"), (8, 1, 0, "delta"), (12, 4, 0, "delta")):
        st = _fixpoint_vs_oracle(pkg, gen, oracle, q, kernel="auto", split_rows=split_rows, split_cands=cands,
                                 split_piece=piece, sparse_mode=mode, schedule=schedule, dense_bits=0xFFFFFFF0,
                                 mfma_bits=0xFFFFFFF0)
        assert st["iterations"] >= 2


def _chain_query(gen, labels, grammar, label_names):
    """A path 0 -> 1 -> ... -> len(labels) with the given edge labels."""
    import numpy as np
    n = len(labels) + 1
    src = np.arange(n - 1, dtype=np.uint32)
    dst = np.arange(1, n, dtype=np.uint32)
    lab = np.asarray(labels, dtype=np.uint8)
    return gen.Query("chain", n, grammar, src, dst, lab, label_names)


@pytest.mark.parametrize("schedule", ["delta", "rules"])
def test_large_chain_closed_forms(pkg, gen, schedule):
    """Sizes the oracle cannot run, checked against closed forms:
    * transitive closure of a 32,768-vertex path under S -> S S | a is the
      strict upper triangle (every pair i < j), 15 steps of doubling;
    * the Dyck grammar on the chain a^k b^k (k = 8,192) derives exactly the k
      nested pairs (k - m, k + m), m = 1..k."""
    import numpy as np
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    with pkg.Engine(0, schedule=schedule) as eng:
        n = 32768
        q = _chain_query(gen, [0] * (n - 1), "S S S\nS a\n", ("a",))
        gr = cf.Grammar(q.grammar, text=True)
        gr.intersection_with_graph(cf.Graph.from_query(q), eng)
        S = gr.matrices[gr.nonterminal_to_index["S"]]
        assert S.popcount() == n * (n - 1) // 2
        rows = S.download()
        # row i holds columns i+1 .. n-1: zero words, one partial word, then all ones
        first = np.arange(1, n + 1, dtype=np.int64)  # first set column of row i
        w = np.arange(rows.shape[1], dtype=np.int64)[None, :]
        fw = (first // 64)[:, None]
        partial = (~((np.uint64(1) << (first % 64).astype(np.uint64)) - np.uint64(1)))[:, None]
        want = np.where(w < fw, np.uint64(0), np.where(w == fw, partial, np.uint64(0xFFFFFFFFFFFFFFFF)))
        assert np.array_equal(rows, want.astype(np.uint64))
        gr.free()
        k = 8192
        q = _chain_query(gen, [0] * k + [1] * k, gen.DYCK_GRAMMAR, gen.DYCK_LABELS)
        gr = cf.Grammar(q.grammar, text=True)
        gr.intersection_with_graph(cf.Graph.from_query(q), eng)
        S = gr.matrices[gr.nonterminal_to_index["S"]].download()
        idx = np.nonzero(np.unpackbits(S.view(np.uint8), axis=1, bitorder="little")[:, :2 * k + 1])
        assert sorted(zip(idx[0].tolist(), idx[1].tolist())) == [(k - m, k + m) for m in range(k, 0, -1)]
        gr.free()


def test_empty_inputs(pkg, gen, engine):
    """No edges, no binary rules, a single vertex: the fixpoint is the initial
    matrices (epsilon diagonals included)."""
    import numpy as np
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    empty = gen.Query("empty", 300, gen.DYCK_GRAMMAR, np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                      np.zeros(0, np.uint8), gen.DYCK_LABELS)
    gr = cf.Grammar(empty.grammar, text=True)
    _, _, st = gr.intersection_with_graph(cf.Graph.from_query(empty), engine)
    assert all(c == 0 for _, c in gr.results().values())
    gr.free()
    q = gen.dyck_er(200, 600, seed=31)
    for grammar, want_counts in (("A a\nB b\n", None), ("S\nA a\n", None)):
        gr = cf.Grammar(grammar, text=True)
        graph = cf.Graph.from_query(q)
        gr.intersection_with_graph(graph, engine)
        res = gr.results()
        if "S" in res:
            assert res["S"][1] == q.n  # epsilon: the diagonal only
        if "A" in res:
            assert res["A"][1] == int((np.asarray(q.labels)[q.label] == "a").sum())
        gr.free()
    one = gen.Query("one", 1, "S S S\nS a\n", np.zeros(1, np.uint32), np.zeros(1, np.uint32), np.zeros(1, np.uint8),
                    ("a",))
    gr = cf.Grammar(one.grammar, text=True)
    gr.intersection_with_graph(cf.Graph.from_query(one), engine)
    assert gr.results()["S"][1] == 1
    gr.free()


def test_left_operand_after_fixpoint_that_did_not_keep_it(pkg, gen, oracle, engine):
    """A rule head that no rule of its fixpoint reads as a left operand (S1 of
    the Dyck grammar) has its column-major form left out of date by that
    fixpoint.  Used afterwards as the left operand of add_mul, a copy or
    another fixpoint, it must give the oracle's products."""
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    q = gen.dyck_er(600, 2000, seed=61)
    g_ref, mats_ref, _ = oracle.solve(q, "seminaive")
    gr = cf.Grammar(q.grammar, text=True)
    gr.intersection_with_graph(cf.Graph.from_query(q), engine)
    assert gr.results() == oracle.fingerprints(g_ref, mats_ref, q.n)
    s1 = gr.matrices[gr.nonterminal_to_index["S1"]]
    b = gr.matrices[gr.nonterminal_to_index["B"]]
    S1, B = mats_ref[g_ref.nt_index["S1"]], mats_ref[g_ref.nt_index["B"]]
    for left in ("itself", "copy"):
        src = s1
        if left == "copy":
            src = engine.matrix(q.n)
            src.copy_from(s1)
        x = engine.matrix(q.n)
        engine.add_mul(x, src, b)
        want = oracle.zeros(q.n)
        oracle.addmul_m4rm(want, S1.copy(), B, q.n)
        assert np.array_equal(x.download(), want), left
        y = engine.matrix(q.n)
        engine.fixpoint([(0, 1, 2)], [y, src, b])   # Y -> S1 B: S1 read as a left operand
        assert np.array_equal(y.download(), want), left
        for m in (x, y):
            m.free()
        if left == "copy":
            src.free()
    gr.free()


@pytest.mark.parametrize("n", [1, 63, 256, 257, 600])
def test_add_mul_m4r_8bit_keys(pkg, oracle, n):
    """The Four Russians path with 8-bit keys and per-wavefront tables over
    64-column slices, forced for every tile-product: bit-exact against the
    oracle's M4RM over densities up to all ones, and with aliasing."""
    rng = np.random.default_rng(900 + n)
    with pkg.Engine(0, kernel="m4r", m4r_keys=8) as eng:
        for density in (0.003, 0.05, 0.3, 0.7, 1.0):
            a, b, c = rand_mat(rng, n, density), rand_mat(rng, n, density), rand_mat(rng, n, 0.01)
            A, B, C = oracle.from_dense(a), oracle.from_dense(b), oracle.from_dense(c)
            want = C.copy()
            want_changed = oracle.addmul_m4rm(want, A, B, n)
            dA, dB, dC = upload(eng, oracle, a), upload(eng, oracle, b), upload(eng, oracle, c)
            assert eng.add_mul(dC, dA, dB) == want_changed
            assert np.array_equal(dC.download(), want), (n, density)
            assert dC.fingerprint() == oracle.fingerprint(want, n)
            for m in (dA, dB, dC):
                m.free()
        x = rand_mat(rng, n, 0.2)
        X = oracle.from_dense(x)
        want = X.copy()
        oracle.addmul_m4rm(want, X.copy(), X.copy(), n)
        dX = upload(eng, oracle, x)
        eng.add_mul(dX, dX, dX)
        assert np.array_equal(dX.download(), want)
        dX.free()


@pytest.mark.parametrize("schedule", ["delta", "rules"])
def test_fixpoint_m4r_8bit_keys(pkg, gen, oracle, schedule):
    stats = _fixpoint_vs_oracle(pkg, gen, oracle, gen.dense_stress(1500), kernel="m4r", m4r_keys=8, schedule=schedule)
    assert stats["m4r_lookups"] > 0 and stats["tile_products_dense"] == stats["tile_products"]


def test_fixpoint_dense_stress_4096_golden_m4r_8bit(pkg, gen):
    import json
    import os
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixpoints.json")))["config2_n4096"]
    q = gen.dense_stress(4096)
    with pkg.Engine(0, schedule="rules", m4r_bits=4096, dense_bits=64, m4r_keys=8) as eng:
        gr = cf.Grammar(q.grammar, text=True)
        _, _, st = gr.intersection_with_graph(cf.Graph.from_query(q), eng)
        assert gr.results() == {k: tuple(v) for k, v in golden["fingerprints"].items()}
        assert st["m4r_lookups"] > 0
        gr.free()


@pytest.mark.parametrize("which", ["sg", "dyck", "self"])
def test_restored_matrices_behave_as_new(pkg, gen, oracle, which):
    """Matrices that held a fixpoint and were restored from copies of their
    inputs must behave like freshly loaded ones: the reference worklist
    (FIFO, one add_mul per popped rule) over them matches the oracle's
    add_mul call by call (changed flag and head matrix), and repeated fixpoints
    with the head update fused into the products give the oracle's result
    every time.  (Column words of a head cleared and filled again must never
    carry rows of its earlier contents.)"""
    cf = __import__("importlib").import_module("cfpq-on-gpgpu_amd.cfpq")
    q = {"sg": lambda: gen.sg_forest(3000, seed=71, types_per_vertex=3), "dyck": lambda: gen.dyck_er(500, 1600, seed=72),
         "self": lambda: gen.dyck_er(300, 800, seed=73)}[which]()
    if which == "self":
        q.grammar = "S\nS A S1\nS1 S B\nS S S\nA a\nB b\n"
    g = oracle.ParsedGrammar(q.grammar)
    g_ref, mats_ref, _ = oracle.solve(q, "seminaive")
    want = oracle.fingerprints(g_ref, mats_ref, q.n)
    for fuse in (False, True):
        with pkg.Engine(0, fuse_apply=fuse) as eng:
            gr = cf.Grammar(q.grammar, text=True)
            mats = gr.load(eng, cf.Graph.from_query(q))
            init = []
            for m in mats:
                c = eng.matrix(q.n)
                c.copy_from(m)
                init.append(c)
            for rep in range(3):
                eng.fixpoint(gr.rule_triples(), mats)
                assert gr.results() == want, (fuse, rep)
                for m, c in zip(mats, init):
                    m.copy_from(c)
            # the reference worklist over the restored matrices, call by call
            om = oracle.initial_matrices(g, q.n, q.src, q.dst, q.label, q.labels)
            idx = {gr.nonterminal_to_index[nm]: g.nt_index[nm] for nm in g.nt_index}
            with_nt = [[] for _ in range(gr.nonterminals_count)]
            todo = []
            for i, (h, (a, b)) in enumerate(gr.rules):
                with_nt[a].append(i)
                with_nt[b].append(i)
                todo.append(i)
            pending = set(todo)
            while todo:
                i = todo.pop(0)
                pending.discard(i)
                h, (a, b) = gr.rules[i]
                ch = eng.add_mul(mats[h], mats[a], mats[b])
                och = oracle.addmul_m4rm(om[idx[h]], om[idx[a]].copy(), om[idx[b]].copy(), q.n)
                assert bool(ch) == bool(och), (fuse, i)
                assert np.array_equal(mats[h].download(), om[idx[h]]), (fuse, i)
                if ch:
                    for j in with_nt[h]:
                        if j not in pending:
                            pending.add(j)
                            todo.append(j)
            assert gr.results() == want
            for c in init:
                c.free()
            gr.free()
