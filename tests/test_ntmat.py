"""The single-matrix formulation (one n x n matrix of nonterminal sets,
cuda/cutlass/CutlassMatrix.cu): host helpers on the CPU, the device closure
cfpq_nt_closure against the oracle's per-nonterminal fixpoint on the GPU.
Bit x of element (i, j) must equal bit (i, j) of the oracle's matrix of
nonterminal x, for every x: the least fixpoint is the same whatever the
layout."""
import importlib

import numpy as np
import pytest

cf = importlib.import_module("cfpq-on-gpgpu_amd.cfpq")


def _planes(elems, nnt):
    return [((elems >> np.uint32(x)) & np.uint32(1)).astype(bool) for x in range(nnt)]


def test_fingerprint_bits_kats_and_oracle(oracle):
    assert cf.fingerprint_bits(np.ones((1, 1), bool)) == (1, 1)
    assert cf.fingerprint_bits(np.eye(2, dtype=bool)) == (144, 2)
    assert cf.fingerprint_bits(np.ones((3, 3), bool)) == (24339, 9)
    rng = np.random.default_rng(3)
    for n in (1, 5, 64, 130, 700):
        d = rng.random((n, n)) < 0.1
        assert cf.fingerprint_bits(d) == oracle.fingerprint(oracle.from_dense(d), n), n


@pytest.mark.parametrize("which", ["dyck", "sg", "eps"])
def test_nt_elements_match_initial_matrices(oracle, gen, which):
    q = {"dyck": lambda: gen.dyck_er(150, 500, seed=4), "sg": lambda: gen.sg_forest(300, seed=5, types_per_vertex=2),
         "eps": lambda: gen.dyck_er(90, 300, seed=6)}[which]()
    if which == "eps":
        q.grammar = "S\nS A S1\nS1 S B\nS S S\nA a\nB b\n"
    gr = cf.Grammar(q.grammar, text=True)
    el = gr.nt_elements(cf.Graph.from_query(q))
    g = oracle.ParsedGrammar(q.grammar)
    mats = oracle.initial_matrices(g, q.n, q.src, q.dst, q.label, q.labels)
    assert g.nt_index == gr.nonterminal_to_index
    for x, plane in enumerate(_planes(el, gr.nonterminals_count)):
        assert np.array_equal(plane, oracle.to_dense(mats[x], q.n)), x


def test_nt_elements_rejects_more_than_32_nonterminals():
    text = "\n".join(f"N{i} N{i + 1} N{i + 2}" for i in range(40))
    gr = cf.Grammar(text, text=True)
    g = cf.Graph(src=[0], dst=[1], labels=["a"], n=2)
    with pytest.raises(ValueError):
        gr.nt_elements(g)


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["dyck", "sg", "dense", "eps", "tiny"])
def test_nt_closure_matches_oracle(pkg, oracle, gen, which):
    q = {"dyck": lambda: gen.dyck_er(700, 2000, seed=8), "sg": lambda: gen.sg_forest(1000, seed=9, types_per_vertex=3),
         "dense": lambda: gen.dense_stress(257, seed=10), "eps": lambda: gen.dyck_er(100, 300, seed=11),
         "tiny": lambda: gen.dyck_er(63, 120, seed=12)}[which]()
    if which == "eps":
        q.grammar = "S\nS A S1\nS1 S B\nS S S\nA a\nB b\n"
    g, mats, _ = oracle.solve(q, "seminaive")
    gr = cf.Grammar(q.grammar, text=True)
    graph = cf.Graph.from_query(q)
    el = gr.nt_elements(graph)
    with pkg.Engine(0) as eng:
        st = eng.nt_closure(gr.rule_triples(), el)
    assert st["iterations"] >= 1
    for x, plane in enumerate(_planes(el, gr.nonterminals_count)):
        assert np.array_equal(plane, oracle.to_dense(mats[x], q.n)), (which, x)
    assert gr.nt_results(el) == oracle.fingerprints(g, mats, q.n)


@pytest.mark.gpu
def test_nt_closure_edge_cases(pkg):
    with pkg.Engine(0) as eng:
        # n = 1 with a self loop: S -> S S over a one-vertex graph
        el = np.array([[1]], dtype=np.uint32)
        eng.nt_closure([(0, 0, 0)], el)
        assert el[0, 0] == 1
        # no rules: nothing changes
        el = np.eye(5, dtype=np.uint32) * 3
        eng.nt_closure([], el)
        assert np.array_equal(el, np.eye(5, dtype=np.uint32) * 3)
        # transitive closure of a path under S -> S S
        n = 130
        el = np.zeros((n, n), dtype=np.uint32)
        el[np.arange(n - 1), np.arange(1, n)] = 1
        st = eng.nt_closure([(0, 0, 0)], el)
        assert np.array_equal(el != 0, np.triu(np.ones((n, n), bool), 1))
        assert st["items"] == n * (n - 1) // 2 - (n - 1)
        with pytest.raises(pkg.CfpqError):
            eng.nt_closure([(32, 0, 0)], np.zeros((4, 4), dtype=np.uint32))
        with pytest.raises(pkg.CfpqError):
            eng.nt_closure([(0, 0, 0)] * 257, np.zeros((4, 4), dtype=np.uint32))
