"""Pins the CPU oracle (CPU only).

The reference ships no golden vectors for this path (every data file is a
Git-LFS pointer), so the oracle is pinned by
  * the fingerprint's closed-form known answers (cfpq-cpp/Grammar.cpp:81-101),
  * numpy's integer matmul (A @ B > 0) -- same semantics as the int8 MFMA path,
  * closed forms of the Boolean product (identity, zero, all-ones, permutations),
  * agreement of four independent fixpoint algorithms (worklist + M4RM,
    worklist + row gather, semi-naive, Hellings' fact worklist),
  * textbook answers (Dyck pairs on a chain, transitive closure),
  * the reference's own driver (oracle/_ref/cfpq_ref: cfpq-cpp/Grammar.cpp,
    Graph.cpp compiled unmodified with the oracle matrix plugged in).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest


def rand_mat(rng, n, density):
    return rng.random((n, n)) < density


def numpy_product(a, b):
    return (a.astype(np.int64) @ b.astype(np.int64)) > 0


# ---------------------------------------------------------------- fingerprint

def test_fingerprint_known_answers(oracle):
    assert oracle.fingerprint(oracle.from_dense(np.ones((1, 1), bool)), 1) == (1, 1)
    assert oracle.fingerprint(oracle.from_dense(np.eye(2, dtype=bool)), 2) == (144, 2)
    assert oracle.fingerprint(oracle.from_dense(np.ones((3, 3), bool)), 3) == (24339, 9)
    assert oracle.fingerprint(oracle.zeros(5), 5) == (0, 0)


def test_fingerprint_matches_python_statement(oracle):
    # an independent pure-Python statement of the loop in Grammar.cpp:84-97
    rng = np.random.default_rng(7)
    for n in (1, 2, 3, 17, 64, 65, 130):
        d = rand_mat(rng, n, 0.3)
        h = 0
        for r in range(n):
            line = 0
            for c in range(n):
                line = (line * 11 + int(d[r, c])) % 1000000007
            h = (h * 13 + line) % 1000000007
        m = oracle.from_dense(d)
        assert oracle.fingerprint(m, n) == (h, int(d.sum()))
        assert oracle.fingerprint(m, n, literal=True) == (h, int(d.sum()))


# ---------------------------------------------------------------- products

@pytest.mark.parametrize("n", [1, 7, 63, 64, 65, 200, 256, 300])
@pytest.mark.parametrize("density", [0.0, 0.02, 0.3, 0.9])
def test_products_match_numpy(oracle, n, density):
    rng = np.random.default_rng(n * 1000 + int(density * 100))
    a, b, c = rand_mat(rng, n, density), rand_mat(rng, n, density), rand_mat(rng, n, 0.05)
    want = c | numpy_product(a, b)
    for fn in (oracle.addmul_m4rm, oracle.addmul_rows):
        C = oracle.from_dense(c)
        changed = fn(C, oracle.from_dense(a), oracle.from_dense(b), n)
        assert np.array_equal(oracle.to_dense(C, n), want)
        assert changed == bool((want & ~c).any())


def test_product_closed_forms(oracle):
    n = 150
    rng = np.random.default_rng(3)
    b = rand_mat(rng, n, 0.2)
    eye, zero, ones = np.eye(n, dtype=bool), np.zeros((n, n), bool), np.ones((n, n), bool)
    perm = np.eye(n, dtype=bool)[rng.permutation(n)]
    for fn in (oracle.addmul_m4rm, oracle.addmul_rows):
        C = oracle.zeros(n)
        fn(C, oracle.from_dense(eye), oracle.from_dense(b), n)           # I*B = B
        assert np.array_equal(oracle.to_dense(C, n), b)
        C = oracle.zeros(n)
        assert not fn(C, oracle.from_dense(b), oracle.from_dense(zero), n)  # A*0 = 0
        C = oracle.zeros(n)
        fn(C, oracle.from_dense(ones), oracle.from_dense(ones), n)       # J*J = J
        assert oracle.to_dense(C, n).all()
        # P^3 by repeated products equals the permutation composed three times
        P = oracle.from_dense(perm)
        acc = P.copy()
        for _ in range(2):
            nxt = oracle.zeros(n)
            fn(nxt, acc, P, n)
            acc = nxt
        assert np.array_equal(oracle.to_dense(acc, n), perm @ perm @ perm)


def test_product_aliasing(oracle):
    # add_mul(C, C, C): the product is over the operands as they were on entry
    n = 100
    rng = np.random.default_rng(11)
    c = rand_mat(rng, n, 0.05)
    want = c | numpy_product(c, c)
    for fn in (oracle.addmul_m4rm, oracle.addmul_rows):
        C = oracle.from_dense(c)
        fn(C, C, C, n)
        assert np.array_equal(oracle.to_dense(C, n), want)


# ---------------------------------------------------------------- fixpoints

def _solve_all(oracle, q):
    out = {}
    for method in ("m4rm", "rows", "seminaive", "seminaive_m4rm", "hellings"):
        g, mats, info = oracle.solve(q, method)
        out[method] = (oracle.fingerprints(g, mats, q.n), info)
    return out


@pytest.mark.parametrize("which", ["dyck", "sg_er", "sg_forest", "random_cnf", "epsilon"])
def test_fixpoint_methods_agree(oracle, gen, which):
    if which == "dyck":
        q = gen.dyck_er(300, 900, seed=5)
    elif which == "sg_er":
        q = gen.batch_queries(4, seed=77, n_lo=200, n_hi=400)[1]
    elif which == "sg_forest":
        q = gen.sg_forest(2000, seed=9, types_per_vertex=3)
    elif which == "random_cnf":
        q = gen.dense_stress(300, seed=13, n_nt=8, n_rules=12, mean_deg=2.0)
    else:  # grammar with an epsilon rule: S -> eps | a S b  (CNF: S -> A S1 | eps ; S1 -> S B)
        q = gen.dyck_er(200, 500, seed=21)
        q.grammar = "S\nS A S1\nS1 S B\nA a\nB b\n"
    res = _solve_all(oracle, q)
    ref = res["seminaive"][0]
    for method, (fp, _) in res.items():
        assert fp == ref, method


def test_dyck_chain_closed_form(oracle, gen):
    # chain 0 -a-> 1 -a-> ... -a-> k -b-> ... -b-> 2k : S(i, j) iff the path
    # spells a^m b^m, i.e. i = k - m, j = k + m for m = 1..k
    k = 40
    n = 2 * k + 1
    src = np.arange(n - 1, dtype=np.uint32)
    dst = src + 1
    lab = (src >= k).astype(np.uint8)
    q = gen.Query("chain", n, gen.DYCK_GRAMMAR, src, dst, lab, gen.DYCK_LABELS)
    g, mats, _ = oracle.solve(q, "m4rm")
    s = oracle.to_dense(mats[g.nt_index["S"]], n)
    want = np.zeros((n, n), bool)
    for m in range(1, k + 1):
        want[k - m, k + m] = True
    assert np.array_equal(s, want)


def test_transitive_closure_closed_form(oracle, gen):
    # S -> S S | a on a random graph = transitive closure of the a-edges
    rng = np.random.default_rng(4)
    n = 120
    src, dst = gen.er_edges(rng, n, 150)
    q = gen.Query("tc", n, "S S S\nS a\n", src, dst, np.zeros(src.shape, np.uint8), ("a",))
    g, mats, _ = oracle.solve(q, "rows")
    got = oracle.to_dense(mats[0], n)
    reach = np.zeros((n, n), bool)
    reach[src, dst] = True
    for k in range(n):  # Warshall
        reach |= reach[:, [k]] & reach[[k], :]
    assert np.array_equal(got, reach)
    # complete graph: every pair
    full = np.array([(i, j) for i in range(30) for j in range(30)], dtype=np.uint32)
    q = gen.Query("full", 30, "S S S\nS a\n", full[:, 0], full[:, 1], np.zeros(900, np.uint8), ("a",))
    g, mats, _ = oracle.solve(q, "m4rm")
    assert oracle.to_dense(mats[0], 30).all()


def test_fixpoint_idempotent(oracle, gen):
    q = gen.dyck_er(256, 700, seed=2)
    g, mats, _ = oracle.solve(q, "seminaive")
    before = [m.copy() for m in mats]
    for h, a, b in g.rules:
        assert not oracle.addmul_rows(mats[h], mats[a], mats[b], q.n)
    assert all(np.array_equal(x, y) for x, y in zip(before, mats))


# ---------------------------------------------------------------- reference driver

@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "cfpq_ref")),
                    reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("which", ["dyck", "sg", "random_cnf"])
def test_reference_driver_agrees(oracle, gen, which):
    q = {"dyck": lambda: gen.dyck_er(512, 2048, seed=31),
         "sg": lambda: gen.sg_forest(4096, seed=32, types_per_vertex=4),
         "random_cnf": lambda: gen.dense_stress(256, seed=33, n_nt=10, n_rules=20, mean_deg=3.0)}[which]()
    with tempfile.TemporaryDirectory() as d:
        gp, ep, op = (os.path.join(d, x) for x in ("g.txt", "e.txt", "o.txt"))
        open(gp, "w").write(q.grammar)
        open(ep, "w").write(q.graph_text())
        ref, info = oracle.run_reference_driver(gp, ep, op)
    n_file = int(max(q.src.max(), q.dst.max())) + 1  # the reference derives N from the edges
    q.n = n_file
    g, mats, winfo = oracle.solve(q, "m4rm")
    assert oracle.fingerprints(g, mats, q.n) == ref
    # same worklist order as the reference => same number of add_mul calls
    assert winfo["add_mul_calls"] == info["add_mul_calls"]


def _cyk_path(g, labels_on_path):
    """CYK over every substring of the word spelled by a simple path: on a path
    graph, CFPQ result NT(i, j) holds iff NT derives the word between i and j."""
    n = len(labels_on_path) + 1
    table = {}  # (i, j) -> set of nonterminals deriving word[i:j]
    for i, lab in enumerate(labels_on_path):
        table[(i, i + 1)] = set(g.term.get(lab, []))
    for length in range(2, n):
        for i in range(0, n - length):
            j = i + length
            cell = set()
            for k in range(i + 1, j):
                left, right = table[(i, k)], table[(k, j)]
                if not left or not right:
                    continue
                for h, a, b in g.rules:
                    if a in left and b in right:
                        cell.add(h)
            table[(i, j)] = cell
    return table


@pytest.mark.parametrize("which", ["dyck", "random_cnf"])
def test_path_graph_equals_cyk(oracle, gen, which):
    # CFPQ on a chain 0 -> 1 -> ... -> L is CYK over all substrings of its word
    rng = np.random.default_rng(17 if which == "dyck" else 18)
    L = 40
    if which == "dyck":
        grammar, labels = gen.DYCK_GRAMMAR, gen.DYCK_LABELS
    else:
        q0 = gen.dense_stress(64, seed=19, n_nt=6, n_rules=14, mean_deg=2.0)
        grammar, labels = q0.grammar, q0.labels
    lab = rng.integers(0, len(labels), size=L).astype(np.uint8)
    src = np.arange(L, dtype=np.uint32)
    q = gen.Query("path", L + 1, grammar, src, src + 1, lab, labels)
    g, mats, _ = oracle.solve(q, "m4rm")
    table = _cyk_path(g, [labels[x] for x in lab])
    assert sum(len(c) for (i, j), c in table.items() if j - i >= 2) > 0  # non-trivial derivations exist
    for name, idx in g.nt_index.items():
        got = oracle.to_dense(mats[idx], L + 1)
        want = np.zeros((L + 1, L + 1), bool)
        for (i, j), cell in table.items():
            if idx in cell:
                want[i, j] = True
        assert np.array_equal(got, want), name


def test_worklist_budget_matches_worklist(oracle, gen):
    """The bounded worklist that times the CPU baseline is the reference
    worklist itself: with a budget it cannot reach, it makes the same add_mul
    calls in the same order and ends at the same matrices; with a zero budget
    it stops after exactly one call, which equals one M4RM add_mul of the
    first popped rule (checked by a numpy product)."""
    import numpy as np
    for q in (gen.dyck_er(300, 900, seed=5), gen.dense_stress(400, seed=9)):
        g = oracle.ParsedGrammar(q.grammar)
        m_full = oracle.initial_matrices(g, q.n, q.src, q.dst, q.label, q.labels)
        m_budget = [m.copy() for m in m_full]
        m_one = [m.copy() for m in m_full]
        calls = oracle.fixpoint_worklist(q.n, g.rules_array(), m_full)
        calls_b, secs = oracle.fixpoint_worklist_budget(q.n, g.rules_array(), m_budget, 1e9)
        assert calls_b == calls and secs >= 0.0
        for a, b in zip(m_full, m_budget):
            assert np.array_equal(a, b)
        calls_1, _ = oracle.fixpoint_worklist_budget(q.n, g.rules_array(), m_one, 0.0)
        assert calls_1 == 1
        init = oracle.initial_matrices(g, q.n, q.src, q.dst, q.label, q.labels)
        changed = [x for x in range(len(init)) if not np.array_equal(init[x], m_one[x])]
        assert len(changed) <= 1
        if changed:  # the one call's head now holds head | left * right of the initial matrices
            x = changed[0]
            cands = [(l, r) for h, l, r in g.rules_array().tolist() if h == x]
            dense = [oracle.to_dense(m, q.n).astype(np.uint8) for m in init]
            assert any(np.array_equal(oracle.to_dense(m_one[x], q.n),
                                      dense[x].astype(bool) | ((dense[l] @ dense[r]) > 0)) for l, r in cands)
