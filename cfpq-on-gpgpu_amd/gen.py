"""Seeded synthetic CFPQ workloads (graphs + CNF grammars).

Every workload is produced from numpy's PCG64 generator seeded with
``1000 + workload index`` (batched queries use ``SeedSequence.spawn``), so the
CPU oracle, the GPU engine and the benchmark all read identical inputs.  A
query is returned as plain text in the reference's own file formats:

* grammar: one rule per line, ``H B1 B2`` (binary), ``NT term`` (terminal) or
  ``NT`` (epsilon) -- the format parsed by cfpq-cpp/Grammar.cpp:17-51;
* graph: one edge per line, ``from label to`` -- cfpq-cpp/Graph.cpp:11-20,
  with ``vertices_count = max id + 1``.

Because the reference derives N from the largest vertex id, every generator
here also returns N explicitly and tests check they agree.
"""
from __future__ import annotations

import dataclasses
import hashlib
import math

import numpy as np

# --------------------------------------------------------------------------
# Grammars (CNF text)
# --------------------------------------------------------------------------

#: Dyck-like ``S -> a S b | a b`` in CNF.
DYCK_GRAMMAR = """S A S1
S A B
S1 S B
A a
B b
"""

#: Same-generation query over RDF-like ``subClassOf``/``type`` edges:
#: ``S -> sc^-1 S sc | t^-1 S t | sc^-1 sc | t^-1 t`` in CNF (7 nonterminals,
#: 6 binary rules).  ``_r`` marks the inverse edge label.
SG_GRAMMAR = """S SCI S1
S TI S2
S SCI SC
S TI T
S1 S SC
S2 S T
SCI subClassOf_r
SC subClassOf
TI type_r
T type
"""

DYCK_LABELS = ("a", "b")
SG_LABELS = ("subClassOf", "subClassOf_r", "type", "type_r")


@dataclasses.dataclass
class Query:
    """One (graph, grammar) CFPQ instance."""

    name: str
    n: int                 # vertex count (explicit, >= max id + 1)
    grammar: str           # CNF grammar text
    src: np.ndarray        # uint32 [E]
    dst: np.ndarray        # uint32 [E]
    label: np.ndarray      # uint8 [E], index into labels
    labels: tuple

    @property
    def n_edges(self) -> int:
        return int(self.src.shape[0])

    def graph_text(self) -> str:
        names = np.asarray(self.labels, dtype=object)[self.label]
        return "".join(f"{s} {l} {d}\n" for s, l, d in zip(self.src.tolist(), names, self.dst.tolist()))

    def digest(self) -> str:
        h = hashlib.sha256()
        h.update(self.grammar.encode())
        h.update(np.uint64(self.n).tobytes())
        h.update(np.ascontiguousarray(self.src, dtype=np.uint32).tobytes())
        h.update(np.ascontiguousarray(self.dst, dtype=np.uint32).tobytes())
        h.update(np.ascontiguousarray(self.label, dtype=np.uint8).tobytes())
        h.update("|".join(self.labels).encode())
        return h.hexdigest()[:16]


def random_cnf_grammar(rng: np.random.Generator, n_nt: int, n_rules: int, n_term: int) -> tuple[str, tuple]:
    """Random CNF grammar: ``n_rules`` distinct binary rules over ``n_nt``
    nonterminals; every terminal ``t<k>`` derives from at least one
    nonterminal (one or two, chosen at random).  Nonterminal 0 is ``S``."""
    names = ["S"] + [f"N{i}" for i in range(1, n_nt)]
    seen = set()
    rules = []
    while len(rules) < n_rules:
        h, l, r = (int(x) for x in rng.integers(0, n_nt, size=3))
        if (h, l, r) in seen:
            continue
        seen.add((h, l, r))
        rules.append((h, l, r))
    lines = [f"{names[h]} {names[l]} {names[r]}" for h, l, r in rules]
    labels = tuple(f"t{k}" for k in range(n_term))
    for k in range(n_term):
        cnt = 1 + int(rng.integers(0, 2))
        nts = rng.choice(n_nt, size=cnt, replace=False)
        for nt in sorted(int(x) for x in nts):
            lines.append(f"{names[nt]} {labels[k]}")
    return "\n".join(lines) + "\n", labels


# --------------------------------------------------------------------------
# Graphs
# --------------------------------------------------------------------------

def er_edges(rng: np.random.Generator, n: int, m: int) -> tuple[np.ndarray, np.ndarray]:
    """``m`` distinct ordered pairs (u != v) drawn uniformly (``m`` <= n(n-1))."""
    m = min(m, n * (n - 1))
    keys = np.empty(0, dtype=np.int64)
    while keys.shape[0] < m:
        need = m - keys.shape[0]
        u = rng.integers(0, n, size=2 * need + 16)
        v = rng.integers(0, n, size=2 * need + 16)
        ok = u != v
        k = (u[ok].astype(np.int64) * n + v[ok])
        keys = np.concatenate([keys, k])
        _, first = np.unique(keys, return_index=True)
        keys = keys[np.sort(first)]
    keys = keys[:m]
    return (keys // n).astype(np.uint32), (keys % n).astype(np.uint32)


def zipf_forest(rng: np.random.Generator, n: int, s: float = 1.5, p_root: float = 0.001) -> np.ndarray:
    """Parent array of a random forest: vertex v >= 1 becomes a new root with
    probability ``p_root``, else picks a parent r < v with weight (r+1)^-s.
    Roots (and vertex 0) have parent -1."""
    w = np.arange(1, n + 1, dtype=np.float64) ** (-s)
    cum = np.cumsum(w)
    v = np.arange(n)
    u = rng.random(n)
    parent = np.full(n, -1, dtype=np.int64)
    tot = np.where(v > 0, cum[np.maximum(v - 1, 0)], 0.0)
    cand = np.searchsorted(cum, u * tot, side="right")
    cand = np.minimum(cand, np.maximum(v - 1, 0))
    root = rng.random(n) < p_root
    sel = (v > 0) & ~root
    parent[sel] = cand[sel]
    return parent


def forest_type_edges(rng: np.random.Generator, n: int, type_frac: float = 0.3,
                      types_per_vertex: int = 1) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Ontology-shaped graph: ``v -subClassOf-> parent(v)`` for a Zipf forest,
    ``v -type-> c`` from a ``type_frac`` share of vertices to random class
    (internal) nodes, ``types_per_vertex`` distinct classes each, plus the
    inverse ``_r`` edge of every edge.  Labels index :data:`SG_LABELS`."""
    parent = zipf_forest(rng, n)
    child = np.nonzero(parent >= 0)[0]
    sc_src = child.astype(np.int64)
    sc_dst = parent[child]
    classes = np.unique(sc_dst)
    typed = np.nonzero(rng.random(n) < type_frac)[0]
    t_src = np.repeat(typed, types_per_vertex)
    t_dst = classes[rng.integers(0, classes.shape[0], size=t_src.shape[0])]
    keys = np.unique(t_src * n + t_dst)
    t_src, t_dst = keys // n, keys % n
    src = np.concatenate([sc_src, sc_dst, t_src, t_dst])
    dst = np.concatenate([sc_dst, sc_src, t_dst, t_src])
    lab = np.concatenate([np.full(sc_src.shape, 0), np.full(sc_src.shape, 1),
                          np.full(t_src.shape, 2), np.full(t_src.shape, 3)])
    return src.astype(np.uint32), dst.astype(np.uint32), lab.astype(np.uint8)


def powerlaw_edges(rng: np.random.Generator, n: int, mean_deg: float, alpha: float = 2.5) -> tuple[np.ndarray, np.ndarray]:
    """Directed graph whose out-degrees follow a discrete power law
    P(d) ~ d^-alpha, rescaled to ``mean_deg``; targets uniform, no self loops,
    duplicate pairs removed."""
    raw = (rng.pareto(alpha - 1.0, size=n) + 1.0)
    deg = np.floor(raw * (mean_deg / raw.mean())).astype(np.int64)
    deg = np.clip(deg, 0, n - 1)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    dst = rng.integers(0, n - 1, size=src.shape[0])
    dst = dst + (dst >= src)             # skip self loop
    keys = np.unique(src * n + dst)
    return (keys // n).astype(np.uint32), (keys % n).astype(np.uint32)


# --------------------------------------------------------------------------
# The five benchmark workloads
# --------------------------------------------------------------------------

def _rng(idx: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(1000 + idx))


def dyck_er(n: int = 1024, m: int = 4096, seed: int = 1000) -> Query:
    """Dyck-like grammar on an Erdos-Renyi graph, labels uniform over {a,b}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    src, dst = er_edges(rng, n, m)
    lab = rng.integers(0, 2, size=src.shape[0]).astype(np.uint8)
    return Query(f"dyck_er_n{n}", n, DYCK_GRAMMAR, src, dst, lab, DYCK_LABELS)


def sg_forest(n: int = 16384, seed: int = 1001, type_frac: float = 0.3,
              types_per_vertex: int = 1) -> Query:
    rng = np.random.Generator(np.random.PCG64(seed))
    src, dst, lab = forest_type_edges(rng, n, type_frac, types_per_vertex)
    return Query(f"sg_forest_n{n}_t{types_per_vertex}", n, SG_GRAMMAR, src, dst, lab, SG_LABELS)


def dense_stress(n: int = 32768, seed: int = 1002, n_nt: int = 24, n_rules: int = 64,
                 n_term: int = 4, mean_deg: float = 8.0) -> Query:
    rng = np.random.Generator(np.random.PCG64(seed))
    grammar, labels = random_cnf_grammar(rng, n_nt, n_rules, n_term)
    src, dst = powerlaw_edges(rng, n, mean_deg)
    lab = rng.integers(0, n_term, size=src.shape[0]).astype(np.uint8)
    return Query(f"dense_stress_n{n}", n, grammar, src, dst, lab, labels)


def batch_queries(count: int = 512, seed: int = 1003, n_lo: int = 256, n_hi: int = 8192,
                  mean_deg: float = 4.0) -> list[Query]:
    """Independent queries: N log-uniform in [n_lo, n_hi], ER with mean
    out-degree ``mean_deg``, Dyck or SG family with probability 1/2 each."""
    subs = np.random.SeedSequence(seed).spawn(count)
    out = []
    for qi, ss in enumerate(subs):
        rng = np.random.Generator(np.random.PCG64(ss))
        n = int(math.floor(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        fam = int(rng.integers(0, 2))
        src, dst = er_edges(rng, n, int(mean_deg * n))
        if fam == 0:
            lab = rng.integers(0, 2, size=src.shape[0]).astype(np.uint8)
            out.append(Query(f"batch{qi}_dyck_n{n}", n, DYCK_GRAMMAR, src, dst, lab, DYCK_LABELS))
        else:
            lab = rng.integers(0, 4, size=src.shape[0]).astype(np.uint8)
            out.append(Query(f"batch{qi}_sg_n{n}", n, SG_GRAMMAR, src, dst, lab, SG_LABELS))
    return out


def sg_forest_dense_types(n: int = 65536, seed: int = 1004, density: float = 1e-4,
                          type_frac: float = 0.3) -> Query:
    """The ontology generator with the number of type edges per typed vertex
    raised until the edge count (inverses included) is about density * N^2."""
    target = density * n * n
    sc_pairs = n - 1
    typed = type_frac * n
    k = max(1, int(round((target / 2 - sc_pairs) / typed)))
    return sg_forest(n, seed, type_frac, k)


def config(idx: int, scale: float = 1.0):
    """Workload ``idx`` (0..4).  ``scale`` < 1 shrinks N (same generators) for
    tests that must finish on the CPU oracle."""
    def sz(n):
        return max(256, int(n * scale))
    if idx == 0:
        return dyck_er(sz(1024), max(16, int(4096 * scale)))
    if idx == 1:
        return sg_forest(sz(16384))
    if idx == 2:
        return dense_stress(sz(32768))
    if idx == 3:
        return batch_queries(max(1, int(512 * scale)))
    if idx == 4:
        return sg_forest_dense_types(sz(65536))
    raise ValueError(idx)
