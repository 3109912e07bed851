"""Python mirror of the reference's cfpq-cpp interface (Grammar / Graph),
driving the gfx950 engine through the C ABI.

Same names, file formats and semantics as the reference:
  Grammar(path)                       cfpq-cpp/Grammar.cpp:17-51
  Graph(path)                         cfpq-cpp/Graph.cpp:11-20
  Grammar.intersection_with_graph     cfpq-cpp/Grammar.h:23-81 -> (total_ms, algorithm_ms)
  Grammar.print_results(path)         cfpq-cpp/Grammar.cpp:81-101 ("name hash count", name order)
The C++ drop-in (include/cfpq/*.h) is the production host; this module is
what the tests and the benchmark use.
"""
from __future__ import annotations

import time

import numpy as np

from . import Engine


class Grammar:
    def __init__(self, source: str, text: bool = False):
        if not text:
            with open(source) as f:
                source = f.read()
        self.nonterminal_to_index: dict[str, int] = {}
        self.epsilon_nonterminals: list[int] = []
        self.terminal_to_nonterminals: dict[str, list[int]] = {}
        self.rules: list[tuple[int, tuple[int, int]]] = []
        for line in source.splitlines():
            terms = line.split()
            if not terms:
                continue
            head = self._index(terms[0])
            if len(terms) == 1:
                self.epsilon_nonterminals.append(head)
            elif len(terms) == 2:
                self.terminal_to_nonterminals.setdefault(terms[1], []).append(head)
            elif len(terms) == 3:
                b1 = self._index(terms[1])
                b2 = self._index(terms[2])
                self.rules.append((head, (b1, b2)))
        self.vertices_count = 0
        self.matrices = []
        self.pair_sets = None  # closed pair keys after mode="pairs"

    def _index(self, name: str) -> int:
        if name not in self.nonterminal_to_index:
            self.nonterminal_to_index[name] = len(self.nonterminal_to_index)
        return self.nonterminal_to_index[name]

    @property
    def nonterminals_count(self) -> int:
        return len(self.nonterminal_to_index)

    def rule_triples(self) -> list[tuple[int, int, int]]:
        return [(h, a, b) for h, (a, b) in self.rules]

    def load(self, engine: Engine, graph: "Graph"):
        """Allocate one device matrix per nonterminal and set terminal and
        epsilon bits (cfpq-cpp/Grammar.h:26-51)."""
        n = graph.vertices_count
        self.vertices_count = n
        self.matrices = [engine.matrix(n) for _ in range(self.nonterminals_count)]
        rows = [[] for _ in range(self.nonterminals_count)]
        cols = [[] for _ in range(self.nonterminals_count)]
        for lab, (src, dst) in graph.edges_by_label().items():
            for nt in self.terminal_to_nonterminals.get(lab, []):
                rows[nt].append(src)
                cols[nt].append(dst)
        for nt in self.epsilon_nonterminals:
            rows[nt].append(np.arange(n, dtype=np.uint32))
            cols[nt].append(np.arange(n, dtype=np.uint32))
        for nt, m in enumerate(self.matrices):
            if rows[nt]:
                m.set_bits(np.concatenate(rows[nt]), np.concatenate(cols[nt]))
        return self.matrices

    def intersection_with_graph(self, graph: "Graph", engine: Engine, mode: str = "fixpoint"):
        """Returns (total_ms, algorithm_ms, stats).  mode "fixpoint" runs the
        device-resident delta fixpoint; "worklist" runs the reference's
        worklist loop (Grammar.h:58-69) over cfpq_add_mul; "pairs" runs the
        sparse pair-list fixpoint (cfpq_pairs_fixpoint, no n x n matrices),
        whose closed sets results() then reads."""
        if mode == "pairs":
            self.free()
            self.vertices_count = graph.vertices_count
            t0 = time.perf_counter()
            self.pair_sets, stats = engine.pairs_fixpoint(self.rule_triples(), self.pair_keys(graph),
                                                          graph.vertices_count)
            total = (time.perf_counter() - t0) * 1e3
            return total, stats["fixpoint_ms"], stats
        self.pair_sets = None
        self.load(engine, graph)
        t0 = time.perf_counter()
        engine.sync()
        ta = time.perf_counter()
        if mode == "fixpoint":
            stats = engine.fixpoint(self.rule_triples(), self.matrices)
        elif mode == "worklist":
            stats = {"add_mul_calls": self._worklist(engine)}
        else:
            raise ValueError(mode)
        engine.sync()
        tb = time.perf_counter()
        return (tb - t0) * 1e3, (tb - ta) * 1e3, stats

    def _worklist(self, engine: Engine) -> int:
        with_nt = [[] for _ in range(self.nonterminals_count)]
        todo = []
        for i, (h, (a, b)) in enumerate(self.rules):
            with_nt[a].append(i)
            with_nt[b].append(i)
            todo.append(i)
        pending = set(todo)
        calls = 0
        while todo:
            i = todo.pop(0)
            pending.discard(i)
            h, (a, b) = self.rules[i]
            calls += 1
            if engine.add_mul(self.matrices[h], self.matrices[a], self.matrices[b]):
                for j in with_nt[h]:
                    if j not in pending:
                        pending.add(j)
                        todo.append(j)
        return calls

    def nt_elements(self, graph: "Graph") -> np.ndarray:
        """The single matrix of nonterminal sets for `graph` (the CUTLASS
        backend's input, cfpq-cpp/Graph.cpp:22-48): bit x of element (i, j) for
        every terminal edge i -a-> j with x -> a, and the diagonal of epsilon
        nonterminals (cfpq-cpp/Grammar.h:47-51).  At most 32 nonterminals."""
        if self.nonterminals_count > 32:
            raise ValueError(f"{self.nonterminals_count} nonterminals: the single-matrix form holds 32")
        n = graph.vertices_count
        el = np.zeros((n, n), dtype=np.uint32)
        for lab, (src, dst) in graph.edges_by_label().items():
            for nt in self.terminal_to_nonterminals.get(lab, []):
                np.bitwise_or.at(el, (src.astype(np.int64), dst.astype(np.int64)), np.uint32(1 << nt))
        for nt in self.epsilon_nonterminals:
            idx = np.arange(n)
            el[idx, idx] |= np.uint32(1 << nt)
        return el

    def nt_results(self, elems: np.ndarray) -> dict[str, tuple[int, int]]:
        """Fingerprints (name -> (hash, count)) of a closed single matrix, as
        print_results gives them for the per-nonterminal matrices."""
        out = {}
        for name, idx in sorted(self.nonterminal_to_index.items(), key=lambda kv: kv[0].encode()):
            out[name] = fingerprint_bits(((elems >> np.uint32(idx)) & np.uint32(1)).astype(bool))
        return out

    def pair_keys(self, graph: "Graph") -> list[np.ndarray]:
        """Per nonterminal, the initial pairs row << 32 | col for the pair-list
        formulation: terminal edges and the epsilon diagonal, as load() sets
        them (cfpq-cpp/Grammar.h:41-51); duplicates are left in."""
        n = graph.vertices_count
        keys = [[] for _ in range(self.nonterminals_count)]
        for lab, (src, dst) in graph.edges_by_label().items():
            for nt in self.terminal_to_nonterminals.get(lab, []):
                keys[nt].append(src.astype(np.uint64) << np.uint64(32) | dst.astype(np.uint64))
        for nt in self.epsilon_nonterminals:
            d = np.arange(n, dtype=np.uint64)
            keys[nt].append(d << np.uint64(32) | d)
        return [np.concatenate(k) if k else np.zeros(0, np.uint64) for k in keys]

    def pair_results(self, sets: list[np.ndarray], n: int) -> dict[str, tuple[int, int]]:
        """Fingerprints (name -> (hash, count)) of closed pair sets, as
        print_results gives them for the matrices."""
        return {name: fingerprint_pairs(sets[idx], n)
                for name, idx in sorted(self.nonterminal_to_index.items(), key=lambda kv: kv[0].encode())}

    def results(self) -> dict[str, tuple[int, int]]:
        """name -> (hash, count), in the reference's name order."""
        if self.pair_sets is not None:
            return self.pair_results(self.pair_sets, self.vertices_count)
        return {name: self.matrices[idx].fingerprint() for name, idx in sorted(self.nonterminal_to_index.items(),
                                                                                key=lambda kv: kv[0].encode())}

    def print_results(self, path: str):
        with open(path, "w") as f:
            for name, (h, c) in self.results().items():
                f.write(f"{name} {h} {c}\n")

    def free(self):
        for m in self.matrices:
            m.free()
        self.matrices = []


HASH_MOD = 1_000_000_007


def fingerprint_bits(bits: np.ndarray) -> tuple[int, int]:
    """(hash, count) of an n x n boolean matrix as cfpq-cpp/Grammar.cpp:81-101
    prints it: row hash = sum_j bit * 11^(n-1-j), matrix hash = sum_i row_i *
    13^(n-1-i), both mod 1e9+7 (the Horner loops written as sums)."""
    n = bits.shape[0]
    p11 = np.ones(n, dtype=np.int64)
    p13 = np.ones(n, dtype=np.int64)
    for j in range(n - 2, -1, -1):
        p11[j] = p11[j + 1] * 11 % HASH_MOD
        p13[j] = p13[j + 1] * 13 % HASH_MOD
    line = (bits.astype(np.int64) @ p11) % HASH_MOD  # < n * 2^30 before the reduction
    return int(((line * p13) % HASH_MOD).sum() % HASH_MOD), int(bits.sum())


def _pow_mod(base: int, e: np.ndarray) -> np.ndarray:
    """base ** e mod HASH_MOD elementwise (square and multiply; every product
    is below 2^60)."""
    out = np.ones(e.shape, dtype=np.int64)
    b = np.full(e.shape, base % HASH_MOD, dtype=np.int64)
    e = e.astype(np.int64).copy()
    while e.any():
        odd = (e & 1).astype(bool)
        out[odd] = out[odd] * b[odd] % HASH_MOD
        b = b * b % HASH_MOD
        e >>= 1
    return out


def fingerprint_pairs(keys: np.ndarray, n: int) -> tuple[int, int]:
    """fingerprint_bits of the matrix whose set bits are the distinct pairs
    row << 32 | col in keys, without building the n x n matrix: each bit
    (i, j) adds 11^(n-1-j) * 13^(n-1-i)."""
    k = np.unique(np.asarray(keys, dtype=np.uint64))
    if k.size == 0:
        return 0, 0
    i = (k >> np.uint64(32)).astype(np.int64)
    j = (k & np.uint64(0xFFFFFFFF)).astype(np.int64)
    h = _pow_mod(11, n - 1 - j) * _pow_mod(13, n - 1 - i) % HASH_MOD
    return int(h.sum() % HASH_MOD), int(k.size)


class Graph:
    def __init__(self, source=None, *, src=None, dst=None, labels=None, n=None):
        """Either a path to a ``from label to`` edge list, or arrays
        (src, dst, labels as a sequence of strings) plus an optional n."""
        if source is not None:
            s, l, d = [], [], []
            with open(source) as f:
                toks = f.read().split()
            for i in range(0, len(toks) - 2, 3):
                s.append(int(toks[i]))
                l.append(toks[i + 1])
                d.append(int(toks[i + 2]))
            src, dst, labels = s, d, l
        self.src = np.asarray(src, dtype=np.uint32)
        self.dst = np.asarray(dst, dtype=np.uint32)
        self.labels = list(labels)
        vc = int(max(self.src.max(initial=0), self.dst.max(initial=0)) + 1) if self.src.size else 0
        self.vertices_count = max(vc, n or 0)

    @classmethod
    def from_query(cls, q) -> "Graph":
        names = np.asarray(q.labels, dtype=object)[q.label]
        return cls(src=q.src, dst=q.dst, labels=names, n=q.n)

    def edges_by_label(self) -> dict[str, tuple[np.ndarray, np.ndarray]]:
        lab = np.asarray(self.labels, dtype=object)
        out = {}
        for name in dict.fromkeys(self.labels):
            sel = lab == name
            out[name] = (self.src[sel], self.dst[sel])
        return out
