"""CPU oracle for the CFPQ fixpoint -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this package, and only as the checker.  The shipped engine never
loads it.

Wraps ``liboracle.so`` (oracle.cpp, a CPU restatement of the reference's
worklist fixpoint cfpq-cpp/Grammar.h:58-69 with the M4RM product of
m4ri/m4ri_matrix.cpp:20-28 and the fingerprint of cfpq-cpp/Grammar.cpp:81-101)
and adds a small, independent grammar parser that follows
cfpq-cpp/Grammar.cpp:17-51.

Matrices are numpy ``uint64`` arrays of shape (n, ceil(n/64)), row-major,
column j in word j//64 bit j%64.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "cfpq_ref")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_lib = None


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_addmul_m4rm.restype = ctypes.c_int
        L.oracle_addmul_m4rm.argtypes = [_u64p, _u64p, _u64p, ctypes.c_uint32]
        L.oracle_addmul_rows.restype = ctypes.c_int
        L.oracle_addmul_rows.argtypes = [_u64p, _u64p, _u64p, ctypes.c_uint32]
        L.oracle_fixpoint_worklist.restype = ctypes.c_uint64
        L.oracle_fixpoint_worklist.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                               ctypes.c_uint32, ctypes.POINTER(_u64p), ctypes.c_int]
        L.oracle_fixpoint_worklist_budget.restype = ctypes.c_double
        L.oracle_fixpoint_worklist_budget.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                                      ctypes.c_uint32, ctypes.POINTER(_u64p), ctypes.c_int,
                                                      ctypes.c_double, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_fixpoint_seminaive.restype = ctypes.c_uint32
        L.oracle_fixpoint_seminaive.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                                ctypes.c_uint32, ctypes.POINTER(_u64p)]
        L.oracle_fixpoint_seminaive_m4rm.restype = ctypes.c_uint32
        L.oracle_fixpoint_seminaive_m4rm.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                                     ctypes.c_uint32, ctypes.POINTER(_u64p),
                                                     ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_hellings.restype = ctypes.c_uint64
        L.oracle_hellings.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.c_uint32, ctypes.POINTER(_u64p)]
        for f in ("oracle_fingerprint", "oracle_fingerprint_literal"):
            getattr(L, f).restype = None
            getattr(L, f).argtypes = [_u64p, ctypes.c_uint32, _u64p, _u64p]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_set_num_threads.restype = None
        L.oracle_set_num_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(_u64p)


def wpr(n: int) -> int:
    return (n + 63) // 64


def zeros(n: int) -> np.ndarray:
    return np.zeros((n, wpr(n)), dtype=np.uint64)


def from_dense(d: np.ndarray) -> np.ndarray:
    """Pack a boolean (n, n) array."""
    n = d.shape[0]
    m = zeros(n)
    r, c = np.nonzero(d)
    np.bitwise_or.at(m, (r, c // 64), np.left_shift(np.uint64(1), (c % 64).astype(np.uint64)))
    return m


def to_dense(m: np.ndarray, n: int) -> np.ndarray:
    bits = np.unpackbits(m.view(np.uint8), axis=1, bitorder="little")
    return bits[:, :n].astype(bool)


def set_bits(m: np.ndarray, rows, cols) -> None:
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    np.bitwise_or.at(m, (rows, cols // 64), np.left_shift(np.uint64(1), (cols % 64).astype(np.uint64)))


# ---------------------------------------------------------------------------
# grammar (cfpq-cpp/Grammar.cpp:17-51)
# ---------------------------------------------------------------------------

class ParsedGrammar:
    def __init__(self, text: str):
        self.nt_index: dict[str, int] = {}
        self.eps: list[int] = []
        self.term: dict[str, list[int]] = {}
        self.rules: list[tuple[int, int, int]] = []
        for line in text.splitlines():
            t = line.split()
            if not t:
                continue
            h = self._idx(t[0])
            if len(t) == 1:
                self.eps.append(h)
            elif len(t) == 2:
                self.term.setdefault(t[1], []).append(h)
            elif len(t) == 3:
                b = self._idx(t[1])
                c = self._idx(t[2])
                self.rules.append((h, b, c))

    def _idx(self, s):
        if s not in self.nt_index:
            self.nt_index[s] = len(self.nt_index)
        return self.nt_index[s]

    @property
    def nnt(self) -> int:
        return len(self.nt_index)

    def rules_array(self) -> np.ndarray:
        return np.asarray(self.rules, dtype=np.uint32).reshape(-1, 3)


def initial_matrices(g: ParsedGrammar, n: int, src, dst, label, labels) -> list[np.ndarray]:
    """Terminal and epsilon bits (cfpq-cpp/Grammar.h:41-51)."""
    mats = [zeros(n) for _ in range(g.nnt)]
    for li, name in enumerate(labels):
        sel = np.asarray(label) == li
        for h in g.term.get(name, []):
            set_bits(mats[h], np.asarray(src)[sel], np.asarray(dst)[sel])
    for h in g.eps:
        set_bits(mats[h], np.arange(n), np.arange(n))
    return mats


def _ptrs(mats):
    arr = (_u64p * len(mats))()
    for i, m in enumerate(mats):
        arr[i] = _p(m)
    return arr


def _rules_ptr(rules: np.ndarray):
    rules = np.ascontiguousarray(rules, dtype=np.uint32)
    return rules, rules.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


# ---------------------------------------------------------------------------
# products and fixpoints
# ---------------------------------------------------------------------------

def addmul_m4rm(C, A, B, n) -> bool:
    return bool(lib().oracle_addmul_m4rm(_p(C), _p(A), _p(B), n))


def addmul_rows(C, A, B, n) -> bool:
    return bool(lib().oracle_addmul_rows(_p(C), _p(A), _p(B), n))


def fixpoint_worklist(n, rules, mats, method: str = "m4rm") -> int:
    """Reference worklist fixpoint in place; returns the add_mul call count."""
    r, rp = _rules_ptr(rules)
    return int(lib().oracle_fixpoint_worklist(n, len(mats), rp, r.shape[0], _ptrs(mats),
                                              0 if method == "m4rm" else 1))


def fixpoint_worklist_budget(n, rules, mats, max_seconds: float, method: str = "m4rm"):
    """Reference worklist for at most ~max_seconds; returns (calls, seconds)."""
    r, rp = _rules_ptr(rules)
    calls = ctypes.c_uint64()
    el = lib().oracle_fixpoint_worklist_budget(n, len(mats), rp, r.shape[0], _ptrs(mats),
                                               0 if method == "m4rm" else 1, max_seconds, ctypes.byref(calls))
    return int(calls.value), float(el)


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def set_num_threads(k: int) -> None:
    lib().oracle_set_num_threads(int(k))


def fixpoint_seminaive(n, rules, mats) -> int:
    r, rp = _rules_ptr(rules)
    return int(lib().oracle_fixpoint_seminaive(n, len(mats), rp, r.shape[0], _ptrs(mats)))


def fixpoint_seminaive_m4rm(n, rules, mats):
    """Semi-naive fixpoint with dense M4RM products; returns (steps, products)."""
    r, rp = _rules_ptr(rules)
    prods = ctypes.c_uint64()
    steps = int(lib().oracle_fixpoint_seminaive_m4rm(n, len(mats), rp, r.shape[0], _ptrs(mats), ctypes.byref(prods)))
    return steps, int(prods.value)


def hellings(n, rules, mats) -> int:
    r, rp = _rules_ptr(rules)
    return int(lib().oracle_hellings(n, len(mats), rp, r.shape[0], _ptrs(mats)))


def fingerprint(m: np.ndarray, n: int, literal: bool = False) -> tuple[int, int]:
    h = ctypes.c_uint64()
    c = ctypes.c_uint64()
    f = lib().oracle_fingerprint_literal if literal else lib().oracle_fingerprint
    f(_p(np.ascontiguousarray(m)), n, ctypes.byref(h), ctypes.byref(c))
    return int(h.value), int(c.value)


def solve(query, method: str = "seminaive"):
    """Least fixpoint of a ``gen.Query``.  Returns (grammar, mats, info)."""
    g = ParsedGrammar(query.grammar)
    mats = initial_matrices(g, query.n, query.src, query.dst, query.label, query.labels)
    t0 = time.perf_counter()
    if method == "seminaive":
        info = {"iterations": fixpoint_seminaive(query.n, g.rules_array(), mats)}
    elif method in ("m4rm", "rows"):
        info = {"add_mul_calls": fixpoint_worklist(query.n, g.rules_array(), mats, method)}
    elif method == "seminaive_m4rm":
        steps, prods = fixpoint_seminaive_m4rm(query.n, g.rules_array(), mats)
        info = {"iterations": steps, "products": prods}
    elif method == "hellings":
        info = {"facts": hellings(query.n, g.rules_array(), mats)}
    else:
        raise ValueError(method)
    info["seconds"] = time.perf_counter() - t0
    return g, mats, info


def fingerprints(g: ParsedGrammar, mats, n) -> dict[str, tuple[int, int]]:
    """``name -> (hash, count)`` as Grammar::print_results writes them."""
    return {name: fingerprint(mats[idx], n) for name, idx in g.nt_index.items()}


def run_reference_driver(grammar_path, graph_path, out_path):
    """Run oracle/_ref/cfpq_ref (reference cfpq-cpp driver + oracle matrix)."""
    p = subprocess.run([REF_BIN, grammar_path, graph_path, out_path], capture_output=True, text=True, check=True)
    total_ms, algo_ms = (int(x) for x in p.stdout.split())
    calls = int(p.stderr.split()[-1])
    res = {}
    with open(out_path) as f:
        for line in f:
            name, h, c = line.split()
            res[name] = (int(h), int(c))
    return res, {"total_ms": total_ms, "algorithm_ms": algo_ms, "add_mul_calls": calls}
