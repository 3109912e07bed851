"""MI355X (gfx950) CFPQ engine -- Python binding of the C ABI in include/cfpq_hip.h.

The engine itself is ``libcfpq_hip.so`` (hand-written HIP kernels, built
in-tree by ``make -C cfpq-on-gpgpu_amd``).  This module only marshals
arguments through ``ctypes``; there is no CPU fallback: if the library is
missing, importing :func:`lib` raises.

The package directory name contains dashes, so import it with
``importlib.import_module("cfpq-on-gpgpu_amd")``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcfpq_hip.so")

KERNEL_AUTO, KERNEL_GATHER, KERNEL_LDS, KERNEL_M4R, KERNEL_MFMA = range(5)
KERNELS = {"auto": KERNEL_AUTO, "gather": KERNEL_GATHER, "lds": KERNEL_LDS, "m4r": KERNEL_M4R, "mfma": KERNEL_MFMA}
#: cfpq_opts.schedule (CFPQ_SCHEDULE_*)
SCHEDULE_DELTA, SCHEDULE_SWEEP, SCHEDULE_RULES = 0, 1, 2
SCHEDULES = {"delta": SCHEDULE_DELTA, "sweep": SCHEDULE_SWEEP, "rules": SCHEDULE_RULES}

E_INVALID, E_DEVICE, E_NOMEM, E_LIMIT = -1, -2, -3, -4


class CfpqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cfpq error {code}: {msg}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [
        ("kernel", ctypes.c_int32),
        ("dense_bits", ctypes.c_uint32),
        ("mfma_bits", ctypes.c_uint32),
        ("m4r_bits", ctypes.c_uint32),
        ("max_iterations", ctypes.c_uint32),
        ("final_full_pass", ctypes.c_int32),
        ("iters_per_sync", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("sparse_batch", ctypes.c_int32),
        ("sparse_waves_per_cu", ctypes.c_int32),
        ("sparse_chunk", ctypes.c_int32),
        ("sched", ctypes.c_int32),
        ("sched_threshold", ctypes.c_uint32),
        ("sparse_mode", ctypes.c_int32),
        ("fuse_apply", ctypes.c_int32),
        ("schedule", ctypes.c_int32),
        ("m4r_keys", ctypes.c_int32),
        ("sched_order", ctypes.c_int32),
        ("order_threshold", ctypes.c_uint32),
        ("split_cands", ctypes.c_uint32),
        ("split_rows", ctypes.c_int32),
        ("split_piece", ctypes.c_uint32),
        ("no_kernel_timing", ctypes.c_int32),
    ]


class Rule(ctypes.Structure):
    _fields_ = [("head", ctypes.c_uint32), ("left", ctypes.c_uint32), ("right", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint32),
        ("full_passes", ctypes.c_uint32),
        ("new_bits", ctypes.c_uint64),
        ("items", ctypes.c_uint64),
        ("tile_products", ctypes.c_uint64),
        ("tile_products_dense", ctypes.c_uint64),
        ("tile_products_mfma", ctypes.c_uint64),
        ("a_rows", ctypes.c_uint64),
        ("b_rows", ctypes.c_uint64),
        ("b_rows_lds", ctypes.c_uint64),
        ("m4r_lookups", ctypes.c_uint64),
        ("out_rows", ctypes.c_uint64),
        ("fixpoint_ms", ctypes.c_double),
        ("product_ms", ctypes.c_double),
        ("product_launches", ctypes.c_uint32),
        ("host_syncs", ctypes.c_uint32),
        ("mfma_instr", ctypes.c_uint64),
        ("sparse_ms", ctypes.c_double),
        ("dense_ms", ctypes.c_double),
        ("mfma_ms", ctypes.c_double),
        ("sched_scored", ctypes.c_uint64),
        ("sched_deferred", ctypes.c_uint64),
        ("sched_full", ctypes.c_uint64),
        ("flush_steps", ctypes.c_uint32),
        ("max_unit_cands", ctypes.c_uint32),
        ("cand_hits", ctypes.c_uint64),
        ("a_rows_dense", ctypes.c_uint64),
        ("b_rows_dense", ctypes.c_uint64),
        ("out_rows_dense", ctypes.c_uint64),
        ("a_cols", ctypes.c_uint64),
        ("units_done", ctypes.c_uint64),
        ("units_idle", ctypes.c_uint64),
        ("unit_clocks", ctypes.c_uint64),
        ("unit_clocks_idle", ctypes.c_uint64),
        ("b_rows_idle", ctypes.c_uint64),
        ("merge_ms", ctypes.c_double),
        ("tile_products_unchanged", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class Query(ctypes.Structure):
    _fields_ = [
        ("mats", ctypes.POINTER(ctypes.c_void_p)),
        ("nmats", ctypes.c_size_t),
        ("rules", ctypes.POINTER(Rule)),
        ("nrules", ctypes.c_size_t),
    ]


class Pairs(ctypes.Structure):
    _fields_ = [
        ("keys", ctypes.POINTER(ctypes.c_uint64)),
        ("count", ctypes.c_uint64),
        ("cap", ctypes.c_uint64),
    ]


E_CAPACITY = -5

#: every symbol include/cfpq_hip.h declares
EXPORTS = (
    "cfpq_default_opts", "cfpq_create", "cfpq_destroy", "cfpq_last_error", "cfpq_set_opts", "cfpq_sync",
    "cfpq_mat_alloc", "cfpq_mat_free", "cfpq_mat_n", "cfpq_mat_upload", "cfpq_mat_download",
    "cfpq_mat_set_bits", "cfpq_mat_popcount", "cfpq_mat_copy", "cfpq_mat_fingerprint", "cfpq_add_mul", "cfpq_add_mul_noalloc", "cfpq_fixpoint",
    "cfpq_batch_fixpoint", "cfpq_nt_closure", "cfpq_pairs_fixpoint", "cfpq_sched_set_weights", "cfpq_sched_get_weights", "cfpq_sched_init",
    "cfpq_sched_eval", "cfpq_sched_train", "cfpq_sched_trace_begin", "cfpq_sched_trace_read",
    "cfpq_sched_trace_end",
)

NOALLOC_MAX_N = 4096
SCHED_PARAMS = 5825
SCHED_FEATURES = 8

_lib = None


def lib():
    """Load ``libcfpq_hip.so`` (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, sz, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int
    u64p = ctypes.POINTER(ctypes.c_uint64)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    fp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "cfpq_default_opts": (None, [ctypes.POINTER(Opts)]),
        "cfpq_create": (i32, [i32, ctypes.POINTER(Opts), ctypes.POINTER(vp)]),
        "cfpq_destroy": (i32, [vp]),
        "cfpq_last_error": (ctypes.c_char_p, [vp]),
        "cfpq_set_opts": (i32, [vp, ctypes.POINTER(Opts)]),
        "cfpq_sync": (i32, [vp]),
        "cfpq_mat_alloc": (i32, [vp, u32, ctypes.POINTER(vp)]),
        "cfpq_mat_free": (i32, [vp, vp]),
        "cfpq_mat_n": (u32, [vp]),
        "cfpq_mat_upload": (i32, [vp, vp, u64p, sz]),
        "cfpq_mat_download": (i32, [vp, vp, u64p, sz]),
        "cfpq_mat_set_bits": (i32, [vp, vp, u32p, u32p, sz]),
        "cfpq_mat_popcount": (i32, [vp, vp, u64p]),
        "cfpq_mat_copy": (i32, [vp, vp, vp]),
        "cfpq_mat_fingerprint": (i32, [vp, vp, u64p, u64p]),
        "cfpq_add_mul": (i32, [vp, vp, vp, vp, ctypes.POINTER(i32)]),
        "cfpq_add_mul_noalloc": (i32, [vp, vp, vp, vp, ctypes.POINTER(i32)]),
        "cfpq_fixpoint": (i32, [vp, ctypes.POINTER(Rule), sz, ctypes.POINTER(vp), sz, ctypes.POINTER(Stats)]),
        "cfpq_batch_fixpoint": (i32, [vp, ctypes.POINTER(Query), sz, ctypes.POINTER(Stats)]),
        "cfpq_nt_closure": (i32, [vp, ctypes.POINTER(Rule), sz, u32p, u32, ctypes.POINTER(Stats)]),
        "cfpq_pairs_fixpoint": (i32, [vp, ctypes.POINTER(Rule), sz, ctypes.POINTER(Pairs), u32, u32,
                                      ctypes.POINTER(Stats)]),
        "cfpq_sched_set_weights": (i32, [vp, ctypes.POINTER(ctypes.c_uint16), sz]),
        "cfpq_sched_get_weights": (i32, [vp, ctypes.POINTER(ctypes.c_uint16), sz]),
        "cfpq_sched_init": (i32, [vp, u64]),
        "cfpq_sched_eval": (i32, [vp, ctypes.POINTER(ctypes.c_uint16), fp, u32p, fp, sz, fp, fp, fp]),
        "cfpq_sched_train": (i32, [vp, fp, u32p, fp, sz, u32, u32, ctypes.c_float, u64, fp]),
        "cfpq_sched_trace_begin": (i32, [vp, sz]),
        "cfpq_sched_trace_read": (i32, [vp, fp, u32p, fp, sz, ctypes.POINTER(sz)]),
        "cfpq_sched_trace_end": (i32, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def make_opts(kernel="auto", dense_bits=0, mfma_bits=0, m4r_bits=0, max_iterations=0, final_full_pass=False,
              iters_per_sync=0, use_graph=False, sparse_batch=0, sparse_waves_per_cu=0,
              sparse_chunk=0, sched=False, sched_threshold=0, sparse_mode=0, fuse_apply=False,
              schedule=0, m4r_keys=0, sched_order=False, order_threshold=0, split_cands=0,
              split_rows=False, split_piece=0, no_kernel_timing=False) -> Opts:
    o = Opts()
    lib().cfpq_default_opts(ctypes.byref(o))
    o.kernel = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    o.dense_bits = dense_bits
    o.mfma_bits = mfma_bits
    o.m4r_bits = m4r_bits
    o.max_iterations = max_iterations
    o.final_full_pass = int(bool(final_full_pass))
    o.iters_per_sync = iters_per_sync
    o.use_graph = int(bool(use_graph))
    o.sparse_batch = sparse_batch
    o.sparse_waves_per_cu = sparse_waves_per_cu
    o.sparse_chunk = sparse_chunk
    o.sparse_mode = sparse_mode
    o.fuse_apply = int(bool(fuse_apply))
    o.schedule = SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
    o.m4r_keys = m4r_keys
    o.sched_order = int(bool(sched_order))
    o.order_threshold = int(round(order_threshold * 65536)) if isinstance(order_threshold, float) else order_threshold
    o.split_cands = split_cands
    o.split_rows = int(bool(split_rows))
    o.split_piece = split_piece
    o.no_kernel_timing = int(bool(no_kernel_timing))
    o.sched = int(bool(sched))
    # a float in (0, 1) is a probability; an int is already in units of 1/65536
    o.sched_threshold = int(round(sched_threshold * 65536)) if isinstance(sched_threshold, float) else sched_threshold
    return o


def _rows_view(a: np.ndarray):
    assert a.dtype == np.uint64 and a.ndim == 2 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), a.shape[1]


class Engine:
    """One device context (``cfpq_ctx``)."""

    def __init__(self, device: int = 0, **opts):
        self._L = lib()
        self._h = ctypes.c_void_p()
        o = make_opts(**opts)
        rc = self._L.cfpq_create(device, ctypes.byref(o), ctypes.byref(self._h))
        if rc != 0:
            raise CfpqError(rc, "cfpq_create failed")

    def _check(self, rc: int):
        if rc != 0:
            raise CfpqError(rc, self._L.cfpq_last_error(self._h).decode())

    def set_opts(self, **opts):
        o = make_opts(**opts)
        self._check(self._L.cfpq_set_opts(self._h, ctypes.byref(o)))

    def close(self):
        if self._h:
            self._L.cfpq_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def matrix(self, n: int) -> "DeviceMatrix":
        return DeviceMatrix(self, n)

    def sync(self):
        self._check(self._L.cfpq_sync(self._h))

    def add_mul(self, c: "DeviceMatrix", a: "DeviceMatrix", b: "DeviceMatrix") -> bool:
        ch = ctypes.c_int(0)
        self._check(self._L.cfpq_add_mul(self._h, c._h, a._h, b._h, ctypes.byref(ch)))
        return bool(ch.value)

    def add_mul_noalloc(self, c: "DeviceMatrix", a: "DeviceMatrix", b: "DeviceMatrix") -> bool:
        """cfpq_add_mul_noalloc: the same product without device allocation
        (n <= NOALLOC_MAX_N)."""
        ch = ctypes.c_int(0)
        self._check(self._L.cfpq_add_mul_noalloc(self._h, c._h, a._h, b._h, ctypes.byref(ch)))
        return bool(ch.value)

    def fixpoint(self, rules, mats) -> dict:
        rr = (Rule * max(1, len(rules)))(*[Rule(*map(int, r)) for r in rules])
        mm = (ctypes.c_void_p * len(mats))(*[m._h.value for m in mats])
        st = Stats()
        self._check(self._L.cfpq_fixpoint(self._h, rr, len(rules), mm, len(mats), ctypes.byref(st)))
        return st.as_dict()

    def batch_fixpoint(self, queries) -> dict:
        """queries: list of (rules, mats)."""
        keep = []
        qs = (Query * max(1, len(queries)))()
        for i, (rules, mats) in enumerate(queries):
            rr = (Rule * max(1, len(rules)))(*[Rule(*map(int, r)) for r in rules])
            mm = (ctypes.c_void_p * len(mats))(*[m._h.value for m in mats])
            keep += [rr, mm]
            qs[i].mats = mm
            qs[i].nmats = len(mats)
            qs[i].rules = rr
            qs[i].nrules = len(rules)
        st = Stats()
        self._check(self._L.cfpq_batch_fixpoint(self._h, qs, len(queries), ctypes.byref(st)))
        return st.as_dict()

    def nt_closure(self, rules, elems: np.ndarray) -> dict:
        """Single-matrix formulation: elems is an (n, n) uint32 array of
        nonterminal sets (bit x = nonterminal x), closed in place under the
        rules (head, left, right) on the device."""
        assert elems.dtype == np.uint32 and elems.ndim == 2 and elems.shape[0] == elems.shape[1]
        assert elems.flags.c_contiguous
        rr = (Rule * max(1, len(rules)))(*[Rule(*map(int, r)) for r in rules])
        st = Stats()
        self._check(self._L.cfpq_nt_closure(self._h, rr, len(rules),
                                            elems.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                            elems.shape[0], ctypes.byref(st)))
        return st.as_dict()

    def pairs_fixpoint(self, rules, sets, n: int, cap=None) -> tuple[list, dict]:
        """Sparse pair-list formulation: sets[x] is a uint64 array of pairs
        row << 32 | col of nonterminal x (any order, duplicates allowed).
        Returns (closed sets, each sorted without duplicates, stats).  `cap`
        bounds the first attempt's output arrays; when a relation outgrows it
        the call is repeated once with the sizes the library reports."""
        ins = [np.ascontiguousarray(k, dtype=np.uint64) for k in sets]
        rr = (Rule * max(1, len(rules)))(*[Rule(*map(int, r)) for r in rules])
        caps = [max(len(k), cap if cap is not None else 4 * len(k) + 4096) for k in ins]
        for _ in range(2):
            bufs = [np.empty(c, dtype=np.uint64) for c in caps]
            ps = (Pairs * max(1, len(ins)))()
            for x, (k, b) in enumerate(zip(ins, bufs)):
                b[: len(k)] = k
                ps[x].keys = b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
                ps[x].count = len(k)
                ps[x].cap = len(b)
            st = Stats()
            rc = self._L.cfpq_pairs_fixpoint(self._h, rr, len(rules), ps, len(ins), int(n), ctypes.byref(st))
            if rc == E_CAPACITY:
                caps = [max(c, int(ps[x].count)) for x, c in enumerate(caps)]
                continue
            self._check(rc)
            return [b[: int(ps[x].count)].copy() for x, b in enumerate(bufs)], st.as_dict()
        self._check(rc)


    # ---- learned tile scheduler ------------------------------------------
    def sched_set_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.uint16)
        self._check(self._L.cfpq_sched_set_weights(self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                                    w.size))

    def sched_get_weights(self) -> np.ndarray:
        w = np.zeros(SCHED_PARAMS, np.uint16)
        self._check(self._L.cfpq_sched_get_weights(self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                                    w.size))
        return w

    def sched_init(self, seed: int = 0):
        self._check(self._L.cfpq_sched_init(self._h, ctypes.c_uint64(seed)))

    def sched_eval(self, w, x, rule, y, want_grad=True):
        """(logits, loss, grad) of the network with bf16 parameters w."""
        w = np.ascontiguousarray(w, dtype=np.uint16)
        x = np.ascontiguousarray(x, dtype=np.float32)
        rule = np.ascontiguousarray(rule, dtype=np.uint32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        n = y.shape[0]
        logits = np.zeros(max(n, 1), np.float32)
        loss = ctypes.c_float()
        grad = np.zeros(SCHED_PARAMS, np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self._L.cfpq_sched_eval(self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                             x.ctypes.data_as(fp), rule.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                             y.ctypes.data_as(fp), n, logits.ctypes.data_as(fp), ctypes.byref(loss),
                                             grad.ctypes.data_as(fp) if want_grad else None))
        return logits[:n], float(loss.value), (grad if want_grad else None)

    def sched_train(self, x, rule, y, epochs=8, batch=4096, lr=3e-3, seed=0) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        rule = np.ascontiguousarray(rule, dtype=np.uint32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = np.zeros(epochs, np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self._L.cfpq_sched_train(self._h, x.ctypes.data_as(fp),
                                              rule.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                              y.ctypes.data_as(fp), y.shape[0], epochs, batch, ctypes.c_float(lr),
                                              ctypes.c_uint64(seed), loss.ctypes.data_as(fp)))
        return loss

    def trace_begin(self, capacity: int):
        self._check(self._L.cfpq_sched_trace_begin(self._h, capacity))

    def trace_read(self, capacity: int):
        """(x, rule, y, rows recorded) of the trace collected so far."""
        x = np.zeros((capacity, SCHED_FEATURES), np.float32)
        rule = np.zeros(capacity, np.uint32)
        y = np.zeros(capacity, np.float32)
        cnt = ctypes.c_size_t()
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self._L.cfpq_sched_trace_read(self._h, x.ctypes.data_as(fp),
                                                   rule.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                   y.ctypes.data_as(fp), capacity, ctypes.byref(cnt)))
        n = min(cnt.value, capacity)
        return x[:n], rule[:n], y[:n], cnt.value

    def trace_end(self):
        self._check(self._L.cfpq_sched_trace_end(self._h))


class DeviceMatrix:
    """An n x n bit matrix resident on the device (``cfpq_mat``)."""

    def __init__(self, eng: Engine, n: int):
        self.eng = eng
        self.n = int(n)
        self._h = ctypes.c_void_p()
        eng._check(eng._L.cfpq_mat_alloc(eng._h, self.n, ctypes.byref(self._h)))

    def free(self):
        if self._h and self.eng._h:
            self.eng._check(self.eng._L.cfpq_mat_free(self.eng._h, self._h))
        self._h = ctypes.c_void_p()

    @property
    def wpr(self) -> int:
        return (self.n + 63) // 64

    def upload(self, rows: np.ndarray):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        p, w = _rows_view(rows)
        if rows.shape[0] != self.n:
            raise ValueError("row count mismatch")
        self.eng._check(self.eng._L.cfpq_mat_upload(self.eng._h, self._h, p, w))

    def download(self) -> np.ndarray:
        out = np.zeros((self.n, self.wpr), dtype=np.uint64)
        p, w = _rows_view(out)
        self.eng._check(self.eng._L.cfpq_mat_download(self.eng._h, self._h, p, w))
        return out

    def set_bits(self, rows, cols):
        r = np.ascontiguousarray(rows, dtype=np.uint32)
        c = np.ascontiguousarray(cols, dtype=np.uint32)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        self.eng._check(self.eng._L.cfpq_mat_set_bits(self.eng._h, self._h, r.ctypes.data_as(u32p),
                                                      c.ctypes.data_as(u32p), r.shape[0]))

    def popcount(self) -> int:
        v = ctypes.c_uint64()
        self.eng._check(self.eng._L.cfpq_mat_popcount(self.eng._h, self._h, ctypes.byref(v)))
        return int(v.value)

    def copy_from(self, src: "DeviceMatrix"):
        self.eng._check(self.eng._L.cfpq_mat_copy(self.eng._h, self._h, src._h))

    def fingerprint(self) -> tuple[int, int]:
        h = ctypes.c_uint64()
        c = ctypes.c_uint64()
        self.eng._check(self.eng._L.cfpq_mat_fingerprint(self.eng._h, self._h, ctypes.byref(h), ctypes.byref(c)))
        return int(h.value), int(c.value)
