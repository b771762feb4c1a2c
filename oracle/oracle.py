"""ctypes binding of the CPU oracle (oracle/ltl_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

ATOM, NOT, AND, OR, NEXT, FIN, GLOB, UNTIL = range(8)


class Problem(ctypes.Structure):
    _fields_ = [("num_pos", ctypes.c_int32), ("num_neg", ctypes.c_int32),
                ("num_atoms", ctypes.c_int32),
                ("lens", ctypes.POINTER(ctypes.c_uint8)),
                ("atoms", ctypes.POINTER(ctypes.c_uint64)),
                ("names", ctypes.c_char_p),
                ("words_per_trace", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class Options(ctypes.Structure):
    _fields_ = [("cost", ctypes.c_int32 * 8), ("max_cost", ctypes.c_int32),
                ("neg_type", ctypes.c_int32), ("hamming_k", ctypes.c_int32),
                ("dedup_mode", ctypes.c_int32), ("keep_store", ctypes.c_int32),
                ("capacity", ctypes.c_int64), ("threads", ctypes.c_int32),
                ("pad_", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("cost", ctypes.c_int32),
                ("all_ltls", ctypes.c_uint64), ("found_id", ctypes.c_int64),
                ("root_op", ctypes.c_int32), ("root_left", ctypes.c_int32),
                ("root_right", ctypes.c_int32), ("on_the_fly", ctypes.c_int32),
                ("n_stored", ctypes.c_int64), ("n_levels", ctypes.c_int32),
                ("pad_", ctypes.c_int32),
                ("level_start", ctypes.POINTER(ctypes.c_int64)),
                ("level_end", ctypes.POINTER(ctypes.c_int64)),
                ("words", ctypes.POINTER(ctypes.c_uint64)),
                ("op", ctypes.POINTER(ctypes.c_uint8)),
                ("left", ctypes.POINTER(ctypes.c_int32)),
                ("right", ctypes.POINTER(ctypes.c_int32)),
                ("formula", ctypes.c_char_p),
                ("seconds", ctypes.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_enumerate.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(Options),
                                    ctypes.POINTER(Result)]
        L.orc_result_free.argtypes = [ctypes.POINTER(Result)]
        L.orc_eval_formula.argtypes = [ctypes.c_char_p, ctypes.POINTER(Problem),
                                       ctypes.POINTER(ctypes.c_uint64)]
        L.orc_formula_cost.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32)]
        L.orc_bruteforce_min_cost.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(Options),
                                              ctypes.c_int64]
        L.orc_mismatches.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int32, ctypes.c_int32]
        L.orc_relaxed_key.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint8),
                                      ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
        L.orc_read_trace_file.argtypes = [ctypes.c_char_p, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.c_char_p]
        for f in ("orc_op_not", "orc_op_glob"):
            getattr(L, f).argtypes = [ctypes.c_uint64, ctypes.c_int]
            getattr(L, f).restype = ctypes.c_uint64
        for f in ("orc_op_next", "orc_op_fin"):
            getattr(L, f).argtypes = [ctypes.c_uint64]
            getattr(L, f).restype = ctypes.c_uint64
        L.orc_op_until.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_op_until.restype = ctypes.c_uint64
        L.orc_trace_op.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.POINTER(ctypes.c_uint64), ctypes.c_int32, ctypes.c_int32,
                                   ctypes.POINTER(ctypes.c_uint64)]
        _lib = L
    return _lib


def _u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class Instance:
    """A learning instance: atom CS words (atom-major, then trace, then the K
    words of a trace), lengths, class sizes."""

    def __init__(self, num_pos, num_neg, lens, atoms, names=None, K=1):
        self.num_pos, self.num_neg, self.K = int(num_pos), int(num_neg), int(K)
        self.lens = np.ascontiguousarray(lens, dtype=np.uint8)
        self.atoms = np.ascontiguousarray(atoms, dtype=np.uint64).reshape(-1, self.W * self.K)
        self.names = names or "".join(chr(ord("a") + i) for i in range(self.atoms.shape[0]))

    @property
    def W(self):
        return self.num_pos + self.num_neg

    def problem(self):
        self._names_b = self.names.encode()
        return Problem(self.num_pos, self.num_neg, self.atoms.shape[0],
                       self.lens.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                       _u64p(self.atoms), self._names_b, self.K, 0)


def options(cost=(1,) * 8, max_cost=10, neg_type=1, hamming_k=0, dedup_mode=0,
            keep_store=False, capacity=0, threads=1):
    o = Options()
    for i, c in enumerate(cost):
        o.cost[i] = int(c)
    o.max_cost, o.neg_type, o.hamming_k = int(max_cost), int(neg_type), int(hamming_k)
    o.dedup_mode, o.keep_store, o.capacity = int(dedup_mode), int(bool(keep_store)), int(capacity)
    o.threads = int(threads)
    return o


def enumerate_(inst, **kw):
    """Run the oracle enumerator; returns a dict (store arrays when keep_store)."""
    L = lib()
    pb = inst.problem()
    opt = options(**kw)
    res = Result()
    L.orc_enumerate(ctypes.byref(pb), ctypes.byref(opt), ctypes.byref(res))
    out = dict(status=res.status, cost=res.cost, all_ltls=res.all_ltls, found_id=res.found_id,
               root=(res.root_op, res.root_left, res.root_right), on_the_fly=res.on_the_fly,
               n_stored=res.n_stored, formula=(res.formula or b"").decode(), seconds=res.seconds)
    if res.status >= 0:
        nl = res.n_levels
        out["level_start"] = np.ctypeslib.as_array(res.level_start, (nl,)).copy()
        out["level_end"] = np.ctypeslib.as_array(res.level_end, (nl,)).copy()
    if opt.keep_store and res.status >= 0 and res.n_stored:
        n, W = res.n_stored, inst.W * inst.K
        out["words"] = np.ctypeslib.as_array(res.words, (n * W,)).reshape(n, W).copy()
        out["op"] = np.ctypeslib.as_array(res.op, (n,)).copy()
        out["left"] = np.ctypeslib.as_array(res.left, (n,)).copy()
        out["right"] = np.ctypeslib.as_array(res.right, (n,)).copy()
    L.orc_result_free(ctypes.byref(res))
    return out


def eval_formula(inst, formula):
    """Textbook per-position evaluation; returns the CS words or None."""
    out = np.zeros(inst.W * inst.K, dtype=np.uint64)
    pb = inst.problem()
    rc = lib().orc_eval_formula(formula.encode(), ctypes.byref(pb), _u64p(out))
    return out if rc == 0 else None


def formula_cost(formula, cost=(1,) * 8):
    c = (ctypes.c_int32 * 8)(*cost)
    return lib().orc_formula_cost(formula.encode(), c)


def mismatches(inst, cs):
    cs = np.ascontiguousarray(cs, dtype=np.uint64)[::inst.K].copy()
    return lib().orc_mismatches(_u64p(cs), inst.num_pos, inst.num_neg)


def trace_op(op, x, y, K, length):
    """Operator op on one trace of K words (y: right operand of & | U)."""
    x = np.ascontiguousarray(x, dtype=np.uint64)
    y = np.ascontiguousarray(y if y is not None else np.zeros(K), dtype=np.uint64)
    out = np.zeros(K, dtype=np.uint64)
    lib().orc_trace_op(int(op), _u64p(x), _u64p(y), int(K), int(length), _u64p(out))
    return out


def bruteforce_min_cost(inst, max_formulas=2_000_000, **kw):
    pb = inst.problem()
    opt = options(**kw)
    return lib().orc_bruteforce_min_cost(ctypes.byref(pb), ctypes.byref(opt), int(max_formulas))


def relaxed_key(inst, cs, mode):
    cs = np.ascontiguousarray(cs, dtype=np.uint64)
    h, l = ctypes.c_uint64(), ctypes.c_uint64()
    lib().orc_relaxed_key(_u64p(cs), inst.lens.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                          inst.W, int(mode), int(inst.lens.astype(np.int64).sum()),
                          ctypes.byref(h), ctypes.byref(l))
    return h.value, l.value


def read_trace_file(path, max_traces=4096):
    lens = np.zeros(max_traces, dtype=np.uint8)
    atoms = np.zeros(26 * max_traces, dtype=np.uint64)
    names = ctypes.create_string_buffer(32)
    P, N, A = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    rc = lib().orc_read_trace_file(path.encode(), max_traces, ctypes.byref(P), ctypes.byref(N),
                                   ctypes.byref(A), lens.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                   _u64p(atoms), names)
    if rc != 0:
        raise ValueError(f"trace file {path}: error {rc}")
    W = P.value + N.value
    return Instance(P.value, N.value, lens[:W], atoms[:A.value * W].reshape(A.value, W),
                    names.value.decode())


def op_not(x, n): return lib().orc_op_not(x, n)
def op_next(x): return lib().orc_op_next(x)
def op_fin(x): return lib().orc_op_fin(x)
def op_glob(x, n): return lib().orc_op_glob(x, n)
def op_until(a, b): return lib().orc_op_until(a, b)
