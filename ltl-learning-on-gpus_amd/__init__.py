"""MI355X LTL learner: Python binding of the C ABI in include/ltl_amd.h.

The library (libltl_amd.so, built by build.py) holds the HIP kernels and the
C++ host scheduler; this module only marshals arguments.  It mirrors the
reference's in-process entry point

    std::string LTLI(costFun, maxCost, RlxUnqChkTyp, NegType, alphabet,
                     int &LTLcost, uint64 &allLTLs, pos, neg)
                                            (code/ltli6463.cu:1043-1052)

as ``LTLI(...)`` returning ``(formula, LTLcost, allLTLs)``, and exposes the
context API (``Learner``) for repeated calls.  There is no CPU fallback: if
the library is missing or no GPU is visible, calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libltl_amd.so")

FOUND, NOT_FOUND = 1, 0
E_ARG, E_LIMIT, E_OOM, E_HIP, E_COLLISION, E_NODEVICE = -1, -2, -3, -4, -5, -6
ATOM, NOT, AND, OR, NEXT, FIN, GLOB, UNTIL = range(8)
ORDER_PLAIN, ORDER_LEARNED = 0, 1
DEDUP_EXACT, DEDUP_REF1, DEDUP_REF2, DEDUP_REF3 = 0, 1, 2, 3
NOALLOC_SCRATCH_BYTES = 128 << 20   # LTL_NOALLOC_SCRATCH_BYTES of include/ltl_amd.h


class LtlError(RuntimeError):
    pass


class Problem(ctypes.Structure):
    _fields_ = [("num_pos", ctypes.c_int32), ("num_neg", ctypes.c_int32),
                ("num_atoms", ctypes.c_int32), ("words_per_trace", ctypes.c_int32),
                ("lens", ctypes.POINTER(ctypes.c_uint8)),
                ("atoms", ctypes.POINTER(ctypes.c_uint64)),
                ("names", ctypes.c_char_p)]


class Options(ctypes.Structure):
    _fields_ = [("cost", ctypes.c_uint16 * 8), ("max_cost", ctypes.c_uint16),
                ("neg_type", ctypes.c_uint8), ("dedup_mode", ctypes.c_uint8),
                ("hamming_k", ctypes.c_int32), ("order", ctypes.c_int32),
                ("hash_slots_log2", ctypes.c_int32), ("precheck", ctypes.c_int32),
                ("group_lanes", ctypes.c_int32), ("capacity", ctypes.c_int64),
                ("hbm_budget_bytes", ctypes.c_uint64), ("device", ctypes.c_int32),
                ("profile", ctypes.c_int32), ("profile_kinds", ctypes.c_uint32), ("reserved", ctypes.c_int32 * 5)]


class Node(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("atom", ctypes.c_int32),
                ("left", ctypes.c_int32), ("right", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("cost", ctypes.c_int32),
                ("all_ltls", ctypes.c_uint64), ("found_id", ctypes.c_int64),
                ("n_stored", ctypes.c_int64), ("on_the_fly", ctypes.c_int32),
                ("n_nodes", ctypes.c_int32), ("nodes", ctypes.POINTER(Node)),
                ("formula", ctypes.c_char_p), ("seconds", ctypes.c_double),
                ("kernel_seconds", ctypes.c_double), ("hbm_bytes", ctypes.c_uint64),
                ("candidates", ctypes.c_uint64), ("checked", ctypes.c_uint64),
                ("words_deferred", ctypes.c_int64), ("deferred_succ", ctypes.c_int32),
                ("retries", ctypes.c_int32), ("words_unwritten", ctypes.c_int64)]


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 24), ("launches", ctypes.c_uint64),
                ("seconds", ctypes.c_double), ("alg_bytes", ctypes.c_uint64),
                ("min_bytes", ctypes.c_uint64), ("items", ctypes.c_uint64)]


class BatchParams(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int32), ("word_block", ctypes.c_int32),
                ("workgroups", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("defer_limit", ctypes.c_int64)]


class BatchTrial(ctypes.Structure):
    _fields_ = [("params", BatchParams), ("seconds", ctypes.c_double),
                ("deferred", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class BatchStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("kernel_seconds", ctypes.c_double),
                ("n", ctypes.c_int32), ("on_device", ctypes.c_int32),
                ("deferred", ctypes.c_int32), ("host", ctypes.c_int32)]


class LaunchRecord(ctypes.Structure):
    _fields_ = [("cost", ctypes.c_int32), ("op", ctypes.c_int32), ("pass_", ctypes.c_int32),
                ("launched", ctypes.c_int32), ("all_ltls", ctypes.c_uint64), ("stored", ctypes.c_int64),
                ("to_check", ctypes.c_uint64)]


LAUNCH_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(LaunchRecord), ctypes.c_void_p)


class GenInstance(ctypes.Structure):
    _fields_ = [("num_pos", ctypes.c_int32), ("num_neg", ctypes.c_int32),
                ("num_atoms", ctypes.c_int32), ("max_cost", ctypes.c_int32),
                ("hamming_k", ctypes.c_int32), ("neg_type", ctypes.c_int32),
                ("lens", ctypes.POINTER(ctypes.c_uint8)),
                ("atoms", ctypes.POINTER(ctypes.c_uint64)),
                ("names", ctypes.c_char * 32), ("planted", ctypes.c_char * 512)]


EXPORTS = ["ltl_default_options", "ltl_create", "ltl_set_options", "ltl_learn", "ltl_learn_batch",
           "ltl_set_batch_params", "ltl_get_batch_params", "ltl_batch_last_stats", "ltl_autotune_batch",
           "ltl_result_free", "ltl_destroy", "ltl_get_kernel_stats", "ltl_store_read", "ltl_store_summaries", "ltl_levels",
           "ltl_read_trace_file", "ltl_read_trace_file_words", "ltl_eval_formula", "ltl_mismatches", "ltl_version",
           "ltl_gen_config", "ltl_gen_custom", "ltl_gen_free",
           "ltl_scorer_create", "ltl_scorer_destroy", "ltl_scorer_get_params", "ltl_scorer_set_params",
           "ltl_scorer_forward", "ltl_scorer_grad", "ltl_scorer_train", "ltl_set_scorer", "ltl_set_log",
           "ltl_get_log", "ltl_learn_dc", "ltl_get_options", "ltl_learn_noalloc", "ltl_read_trace_file_names",
           "ltl_learn_dc_batch", "ltl_match", "ltl_set_launch_log"]

DC_DET, DC_RAND = 0, 1

SCORER_FEATURES = 24
SCORER_PARAMS = 128 * 32 + 128 + 128 * 128 + 128 + 128 + 1

_lib = None


def lib():
    """Load libltl_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LtlError(f"{LIB_PATH} is missing: run build.py (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    L.ltl_default_options.argtypes = [P(Options)]
    L.ltl_create.argtypes = [P(Options), P(ctypes.c_int32)]
    L.ltl_create.restype = ctypes.c_void_p
    L.ltl_set_options.argtypes = [ctypes.c_void_p, P(Options)]
    L.ltl_learn.argtypes = [ctypes.c_void_p, P(Problem), P(Result)]
    L.ltl_learn_batch.argtypes = [ctypes.c_void_p, P(Problem), ctypes.c_int32, P(Result)]
    L.ltl_learn_noalloc.argtypes = [ctypes.c_void_p, P(Problem), P(Result)]
    L.ltl_result_free.argtypes = [P(Result)]
    L.ltl_set_batch_params.argtypes = [ctypes.c_void_p, P(BatchParams)]
    L.ltl_get_batch_params.argtypes = [ctypes.c_void_p, P(BatchParams)]
    L.ltl_batch_last_stats.argtypes = [ctypes.c_void_p, P(BatchStats)]
    L.ltl_autotune_batch.argtypes = [ctypes.c_void_p, P(Problem), ctypes.c_int32, P(BatchTrial), ctypes.c_int32]
    L.ltl_destroy.argtypes = [ctypes.c_void_p]
    L.ltl_get_kernel_stats.argtypes = [ctypes.c_void_p, P(KernelStat), ctypes.c_int32]
    L.ltl_store_read.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, P(ctypes.c_uint64),
                                 P(ctypes.c_uint8), P(ctypes.c_int32), P(ctypes.c_int32)]
    L.ltl_store_summaries.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, P(ctypes.c_uint64)]
    L.ltl_levels.argtypes = [ctypes.c_void_p, P(ctypes.c_int64), P(ctypes.c_int64), ctypes.c_int32]
    L.ltl_read_trace_file.argtypes = [ctypes.c_char_p, ctypes.c_int32, P(ctypes.c_int32), P(ctypes.c_int32),
                                      P(ctypes.c_int32), P(ctypes.c_uint8), P(ctypes.c_uint64), ctypes.c_char_p]
    L.ltl_read_trace_file_words.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, P(ctypes.c_int32),
                                            P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int32),
                                            P(ctypes.c_uint8), P(ctypes.c_uint64), ctypes.c_char_p]
    L.ltl_read_trace_file_names.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, P(ctypes.c_int32),
                                            P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int32),
                                            P(ctypes.c_uint8), P(ctypes.c_uint64), ctypes.c_char_p, ctypes.c_int32]
    L.ltl_eval_formula.argtypes = [ctypes.c_char_p, P(Problem), P(ctypes.c_uint64)]
    L.ltl_mismatches.argtypes = [P(ctypes.c_uint64), ctypes.c_int32, ctypes.c_int32]
    L.ltl_version.restype = ctypes.c_char_p
    L.ltl_gen_config.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, P(GenInstance)]
    L.ltl_gen_custom.argtypes = [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                 ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_char_p,
                                 ctypes.c_int32, P(GenInstance)]
    L.ltl_gen_free.argtypes = [P(GenInstance)]
    fp = P(ctypes.c_float)
    L.ltl_scorer_create.argtypes = [ctypes.c_int32, ctypes.c_uint64, P(ctypes.c_int32)]
    L.ltl_scorer_create.restype = ctypes.c_void_p
    L.ltl_scorer_destroy.argtypes = [ctypes.c_void_p]
    L.ltl_scorer_get_params.argtypes = [ctypes.c_void_p, fp]
    L.ltl_scorer_set_params.argtypes = [ctypes.c_void_p, fp]
    L.ltl_scorer_forward.argtypes = [ctypes.c_void_p, fp, ctypes.c_int64, fp]
    L.ltl_scorer_grad.argtypes = [ctypes.c_void_p, fp, fp, ctypes.c_int64, fp, fp]
    L.ltl_scorer_train.argtypes = [ctypes.c_void_p, fp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_float, ctypes.c_uint64, fp]
    L.ltl_set_scorer.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.ltl_set_log.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32]
    L.ltl_get_log.argtypes = [ctypes.c_void_p, fp, ctypes.c_int64]
    L.ltl_get_log.restype = ctypes.c_int64
    L.ltl_learn_dc.argtypes = [ctypes.c_void_p, P(Problem), ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64,
                               P(Result)]
    L.ltl_learn_dc_batch.argtypes = [ctypes.c_void_p, P(Problem), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_uint64, P(Result)]
    L.ltl_match.argtypes = [ctypes.c_void_p, P(Problem), P(ctypes.c_char_p), ctypes.c_int32,
                            P(ctypes.c_uint8)]
    L.ltl_get_options.argtypes = [ctypes.c_void_p, P(Options)]
    L.ltl_set_launch_log.argtypes = [ctypes.c_void_p, LAUNCH_FN, ctypes.c_void_p]
    _lib = L
    return L


def _u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class Instance:
    """Atom characteristic sets (atom-major, one uint64 per trace, or K per
    trace for traces longer than 64 positions: word k = positions
    64k..64k+63), lengths, class sizes (positives first) and atom letters."""

    def __init__(self, num_pos, num_neg, lens, atoms, names=None, planted=None,
                 max_cost=None, hamming_k=0, neg_type=1, K=1):
        self.num_pos, self.num_neg, self.K = int(num_pos), int(num_neg), int(K)
        self.lens = np.ascontiguousarray(lens, dtype=np.uint8)
        self.atoms = np.ascontiguousarray(atoms, dtype=np.uint64).reshape(-1, self.W * self.K)
        A = self.atoms.shape[0]
        # (the library's names: letters up to 26 atoms, else p0, p1, ... as a
        # comma-separated list)
        self.names = names or ("".join(chr(ord("a") + i) for i in range(A)) if A <= 26
                               else ",".join(f"p{i}" for i in range(A)))
        self.planted, self.max_cost, self.hamming_k, self.neg_type = planted, max_cost, hamming_k, neg_type

    @property
    def W(self):
        return self.num_pos + self.num_neg

    def problem(self):
        self._names_b = self.names.encode()
        return Problem(self.num_pos, self.num_neg, self.atoms.shape[0], self.K,
                       self.lens.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                       _u64p(self.atoms), self._names_b)


def _from_gen(g):
    W = g.num_pos + g.num_neg
    lens = np.ctypeslib.as_array(g.lens, (W,)).copy()
    atoms = np.ctypeslib.as_array(g.atoms, (g.num_atoms * W,)).reshape(g.num_atoms, W).copy()
    inst = Instance(g.num_pos, g.num_neg, lens, atoms, g.names.decode(), g.planted.decode(),
                    g.max_cost, g.hamming_k, g.neg_type)
    lib().ltl_gen_free(ctypes.byref(g))
    return inst


def generate(config, seed=1, index=0):
    """Workload `config` (1..5) from the seeded generator (include/ltl_gen.h)."""
    g = GenInstance()
    rc = lib().ltl_gen_config(int(config), ctypes.c_uint64(seed), int(index), ctypes.byref(g))
    if rc != 0:
        raise LtlError(f"ltl_gen_config({config}) -> {rc}")
    return _from_gen(g)


def generate_custom(seed, num_pos, num_neg, num_atoms, len_min, len_max, rate=0.5,
                    planted=None, planted_size=5):
    g = GenInstance()
    rc = lib().ltl_gen_custom(ctypes.c_uint64(seed), num_pos, num_neg, num_atoms, len_min, len_max,
                              float(rate), planted.encode() if planted else None, int(planted_size),
                              ctypes.byref(g))
    if rc != 0:
        raise LtlError(f"ltl_gen_custom -> {rc}")
    return _from_gen(g)


def read_trace_file(path, max_traces=4096):
    """Reference trace-file format (code/ltli6463.cu:1500-1573), traces of up
    to 255 positions (the reference stops at 63, :1078): one word per trace
    when the longest fits 64 positions, else K = ceil(longest / 64)."""
    lens = np.zeros(max_traces, dtype=np.uint8)
    atoms = np.zeros(64 * max_traces * 4, dtype=np.uint64)
    names = ctypes.create_string_buffer(64 * 40)
    P, N, A, M = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    # (an alphabet line of comma-separated names names more than 26
    # propositions; the names come back as a comma-separated list)
    rc = lib().ltl_read_trace_file_names(path.encode(), max_traces, 4, ctypes.byref(P), ctypes.byref(N),
                                         ctypes.byref(A), ctypes.byref(M),
                                         lens.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _u64p(atoms), names,
                                         len(names))
    if rc != 0:
        raise LtlError(f"{path}: error {rc}")
    W, K = P.value + N.value, (max(M.value, 1) + 63) // 64
    words = atoms[:A.value * W * 4].reshape(A.value, W, 4)[:, :, :K]
    return Instance(P.value, N.value, lens[:W], words.reshape(A.value, W * K), names.value.decode(), K=K)


def eval_formula(inst, formula):
    out = np.zeros(inst.W * inst.K, dtype=np.uint64)
    pb = inst.problem()
    rc = lib().ltl_eval_formula(formula.encode(), ctypes.byref(pb), _u64p(out))
    if rc != 0:
        raise LtlError(f"cannot parse {formula!r}")
    return out


def mismatches(inst, cs):
    cs = np.ascontiguousarray(np.asarray(cs, dtype=np.uint64)[::inst.K])
    return lib().ltl_mismatches(_u64p(cs), inst.num_pos, inst.num_neg)


def make_options(cost=(1,) * 8, max_cost=10, neg_type=1, hamming_k=0, order=ORDER_PLAIN,
                 hash_slots_log2=0, precheck=1, capacity=0, budget=0, device=0, profile=0, dedup_mode=0,
                 profile_kinds=0):
    o = Options()
    lib().ltl_default_options(ctypes.byref(o))
    for i, c in enumerate(cost):
        o.cost[i] = int(c)
    o.max_cost, o.neg_type, o.hamming_k, o.order = int(max_cost), int(neg_type), int(hamming_k), int(order)
    o.dedup_mode = int(dedup_mode)
    o.hash_slots_log2, o.precheck, o.capacity = int(hash_slots_log2), int(precheck), int(capacity)
    o.hbm_budget_bytes, o.device, o.profile = int(budget), int(device), int(profile)
    o.profile_kinds = int(profile_kinds)
    return o


class Learner:
    """A context (ltl_create ... ltl_destroy) with reusable device memory."""

    def __init__(self, budget=1 << 30, **opts):
        self._opts = dict(budget=budget, **opts)
        st = ctypes.c_int32(0)
        o = make_options(**self._opts)
        self._ctx = lib().ltl_create(ctypes.byref(o), ctypes.byref(st))
        if not self._ctx:
            raise LtlError(f"ltl_create failed: {st.value}")

    def close(self):
        if self._ctx:
            lib().ltl_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_options(self, **opts):
        """Apply the constructor's options overridden by `opts` (not sticky)."""
        o = make_options(**dict(self._opts, **opts))
        rc = lib().ltl_set_options(self._ctx, ctypes.byref(o))
        if rc != 0:
            raise LtlError(f"ltl_set_options -> {rc}")

    def learn(self, inst, **opts):
        """ltl_learn with the constructor's options overridden by `opts`."""
        return self._learn(lib().ltl_learn, "ltl_learn", inst, **opts)

    def learn_noalloc(self, inst, **opts):
        """ltl_learn_noalloc (no device allocation: the library's static
        scratch region of NOALLOC_SCRATCH_BYTES per GPU)."""
        return self._learn(lib().ltl_learn_noalloc, "ltl_learn_noalloc", inst, **opts)

    def _learn(self, fn, name, inst, **opts):
        self.set_options(**opts)
        pb = inst.problem()
        res = Result()
        rc = fn(self._ctx, ctypes.byref(pb), ctypes.byref(res))
        out = dict(status=res.status, cost=res.cost, all_ltls=res.all_ltls, found_id=res.found_id,
                   n_stored=res.n_stored, on_the_fly=res.on_the_fly,
                   formula=(res.formula or b"").decode(), seconds=res.seconds,
                   kernel_seconds=res.kernel_seconds, hbm_bytes=res.hbm_bytes,
                   candidates=res.candidates, checked=res.checked, words_deferred=res.words_deferred,
                   words_unwritten=res.words_unwritten,
                   deferred_succ=res.deferred_succ, retries=res.retries,
                   nodes=[(res.nodes[i].op, res.nodes[i].atom, res.nodes[i].left, res.nodes[i].right)
                          for i in range(res.n_nodes)])
        lib().ltl_result_free(ctypes.byref(res))
        if rc < 0:
            raise LtlError(f"{name} -> {rc}")
        return out

    def learn_dc(self, inst, window=64, algo=DC_DET, seed=0, **opts):
        """Divide and conquer over windows of at most `window` traces
        (ltl_learn_dc; the reference's DetSplit / RandSplit)."""
        self.set_options(**opts)
        pb = inst.problem()
        res = Result()
        rc = lib().ltl_learn_dc(self._ctx, ctypes.byref(pb), int(window), int(algo), int(seed), ctypes.byref(res))
        out = self._result(res)
        out["calls"] = res.n_stored
        lib().ltl_result_free(ctypes.byref(res))
        if rc < 0:
            raise LtlError(f"ltl_learn_dc -> {rc}")
        return out

    def learn_dc_batch(self, insts, window=64, algo=DC_DET, seed=0, **opts):
        """Divide and conquer over several instances in lock-step
        (ltl_learn_dc_batch): the pending windows of all instances learned
        together by ltl_learn_batch, formula matches on the device.  Returns
        one dict per instance, as learn_dc; an instance that fails has its
        LTL_E_* code as status (the others are unaffected)."""
        self.set_options(**opts)
        n = len(insts)
        pbs = (Problem * n)(*[i.problem() for i in insts])
        res = (Result * n)()
        rc = lib().ltl_learn_dc_batch(self._ctx, pbs, n, int(window), int(algo), int(seed), res)
        out = []
        for i in range(n):
            out.append(self._result(res[i]))
            out[-1]["calls"] = res[i].n_stored
            lib().ltl_result_free(ctypes.byref(res[i]))
        if rc < 0:
            raise LtlError(f"ltl_learn_dc_batch -> {rc}")
        return out

    def match(self, inst, formulas):
        """Value at position 0 of each formula on every trace, evaluated on
        the device (ltl_match; the reference's LTLmatch, code/dc.py:185):
        uint8 array [len(formulas), W]."""
        n = len(formulas)
        arr = (ctypes.c_char_p * max(n, 1))(*[f.encode() for f in formulas])
        out = np.zeros((n, inst.W), dtype=np.uint8)
        pb = inst.problem()
        rc = lib().ltl_match(self._ctx, ctypes.byref(pb), arr, n,
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        if rc != 0:
            raise LtlError(f"ltl_match -> {rc}")
        return out

    @staticmethod
    def _result(res):
        return dict(status=res.status, cost=res.cost, all_ltls=res.all_ltls, found_id=res.found_id,
                    n_stored=res.n_stored, on_the_fly=res.on_the_fly,
                    formula=(res.formula or b"").decode(), seconds=res.seconds,
                    nodes=[(res.nodes[i].op, res.nodes[i].atom, res.nodes[i].left, res.nodes[i].right)
                           for i in range(res.n_nodes)])

    def learn_batch(self, insts, **opts):
        """ltl_learn_batch over a list of instances; returns a list of dicts."""
        self.set_options(**opts)
        n = len(insts)
        pbs = (Problem * n)(*[i.problem() for i in insts])
        res = (Result * n)()
        rc = lib().ltl_learn_batch(self._ctx, pbs, n, res)
        out = [self._result(res[i]) for i in range(n)]
        for i in range(n):
            lib().ltl_result_free(ctypes.byref(res[i]))
        if rc < 0:
            raise LtlError(f"ltl_learn_batch -> {rc}")
        return out

    def set_batch_params(self, tile=1024, word_block=4, workgroups=0, defer_limit=1 << 18):
        p = BatchParams(tile, word_block, workgroups, 0, defer_limit)
        rc = lib().ltl_set_batch_params(self._ctx, ctypes.byref(p))
        if rc != 0:
            raise LtlError(f"ltl_set_batch_params -> {rc}")

    def batch_params(self):
        p = BatchParams()
        lib().ltl_get_batch_params(self._ctx, ctypes.byref(p))
        return dict(tile=p.tile, word_block=p.word_block, workgroups=p.workgroups, defer_limit=p.defer_limit)

    def batch_stats(self):
        s = BatchStats()
        lib().ltl_batch_last_stats(self._ctx, ctypes.byref(s))
        return dict(seconds=s.seconds, kernel_seconds=s.kernel_seconds, n=s.n, on_device=s.on_device,
                    deferred=s.deferred, host=s.host)

    def autotune_batch(self, insts, **opts):
        """Project autotuner: returns the trials; the fastest stays set."""
        self.set_options(**opts)
        n = len(insts)
        pbs = (Problem * n)(*[i.problem() for i in insts])
        trials = (BatchTrial * 64)()
        k = lib().ltl_autotune_batch(self._ctx, pbs, n, trials, 64)
        if k < 0:
            raise LtlError(f"ltl_autotune_batch -> {k}")
        return [dict(tile=t.params.tile, word_block=t.params.word_block, defer_limit=t.params.defer_limit,
                     seconds=t.seconds, deferred=t.deferred) for t in trials[:k]]

    def kernel_stats(self):
        arr = (KernelStat * 32)()
        n = lib().ltl_get_kernel_stats(self._ctx, arr, 32)
        return {arr[i].name.decode(): dict(launches=arr[i].launches, seconds=arr[i].seconds,
                                           alg_bytes=arr[i].alg_bytes, min_bytes=arr[i].min_bytes,
                                           items=arr[i].items) for i in range(min(n, 32))}

    def set_scorer(self, scorer):
        """Attach a Scorer (used with order=ORDER_LEARNED); None detaches."""
        rc = lib().ltl_set_scorer(self._ctx, scorer._sc if scorer is not None else None)
        if rc != 0:
            raise LtlError(f"ltl_set_scorer -> {rc}")
        self._scorer = scorer  # keep it alive while attached

    def launch_log(self, enable=True):
        """Per-launch log (ltl_set_launch_log; the reference's "Cost N | (op)
        | AllLTLs | StoredLTLs | ToBeChecked" lines): while enabled, the
        records of every learn call are appended to self.launches (dicts:
        cost, op, pass (1 generation, 0 test only), launched, all_ltls,
        stored, to_check)."""
        self.launches = []
        if enable:
            def cb(rec, _user):
                r = rec.contents
                self.launches.append(dict(cost=r.cost, op=r.op, pass_=r.pass_, launched=r.launched,
                                          all_ltls=r.all_ltls, stored=r.stored, to_check=r.to_check))
            self._launch_cb = LAUNCH_FN(cb)
            rc = lib().ltl_set_launch_log(self._ctx, self._launch_cb, None)
        else:
            rc = lib().ltl_set_launch_log(self._ctx, LAUNCH_FN(), None)
            self._launch_cb = None
        if rc != 0:
            raise LtlError(f"ltl_set_launch_log -> {rc}")

    def set_log(self, max_rows, per_block=64):
        """Log up to per_block tile-pair rows per binary block of every level."""
        rc = lib().ltl_set_log(self._ctx, int(max_rows), int(per_block))
        if rc != 0:
            raise LtlError(f"ltl_set_log -> {rc}")
        self._log_cap = int(max_rows)

    def get_log(self):
        """Rows (n, 25) of the last call: 24 features, then the label."""
        cap = getattr(self, "_log_cap", 0)
        rows = np.zeros((max(cap, 1), SCORER_FEATURES + 1), dtype=np.float32)
        n = lib().ltl_get_log(self._ctx, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), cap)
        if n < 0:
            raise LtlError(f"ltl_get_log -> {n}")
        return rows[:min(n, cap)].copy()

    def levels(self, max_levels=64):
        s = np.zeros(max_levels, dtype=np.int64)
        e = np.zeros(max_levels, dtype=np.int64)
        n = lib().ltl_levels(self._ctx, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                             e.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), max_levels)
        return s[:n], e[:n]

    def summaries(self, first, count, W):
        """Per-trace summaries of stored sets [first, first + count): an
        array (count, 4, ceil(W / 64)) -- position 0, position 1, some
        position, every position -- in trace order."""
        nsw = (W + 63) // 64
        out = np.zeros(count * 4 * nsw, dtype=np.uint64)
        rc = lib().ltl_store_summaries(self._ctx, first, count, _u64p(out))
        if rc != 0:
            raise LtlError(f"ltl_store_summaries -> {rc}")
        return out.reshape(count, 4, nsw)

    def store(self, first, count, W):
        """Stored sets [first, first + count): W words each (traces x words
        per trace)."""
        cs = np.zeros(count * W, dtype=np.uint64)
        op = np.zeros(count, dtype=np.uint8)
        left = np.zeros(count, dtype=np.int32)
        right = np.zeros(count, dtype=np.int32)
        rc = lib().ltl_store_read(self._ctx, first, count, _u64p(cs),
                                  op.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                  left.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                  right.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        if rc != 0:
            raise LtlError(f"ltl_store_read -> {rc}")
        return cs.reshape(count, W), op, left, right


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Scorer:
    """The learned-order scorer: BF16 MLP 24 -> 128 -> 128 -> 1 on MFMA."""

    def __init__(self, seed=1, device=0):
        st = ctypes.c_int32(0)
        self._sc = lib().ltl_scorer_create(int(device), int(seed), ctypes.byref(st))
        if not self._sc:
            raise LtlError(f"ltl_scorer_create failed: {st.value}")

    def close(self):
        if self._sc:
            lib().ltl_scorer_destroy(self._sc)
            self._sc = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(SCORER_PARAMS, dtype=np.float32)
        rc = lib().ltl_scorer_get_params(self._sc, _fp(p))
        if rc != 0:
            raise LtlError(f"ltl_scorer_get_params -> {rc}")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (SCORER_PARAMS,):
            raise ValueError("parameter vector of the wrong size")
        rc = lib().ltl_scorer_set_params(self._sc, _fp(p))
        if rc != 0:
            raise LtlError(f"ltl_scorer_set_params -> {rc}")

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, SCORER_FEATURES)
        y = np.zeros(x.shape[0], dtype=np.float32)
        rc = lib().ltl_scorer_forward(self._sc, _fp(x), x.shape[0], _fp(y))
        if rc != 0:
            raise LtlError(f"ltl_scorer_forward -> {rc}")
        return y

    def grad(self, x, t):
        """(gradient of mean (y - t)^2 over the parameter vector, loss)."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, SCORER_FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        g = np.zeros(SCORER_PARAMS, dtype=np.float32)
        loss = ctypes.c_float(0)
        rc = lib().ltl_scorer_grad(self._sc, _fp(x), _fp(t), x.shape[0], _fp(g), ctypes.byref(loss))
        if rc != 0:
            raise LtlError(f"ltl_scorer_grad -> {rc}")
        return g, loss.value

    def train(self, rows, epochs=20, batch=1024, lr=1e-3, seed=1):
        """Adam on log rows (n, 25); returns the mean batch loss of the last epoch."""
        rows = np.ascontiguousarray(rows, dtype=np.float32).reshape(-1, SCORER_FEATURES + 1)
        loss = ctypes.c_float(0)
        rc = lib().ltl_scorer_train(self._sc, _fp(rows), rows.shape[0], int(epochs), int(batch), float(lr),
                                    int(seed), ctypes.byref(loss))
        if rc != 0:
            raise LtlError(f"ltl_scorer_train -> {rc}")
        return loss.value


def _instance_from_tokens(pos, neg, alphabet):
    """Encode the reference's token traces (code/ltli6463.cu:1100-1133); K
    words per trace when the longest trace exceeds 64 positions."""
    names = "".join(sorted(set(alphabet)))
    traces = list(pos) + list(neg)
    W = len(traces)
    longest = max((len(t) for t in traces), default=1)
    K = (max(longest, 1) + 63) // 64
    lens = np.array([min(len(t), 255) for t in traces], dtype=np.uint8)
    atoms = np.zeros((len(names), W, K), dtype=np.uint64)
    for a, ch in enumerate(names):
        for i, tr in enumerate(traces):
            w = 0
            for k, tok in enumerate(tr):
                if ch in tok:
                    w |= 1 << k
            for j in range(K):
                atoms[a, i, j] = (w >> (64 * j)) & 0xFFFFFFFFFFFFFFFF
    return Instance(len(pos), len(neg), lens, atoms.reshape(len(names), W * K), names, K=K)


def LTLI(costFun, maxCost, RlxUnqChkTyp, NegType, alphabet, pos, neg, learner=None, reference_keys=False):
    """Drop-in for the reference's LTLI() (code/ltli6463.cu:1043-1052).

    pos/neg: lists of traces, a trace a list of tokens, a token a string of
    the letters holding at that position.  Returns (formula, LTLcost,
    allLTLs) where formula is "not_found" or "see_the_error" on failure, as
    in the reference.  Dedup is exact by default; reference_keys=True uses
    the reference's relaxed uniqueness key RlxUnqChkTyp (1-3), which is lossy
    above 126 total trace positions (code/ltli6463.cu:57-141).
    """
    if any(len(t) > 255 for t in list(pos) + list(neg)):
        return "see_the_error", int(costFun[0]), 0
    inst = _instance_from_tokens(pos, neg, alphabet)
    if inst.W == 0 or (inst.K > 1 and reference_keys):
        return "see_the_error", int(costFun[0]), 0
    if reference_keys and int(RlxUnqChkTyp) not in (1, 2, 3):
        return "see_the_error", int(costFun[0]), 0
    own = learner is None
    L = learner or Learner(budget=1 << 30)
    try:
        r = L.learn(inst, cost=tuple(costFun), max_cost=maxCost, neg_type=NegType,
                    dedup_mode=int(RlxUnqChkTyp) if reference_keys else DEDUP_EXACT)
    except LtlError:
        return "see_the_error", int(costFun[0]), 0
    finally:
        if own:
            L.close()
    return r["formula"], r["cost"], r["all_ltls"]


def collect_logs(learner=None, per_block=64, config1_seeds=range(1, 21), config2_max_cost=7,
                 config3_max_cost=6, max_rows=1 << 20, config5_seeds=(), config5_max_cost=10):
    """Enumeration logs of seeded runs of configurations 1-3 and of the
    configuration-5 instances of config5_seeds (plain order, each up to its
    planted size or config5_max_cost): rows (n, 25) of tile-pair features and
    labels for Scorer.train."""
    own = learner is None
    L = learner or Learner(budget=8 << 30)
    out = []
    try:
        L.set_log(max_rows, per_block)
        for s in config1_seeds:
            L.learn(generate(1, seed=s), max_cost=10, order=ORDER_PLAIN)
            out.append(L.get_log())
        L.learn(generate(2, seed=1), max_cost=config2_max_cost, order=ORDER_PLAIN)
        out.append(L.get_log())
        inst3 = generate(3, seed=1)
        L.learn(inst3, max_cost=config3_max_cost, hamming_k=inst3.hamming_k, order=ORDER_PLAIN)
        out.append(L.get_log())
        for s in config5_seeds:
            for idx in range(16):
                inst = generate(5, seed=s, index=idx)
                L.learn(inst, max_cost=min(inst.max_cost, config5_max_cost), order=ORDER_PLAIN)
                out.append(L.get_log())
        L.set_log(0, 0)
    finally:
        if own:
            L.close()
    return np.concatenate(out) if out else np.zeros((0, SCORER_FEATURES + 1), np.float32)


def train_scorer(rows=None, seed=1, epochs=20, batch=512, lr=1e-3, learner=None):
    """A Scorer trained on the device from enumeration logs (collect_logs()
    when rows is None).  Returns (scorer, final mean batch loss, rows used)."""
    if rows is None:
        rows = collect_logs(learner=learner)
    sc = Scorer(seed=seed)
    loss = sc.train(rows, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return sc, loss, len(rows)


def first_hit_ranks(learner, inst, scorers, max_cost, max_rows=1 << 18, **opts):
    """Ranking quality on the last level of a plain-order run: the level's
    tile pairs (all of them, from the enumeration log) in canonical order and
    in each scorer's descending order; returns, per order, the position of
    the first tile pair that holds an accepted candidate (None if the run
    found nothing in a binary block), and the number of tile pairs."""
    learner.set_log(max_rows, max_rows)
    try:
        r = learner.learn(inst, max_cost=max_cost, order=ORDER_PLAIN, **opts)
        rows = learner.get_log()
    finally:
        learner.set_log(0, 0)
    if r["status"] != 1 or len(rows) == 0:
        return None
    last = rows[np.isclose(rows[:, 19] * 32, r["cost"])]
    hit = last[:, 24] >= 1.0 - opts.get("hamming_k", 0) / inst.W - 1e-6
    if not hit.any():
        return None
    out = {"tile_pairs": int(len(last)), "canonical": int(np.argmax(hit))}
    for name, sc in scorers.items():
        order = np.argsort(-sc.forward(last[:, :SCORER_FEATURES]), kind="stable")
        out[name] = int(np.argmax(hit[order]))
    return out
