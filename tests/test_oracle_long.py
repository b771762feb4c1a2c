"""Pins of the oracle's multi-word traces (longer than 64 positions; no GPU).

A trace of length up to 64K takes K words (word k = positions 64k..64k+63).
The reference stops at 63 positions (code/ltli6463.cu:1078), so nothing of
its own covers this; the pins are what the mathematics fixes:
  - each operator on K words against the per-position finite-trace
    definitions (Python integers of any width), at lengths around every word
    boundary, with runs of ones across the boundaries so that every carry
    matters;
  - a plausible slip in each carry, planted in a CPU copy, fails that pin;
  - the same instance with a zero word appended to every trace enumerates
    exactly as with one word (formula, counts, every stored set);
  - minimal costs against brute force over all syntactic formulas;
  - every stored set of an enumeration against the textbook evaluation of
    its rebuilt formula.
"""
import random

import numpy as np
import pytest

from test_oracle import pos_semantics

OPS = ["not", "next", "fin", "glob", "and", "or", "until"]
CODE = {"not": 1, "and": 2, "or": 3, "next": 4, "fin": 5, "glob": 6, "until": 7}
M64 = (1 << 64) - 1


def to_words(x, K):
    return np.array([(x >> (64 * k)) & M64 for k in range(K)], dtype=np.uint64)


def from_words(w):
    return sum(int(v) << (64 * k) for k, v in enumerate(w))


def expected(op, a, b, n):
    if op == "and":
        return a & b
    if op == "or":
        return a | b
    return pos_semantics(op, a, b, n)


def rand_trace(rng, n):
    """Random bits with long runs of ones, so that until and glob carry."""
    x = rng.getrandbits(n)
    if rng.random() < 0.6:
        lo = rng.randrange(n)
        hi = rng.randrange(lo, n + 1)
        x |= ((1 << hi) - 1) ^ ((1 << lo) - 1)
    if rng.random() < 0.2:
        x = ((1 << n) - 1) ^ (1 << rng.randrange(n))
    if rng.random() < 0.1:
        x = 1 << rng.randrange(n)
    return x & ((1 << n) - 1)


LENS = [1, 5, 63, 64, 65, 100, 127, 128, 129, 150, 191, 192, 193, 230, 255]


@pytest.mark.parametrize("n", LENS)
def test_trace_ops_vs_positions(oracle, n):
    rng = random.Random(n)
    for K in range(max(1, (n + 63) // 64), 5):
        for _ in range(60):
            a, b = rand_trace(rng, n), rand_trace(rng, n)
            for op in OPS:
                got = from_words(oracle.trace_op(CODE[op], to_words(a, K), to_words(b, K), K, n))
                assert got == expected(op, a, b, n), (op, n, K, hex(a), hex(b))


def test_trace_ops_k1_is_single_word(oracle):
    rng = random.Random(3)
    for _ in range(300):
        n = rng.randint(1, 64)
        a, b = rand_trace(rng, n), rand_trace(rng, n)
        w = lambda op, y=0: int(oracle.trace_op(CODE[op], to_words(a, 1), to_words(y, 1), 1, n)[0])
        assert w("not") == oracle.op_not(a, n)
        assert w("next") == oracle.op_next(a)
        assert w("fin") == oracle.op_fin(a)
        assert w("glob") == oracle.op_glob(a, n)
        assert w("until", b) == oracle.op_until(a, b)


# --- planted slips (CPU copies of the multi-word operators) ---------------
def _fin1(x):
    for s in (1, 2, 4, 8, 16, 32):
        x |= x >> s
    return x


def _until1(l, r):
    run, y = l, r
    for s in (1, 2, 4, 8, 16, 32):
        y |= run & (y >> s)
        run &= run >> s
    return y


def _wmask(n, k):
    lo = 64 * k
    return 0 if n <= lo else (M64 if n - lo >= 64 else (1 << (n - lo)) - 1)


def mw_copy(op, a, b, n, K, slip=None):
    x, y = to_words(a, K).tolist(), to_words(b, K).tolist()
    out, carry = [0] * K, 0
    if op == "next":
        for k in range(K):
            hi = 0 if slip == "next" or k + 1 >= K else (x[k + 1] << 63) & M64
            out[k] = (x[k] >> 1) | hi
    elif op == "fin":
        for k in reversed(range(K)):
            out[k] = _fin1(x[k]) | (_wmask(n, k) if carry and slip != "fin" else 0)
            carry |= x[k] != 0
    elif op == "glob":
        carry = 1
        for k in reversed(range(K)):
            m = _wmask(n, k)
            out[k] = x[k] & ~_fin1(~x[k] & m & M64) & M64 if (carry or slip == "glob") else 0
            carry = carry and x[k] == m
    else:
        for k in reversed(range(K)):
            u = _until1(x[k], y[k])
            if carry and slip != "until":
                u |= ~_fin1(~x[k] & M64) & M64
            out[k] = u
            carry = u & 1
    return from_words(out)


def test_carry_slips_fail_the_pin():
    rng = random.Random(9)
    for slip in ("next", "fin", "glob", "until"):
        bad = 0
        for _ in range(300):
            n = rng.choice([65, 100, 128, 129, 200, 255])
            K = (n + 63) // 64
            a, b = rand_trace(rng, n), rand_trace(rng, n)
            assert mw_copy(slip, a, b, n, K) == expected(slip, a, b, n)   # the copy itself is right
            bad += mw_copy(slip, a, b, n, K, slip) != expected(slip, a, b, n)
        assert bad > 0, slip


# --- enumeration ----------------------------------------------------------
def long_instance(oracle, seed, T, A, lo, hi, K=None):
    rng = random.Random(seed)
    lens = [rng.randint(lo, hi) for _ in range(T)]
    K = K or (max(lens) + 63) // 64
    P = rng.randint(1, T - 1)
    atoms = np.zeros((A, T * K), dtype=np.uint64)
    for a in range(A):
        for t, n in enumerate(lens):
            atoms[a, t * K:(t + 1) * K] = to_words(rand_trace(rng, n), K)
    return oracle.Instance(P, T - P, np.array(lens, dtype=np.uint8), atoms, K=K)


def pad(oracle, inst, K):
    A, T = inst.atoms.shape[0], inst.W
    atoms = np.zeros((A, T, K), dtype=np.uint64)
    atoms[:, :, :inst.K] = inst.atoms.reshape(A, T, inst.K)
    return oracle.Instance(inst.num_pos, inst.num_neg, inst.lens, atoms.reshape(A, T * K), inst.names, K=K)


@pytest.mark.parametrize("seed", range(4))
def test_padding_word_changes_nothing(oracle, seed):
    one = long_instance(oracle, seed, T=30, A=3, lo=5, hi=64, K=1)
    two = pad(oracle, one, 2)
    r1 = oracle.enumerate_(one, max_cost=6, keep_store=True)
    r2 = oracle.enumerate_(two, max_cost=6, keep_store=True)
    for k in ("status", "cost", "all_ltls", "found_id", "n_stored", "formula"):
        assert r1[k] == r2[k], k
    w2 = r2["words"].reshape(r2["n_stored"], one.W, 2)
    assert np.array_equal(w2[:, :, 0], r1["words"])
    assert not w2[:, :, 1].any()
    assert np.array_equal(r1["left"], r2["left"]) and np.array_equal(r1["op"], r2["op"])


def tiny_long(oracle):
    out = []
    for s in range(10):
        rng = random.Random(100 + s)
        T = rng.randint(2, 5)
        out.append(long_instance(oracle, 100 + s, T=T, A=rng.randint(1, 2), lo=40, hi=200))
    return out


@pytest.mark.parametrize("neg_type", [1, 2])
def test_long_min_cost_vs_bruteforce(oracle, neg_type):
    found = 0
    for inst in tiny_long(oracle):
        r = oracle.enumerate_(inst, max_cost=5, neg_type=neg_type)
        bf = oracle.bruteforce_min_cost(inst, max_cost=5, neg_type=neg_type, max_formulas=3_000_000)
        assert bf != -2
        assert (r["cost"] if r["status"] == 1 else -1) == bf
        if r["status"] == 1:
            found += 1
            assert oracle.mismatches(inst, oracle.eval_formula(inst, r["formula"])) == 0
    assert found >= 3


def formula_of(res, names, k, memo):
    if k in memo:
        return memo[k]
    op, l, r = int(res["op"][k]), int(res["left"][k]), int(res["right"][k])
    if op == 0:
        f = names[k]
    elif op in (1, 4, 5, 6):
        f = "~XFG"[[1, 4, 5, 6].index(op)] + "(" + formula_of(res, names, l, memo) + ")"
    else:
        mid = {2: ")&(", 3: ")|(", 7: ")U("}[op]
        f = "(" + formula_of(res, names, l, memo) + mid + formula_of(res, names, r, memo) + ")"
    memo[k] = f
    return f


@pytest.mark.parametrize("K,lo,hi", [(2, 60, 128), (3, 100, 192), (4, 1, 255)])
def test_long_store_vs_textbook(oracle, K, lo, hi):
    inst = long_instance(oracle, 7 * K, T=12, A=3, lo=lo, hi=hi)
    r = oracle.enumerate_(inst, max_cost=5, keep_store=True)
    n = r["n_stored"]
    assert n > 100
    words = r["words"]
    assert len({w.tobytes() for w in words}) == n                 # exact dedup: all distinct
    memo = {}
    rng = random.Random(K)
    for k in sorted(rng.sample(range(n), min(n, 250))):
        f = formula_of(r, inst.names, k, memo)
        assert np.array_equal(oracle.eval_formula(inst, f), words[k]), f


def test_long_parallel_equals_sequential(oracle):
    inst = long_instance(oracle, 77, T=40, A=3, lo=50, hi=250)
    a = oracle.enumerate_(inst, max_cost=6, keep_store=True)
    b = oracle.enumerate_(inst, max_cost=6, keep_store=True, threads=4)
    for k in ("status", "cost", "all_ltls", "found_id", "n_stored", "formula"):
        assert a[k] == b[k], k
    assert np.array_equal(a["words"], b["words"])


def test_long_rejects_relaxed_keys(oracle):
    inst = long_instance(oracle, 5, T=6, A=2, lo=70, hi=120)
    assert oracle.enumerate_(inst, max_cost=4, dedup_mode=3)["status"] < 0
    assert oracle.enumerate_(inst, max_cost=4)["status"] >= 0
