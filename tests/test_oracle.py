"""Pins of the CPU oracle (no GPU).

The reference ships no tests and no expected outputs, so the oracle is pinned
by what the mathematics fixes:
  - every bit-parallel operator against textbook per-position finite-trace
    semantics, exhaustively over all words of length <= 8 (and sampled at
    length 64);
  - identities: F f = T U f, G f = ~F~f, X as a shift, zero padding;
  - the enumerator's minimal cost against brute force over all syntactic
    formulas (no dedup) on tiny instances, for NegType 1/2, non-unit costs
    and Hamming acceptance;
  - the returned formula re-evaluated by the textbook evaluator;
  - the relaxed-key mode-3 mixer against the published SplitMix64 vector;
  - the committed golden results of the reference trace fixtures;
  - the OpenMP enumerator against the sequential one, set for set.
"""
import glob
import itertools
import json
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN


def pos_semantics(op, a, b, n):
    """Per-position definition of each operator on a trace of length n."""
    bit = lambda x, k: (x >> k) & 1 if k < n else 0
    out = 0
    for k in range(n):
        if op == "not":
            v = 1 - bit(a, k)
        elif op == "next":
            v = bit(a, k + 1)
        elif op == "fin":
            v = int(any(bit(a, j) for j in range(k, n)))
        elif op == "glob":
            v = int(all(bit(a, j) for j in range(k, n)))
        else:  # until
            v = int(any(bit(b, j) and all(bit(a, m) for m in range(k, j)) for j in range(k, n)))
        out |= v << k
    return out


@pytest.mark.parametrize("n", range(1, 9))
def test_unary_ops_exhaustive(oracle, n):
    for a in range(1 << n):
        assert oracle.op_not(a, n) == pos_semantics("not", a, 0, n)
        assert oracle.op_next(a) == pos_semantics("next", a, 0, n)
        assert oracle.op_fin(a) == pos_semantics("fin", a, 0, n)
        assert oracle.op_glob(a, n) == pos_semantics("glob", a, 0, n)


@pytest.mark.parametrize("n", range(1, 7))
def test_until_exhaustive(oracle, n):
    for a in range(1 << n):
        for b in range(1 << n):
            assert oracle.op_until(a, b) == pos_semantics("until", a, b, n), (a, b, n)


def test_ops_sampled_len64(oracle):
    rng = random.Random(5)
    for _ in range(200):
        n = rng.choice([63, 64, 40, 17])
        m = (1 << n) - 1
        a = rng.getrandbits(64) & m
        b = rng.getrandbits(64) & m
        # sparse runs make until non-trivial
        if rng.random() < 0.5:
            a |= m ^ (1 << rng.randrange(n))
        assert oracle.op_until(a, b) == pos_semantics("until", a, b, n)
        assert oracle.op_glob(a, n) == pos_semantics("glob", a, 0, n)
        assert oracle.op_fin(b) == pos_semantics("fin", b, 0, n)


@pytest.mark.parametrize("n", [1, 5, 8, 33, 64])
def test_identities(oracle, n):
    rng = random.Random(n)
    m = (1 << n) - 1
    for _ in range(300):
        a = rng.getrandbits(64) & m
        assert oracle.op_fin(a) == oracle.op_until(m, a)                       # F f = T U f
        assert oracle.op_glob(a, n) == oracle.op_not(oracle.op_fin(oracle.op_not(a, n)), n)
        assert oracle.op_next(a) == a >> 1
        for v in (oracle.op_not(a, n), oracle.op_fin(a), oracle.op_glob(a, n), oracle.op_next(a)):
            assert v & ~m == 0                                                 # padding stays zero


def tiny_instances(oracle):
    rng = random.Random(11)
    out = []
    for t in range(14):
        P, N = rng.randint(1, 3), rng.randint(1, 3)
        A = rng.randint(1, 2)
        lens = np.array([rng.randint(1, 4) for _ in range(P + N)], dtype=np.uint8)
        atoms = np.array([[rng.getrandbits(int(l)) for l in lens] for _ in range(A)], dtype=np.uint64)
        out.append(oracle.Instance(P, N, lens, atoms))
    return out


@pytest.mark.parametrize("neg_type", [1, 2])
def test_min_cost_vs_bruteforce(oracle, neg_type):
    for inst in tiny_instances(oracle):
        r = oracle.enumerate_(inst, max_cost=6, neg_type=neg_type)
        bf = oracle.bruteforce_min_cost(inst, max_cost=6, neg_type=neg_type, max_formulas=3_000_000)
        assert bf != -2
        if r["status"] == 1:
            assert r["cost"] == bf
            cs = oracle.eval_formula(inst, r["formula"])
            assert oracle.mismatches(inst, cs) == 0
            assert oracle.formula_cost(r["formula"]) == r["cost"]
        else:
            assert bf == -1


@pytest.mark.parametrize("cost", [(1, 2, 1, 1, 2, 3, 1, 2), (2, 1, 2, 3, 1, 1, 2, 1)])
def test_min_cost_nonunit_vs_bruteforce(oracle, cost):
    for inst in tiny_instances(oracle)[:8]:
        r = oracle.enumerate_(inst, max_cost=8, cost=cost)
        bf = oracle.bruteforce_min_cost(inst, max_cost=8, cost=cost, max_formulas=3_000_000)
        assert bf != -2
        assert (r["cost"] if r["status"] == 1 else -1) == bf
        if r["status"] == 1:
            assert oracle.formula_cost(r["formula"], cost) == r["cost"]


def test_hamming_vs_bruteforce(oracle):
    for inst in tiny_instances(oracle)[:10]:
        for k in (1, 2):
            r = oracle.enumerate_(inst, max_cost=5, hamming_k=k)
            bf = oracle.bruteforce_min_cost(inst, max_cost=5, hamming_k=k, max_formulas=3_000_000)
            assert (r["cost"] if r["status"] == 1 else -1) == bf
            if r["status"] == 1:
                assert oracle.mismatches(inst, oracle.eval_formula(inst, r["formula"])) <= k


def _until_missing_run_update(a, b):
    y = b
    for s in (1, 2, 4, 8, 16, 32):
        y |= a & (y >> s)          # plausible slip: the run mask never narrows
    return y & 0xFFFFFFFFFFFFFFFF


def _glob_without_mask(a, n):
    c = ~a & 0xFFFFFFFFFFFFFFFF    # plausible slip: complement not masked to the length
    for s in (1, 2, 4, 8, 16, 32):
        c |= c >> s
    return a & ~c & 0xFFFFFFFFFFFFFFFF


def _next_wrong_direction(a, n):
    return (a << 1) & ((1 << n) - 1)


def test_pins_reject_planted_mistakes():
    """Each plausible operator slip, planted in a CPU copy, fails the
    per-position pin used above (so the pin is not a retyped formula)."""
    n = 6
    assert any(_until_missing_run_update(a, b) != pos_semantics("until", a, b, n)
               for a in range(1 << n) for b in range(1 << n))
    assert any(_glob_without_mask(a, n) != pos_semantics("glob", a, 0, n) for a in range(1 << n))
    assert any(_next_wrong_direction(a, n) != pos_semantics("next", a, 0, n) for a in range(1 << n))


def test_bruteforce_reads_positions(oracle):
    """Reversing every trace changes minimal costs on some tiny instance:
    the search depends on the temporal order, not only on the labels."""
    diffs = 0
    for inst in tiny_instances(oracle):
        a = oracle.bruteforce_min_cost(inst, max_cost=5)
        rev = oracle.Instance(inst.num_pos, inst.num_neg, inst.lens,
                              np.array([[int(format(int(w), f"0{int(l)}b")[::-1], 2)
                                         for w, l in zip(row, inst.lens)] for row in inst.atoms],
                                       dtype=np.uint64))
        diffs += a != oracle.bruteforce_min_cost(rev, max_cost=5)
    assert diffs > 0


def test_splitmix_mode3_vector(oracle):
    """Mode 3 mixes each trace word with the SplitMix64 finalizer; its first
    published output for seed 0 is fin(0x9e3779b97f4a7c15) = 0xe220a8397b1dcdaf.
    With two traces, h gets fin(word0) and l gets fin(word1)."""
    lens = np.array([64, 64] + [64] * 2, dtype=np.uint8)
    inst = oracle.Instance(2, 2, lens, np.zeros((1, 4), dtype=np.uint64))
    cs = np.array([0x9e3779b97f4a7c15, 0, 0, 0], dtype=np.uint64)
    h, l = oracle.relaxed_key(inst, cs, 3)
    assert h == 0xe220a8397b1dcdaf ^ 0  # fin(0) = 0
    assert l == 0


def test_exact_key_is_injective_below_126(oracle):
    rng = random.Random(3)
    lens = np.array([5, 7, 3, 9], dtype=np.uint8)
    inst = oracle.Instance(2, 2, lens, np.zeros((1, 4), dtype=np.uint64))
    seen = {}
    for _ in range(3000):
        cs = np.array([rng.getrandbits(int(n)) for n in lens], dtype=np.uint64)
        for mode in (1, 2, 3):
            assert oracle.relaxed_key(inst, cs, mode) == oracle.relaxed_key(inst, cs, 1)
        k = oracle.relaxed_key(inst, cs, 1)
        assert seen.setdefault(k, tuple(cs)) == tuple(cs)


def test_relaxed_mode_dedup_is_coarser(oracle):
    """Above 126 positions the reference keys are lossy: the stored set can
    only shrink relative to exact dedup (same order, fewer survivors)."""
    rng = np.random.default_rng(2)
    W, A = 24, 3
    lens = rng.integers(8, 16, W).astype(np.uint8)
    atoms = np.array([[int(rng.integers(0, 1 << int(n))) for n in lens] for _ in range(A)], dtype=np.uint64)
    inst = oracle.Instance(12, 12, lens, atoms)
    ex = oracle.enumerate_(inst, max_cost=5)
    for mode in (1, 2, 3):
        r = oracle.enumerate_(inst, max_cost=5, dedup_mode=mode)
        assert r["n_stored"] <= ex["n_stored"]
    assert oracle.enumerate_(inst, max_cost=5, dedup_mode=1)["n_stored"] < ex["n_stored"]


def test_parallel_equals_sequential(oracle):
    rng = np.random.default_rng(7)
    W, A = 40, 3
    lens = rng.integers(5, 20, W).astype(np.uint8)
    atoms = np.array([[int(rng.integers(0, 1 << int(n))) for n in lens] for _ in range(A)], dtype=np.uint64)
    inst = oracle.Instance(20, 20, lens, atoms)
    for mc in (5, 6):
        a = oracle.enumerate_(inst, max_cost=mc, keep_store=True)
        b = oracle.enumerate_(inst, max_cost=mc, keep_store=True, threads=4)
        for k in ("status", "cost", "all_ltls", "found_id", "formula", "n_stored"):
            assert a[k] == b[k]
        assert np.array_equal(a["words"], b["words"])
        assert np.array_equal(a["left"], b["left"]) and np.array_equal(a["right"], b["right"])


def test_overflow_semantics(oracle):
    """A capacity below the level sizes switches to check-only mode and, by
    the last-round rule, stops one level after the overflow."""
    rng = np.random.default_rng(9)
    W = 16
    lens = rng.integers(4, 12, W).astype(np.uint8)
    atoms = np.array([[int(rng.integers(0, 1 << int(n))) for n in lens] for _ in range(3)], dtype=np.uint64)
    inst = oracle.Instance(8, 8, lens, atoms)
    free = oracle.enumerate_(inst, max_cost=7)
    cap = oracle.enumerate_(inst, max_cost=7, capacity=100)
    assert cap["on_the_fly"] == 1 and cap["n_stored"] == 100
    if free["status"] == 1 and cap["status"] == 1:
        assert cap["cost"] >= free["cost"]


def golden_cases():
    with open(os.path.join(GOLDEN, "expected.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: f"{c['file']}-neg{c['neg_type']}")
def test_golden_fixtures(oracle, case):
    inst = oracle.read_trace_file(os.path.join(GOLDEN, "traces", case["file"]))
    r = oracle.enumerate_(inst, max_cost=case["max_cost"], neg_type=case["neg_type"])
    for k in ("status", "cost", "all_ltls", "found_id", "formula", "n_stored"):
        assert r[k] == case[k], k
    if case.get("bruteforce") is not None:
        assert (r["cost"] if r["status"] == 1 else -1) == case["bruteforce"]


def test_trace_reader_known_values(oracle):
    inst = oracle.read_trace_file(os.path.join(GOLDEN, "traces", "Ab0.trace"))
    assert (inst.num_pos, inst.num_neg, inst.names) == (10, 10, "pq")
    # first positive line "0,0;0,1;0,0;0,1;0,0": q at positions 1 and 3
    assert int(inst.atoms[1, 0]) == 0b01010 and int(inst.atoms[0, 0]) == 0
    # first negative line "1,0;0,1;0,0;0,1;0,1": p at 0, q at 1, 3, 4
    assert int(inst.atoms[0, 10]) == 0b00001 and int(inst.atoms[1, 10]) == 0b11010
    assert list(inst.lens) == [5] * 20


# ---- relaxed keys above 126 positions: hand-derived vectors ---------------
# Key bit j is bit j of l (j < 63) or bit j - 63 of h (code/ltli6463.cu:57-141).
def _key(oracle, lens, trace, pos, mode):
    """Key of the set holding exactly position `pos` of trace `trace`."""
    lens = np.array(lens, dtype=np.uint8)
    inst = oracle.Instance(len(lens), 0, lens, np.zeros((1, len(lens)), dtype=np.uint64))
    cs = np.zeros(len(lens), dtype=np.uint64)
    cs[trace] = np.uint64(1 << pos)
    return oracle.relaxed_key(inst, cs, mode)


L = lambda j: (0, 1 << j)          # key bit j < 63
H = lambda j: (1 << (j - 63), 0)   # key bit j >= 63


def test_relaxed_mode1_vectors(oracle):
    # 5 traces of 52 positions: lenSum 260, stride 260 / 126 = 2; every
    # trace gives 26 key bits (positions 0, 2, ..., 50) until 126 are taken,
    # so trace t position p (p even) is key bit 26 t + p / 2, and trace 4
    # stops after position 42 (key bit 125).
    lens = [52] * 5
    assert _key(oracle, lens, 0, 50, 1) == L(25)
    assert _key(oracle, lens, 1, 2, 1) == L(27)
    assert _key(oracle, lens, 1, 3, 1) == (0, 0)       # odd positions are not sampled
    assert _key(oracle, lens, 2, 0, 1) == L(52)
    assert _key(oracle, lens, 3, 0, 1) == H(78)
    assert _key(oracle, lens, 4, 42, 1) == H(125)
    assert _key(oracle, lens, 4, 44, 1) == (0, 0)      # the 127th sample is dropped
    # 3 traces of 60 positions: lenSum 180, stride 1: trace 2 contributes
    # positions 0..5 only (key bits 120..125).
    lens = [60] * 3
    assert _key(oracle, lens, 1, 59, 1) == H(119)
    assert _key(oracle, lens, 2, 5, 1) == H(125)
    assert _key(oracle, lens, 2, 6, 1) == (0, 0)


def test_relaxed_mode2_vectors(oracle):
    # 3 traces of 60 positions: each takes (126 * 60 + 180) / 180 = 43 key
    # bits (positions 0..42), the third only the 40 that remain.
    lens = [60] * 3
    assert _key(oracle, lens, 0, 42, 2) == L(42)
    assert _key(oracle, lens, 1, 0, 2) == L(43)
    assert _key(oracle, lens, 1, 42, 2) == H(85)
    assert _key(oracle, lens, 1, 43, 2) == (0, 0)
    assert _key(oracle, lens, 2, 39, 2) == H(125)
    assert _key(oracle, lens, 2, 40, 2) == (0, 0)
    # lengths 10, 63, 63 (lenSum 136): takes (1260 + 136) / 136 = 10 and
    # (7938 + 136) / 136 = 59, the last trace the 57 bits that remain.
    lens = [10, 63, 63]
    assert _key(oracle, lens, 0, 9, 2) == L(9)
    assert _key(oracle, lens, 1, 58, 2) == H(68)
    assert _key(oracle, lens, 1, 59, 2) == (0, 0)
    assert _key(oracle, lens, 2, 56, 2) == H(125)
    assert _key(oracle, lens, 2, 57, 2) == (0, 0)


def _key_copy(cs, lens, mode, slip=None):
    """CPU copy of the relaxed keys of modes 1 and 2, with a planted slip."""
    lsum = sum(lens)
    h = l = 0
    j = 0
    for i, n in enumerate(lens):
        if mode == 1:
            stride = lsum // 126 + (1 if slip == "stride" else 0)
            ks = range(0, n, stride)
        else:
            take = (126 * n + (0 if slip == "take" else lsum)) // lsum
            ks = range(take)
        for k in ks:
            if j >= 126:
                break
            if (int(cs[i]) >> k) & 1 if k < 64 else 0:
                if j < 63:
                    l |= 1 << j
                else:
                    h |= 1 << (j - 63)
            j += 1
    return h, l


def test_relaxed_key_slips_fail_the_vectors(oracle):
    """A stride off by one (mode 1) or the take without the + lenSum rounding
    (mode 2), planted in a CPU copy, fails the hand-derived vectors; the copy
    without a slip equals the oracle on random sets."""
    rng = np.random.default_rng(4)
    for lens, mode in (([52] * 5, 1), ([60] * 3, 1), ([60] * 3, 2), ([10, 63, 63], 2)):
        inst = oracle.Instance(len(lens), 0, np.array(lens, dtype=np.uint8),
                               np.zeros((1, len(lens)), dtype=np.uint64))
        for _ in range(200):
            cs = np.array([int(rng.integers(0, 1 << n, dtype=np.uint64)) for n in lens], dtype=np.uint64)
            assert _key_copy(cs, lens, mode) == oracle.relaxed_key(inst, cs, mode)
    one = lambda lens, t, p: np.array([1 << p if i == t else 0 for i in range(len(lens))], dtype=np.uint64)
    assert _key_copy(one([52] * 5, 1, 2), [52] * 5, 1, "stride") != L(27)
    assert _key_copy(one([60] * 3, 1, 42), [60] * 3, 2, "take") != H(85)
    assert _key_copy(one([10, 63, 63], 0, 9), [10, 63, 63], 2, "take") != L(9)


def envelope_cases():
    with open(os.path.join(GOLDEN, "envelope.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", envelope_cases(), ids=lambda c: c["file"][-40:])
def test_envelope_fixtures(oracle, case):
    """The 69 reference trace files within the reference's single-run limits
    (tests/golden/envelope/): the oracle reproduces its committed results,
    exact and with the reference's mode-3 key."""
    inst = oracle.read_trace_file(os.path.join(GOLDEN, "envelope", case["file"]))
    assert inst.W == case["W"]
    r = oracle.enumerate_(inst, max_cost=case["max_cost"], neg_type=2, threads=8)
    for k in ("status", "cost", "all_ltls", "found_id", "formula", "n_stored"):
        assert r[k] == case[k], k
    r3 = oracle.enumerate_(inst, max_cost=case["ref3"]["max_cost"], neg_type=2, dedup_mode=3)
    for k in ("status", "cost", "all_ltls", "formula", "n_stored"):
        assert r3[k] == case["ref3"][k], k
    if r["status"] == 1:
        assert oracle.mismatches(inst, oracle.eval_formula(inst, r["formula"])) == 0


def _twin4_oracle(oracle, seed, n=40):
    """The mirrored four-trace instance of the GPU parity tests: t1 is t0 with
    a and b exchanged, t3 is t2 with a and b exchanged (so a set and its a<->b
    mirror have equal mode-3 keys); at position 1 b holds on t0 only."""
    rng = np.random.default_rng(seed)
    a0, b0, a2, b2 = (rng.integers(0, 2, n) for _ in range(4))
    a0[1], b0[1], a0[0], b0[0], a2[1], b2[1] = 0, 1, 1, 1, 0, 0
    w = lambda v: int(sum(int(x) << k for k, x in enumerate(v)))
    atoms = np.array([[w(a0), w(b0), w(a2), w(b2)], [w(b0), w(a0), w(b2), w(a2)]], dtype=np.uint64)
    return oracle.Instance(1, 3, np.full(4, n, dtype=np.uint8), atoms, "ab")


def test_relaxed_acceptance_tests_key_new_candidates_only(oracle):
    """The reference tests a candidate only when its relaxed key is new
    (code/ltli6463.cu:231-243).  X(b) separates this instance (the brute-force
    minimum, cost 2), but its mode-3 key equals that of X(a), an earlier
    candidate of the same level: in mode 3 it is never tested, and the answer
    is a separating formula of higher cost."""
    inst = _twin4_oracle(oracle, 3)
    assert oracle.bruteforce_min_cost(inst) == 2
    exact = oracle.enumerate_(inst, max_cost=6)
    assert (exact["cost"], exact["formula"]) == (2, "X(b)")
    xa, xb = oracle.eval_formula(inst, "X(a)"), oracle.eval_formula(inst, "X(b)")
    assert oracle.relaxed_key(inst, xa, 3) == oracle.relaxed_key(inst, xb, 3)
    assert oracle.mismatches(inst, xa) > 0 and oracle.mismatches(inst, xb) == 0
    r3 = oracle.enumerate_(inst, max_cost=6, dedup_mode=3)
    assert (r3["cost"], r3["formula"]) == (4, "X((a)U(b))")
    assert oracle.mismatches(inst, oracle.eval_formula(inst, r3["formula"])) == 0
    assert oracle.formula_cost(r3["formula"]) == 4
