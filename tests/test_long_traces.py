"""Traces longer than 64 positions: K words per trace through the product.

CPU tests: the host evaluator (the product's LTLmatch) and the trace reader
against the oracle's textbook evaluator and a plain per-position parse.
GPU tests (marked): the enumeration through the C ABI against the oracle,
bit-exact as in the one-word parity tests (status, cost, allLTLs, canonical
id, formula, levels, every stored set with its parents); a one-word instance
padded to K words gives the one-word results; the learned order, the
batched call, the divide-and-conquer driver and the CLI take long traces.
The reference stops at 63 positions (code/ltli6463.cu:1078), so there are no
reference outputs to compare with; the oracle's multi-word operators are
pinned in test_oracle_long.py.
"""
import os
import random
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR

M64 = (1 << 64) - 1


def rand_trace(rng, n, rate=0.5):
    x = 0
    for k in range(n):
        if rng.random() < rate:
            x |= 1 << k
    if rng.random() < 0.3:   # a long run of ones: temporal operators across words
        lo = rng.randrange(n)
        hi = rng.randrange(lo, n + 1)
        x |= ((1 << hi) - 1) ^ ((1 << lo) - 1)
    return x


def planted_instance(ltl, O, seed, P, N, A, lo, hi, planted, rate=0.5):
    """Random traces labelled by a planted formula (oracle evaluator)."""
    rng = random.Random(seed)
    K = (hi + 63) // 64
    pos, neg = [], []
    for _ in range(200 * (P + N)):
        if len(pos) >= P and len(neg) >= N:
            break
        n = rng.randint(lo, hi)
        tr = [rand_trace(rng, n, rate) for _ in range(A)]
        atoms = np.array([[(x >> (64 * k)) & M64 for k in range(K)] for x in tr], dtype=np.uint64)
        one = O.Instance(1, 0, np.array([n], dtype=np.uint8), atoms.reshape(A, K), K=K)
        holds = int(O.eval_formula(one, planted)[0]) & 1
        (pos if holds else neg).append((n, atoms))
    assert len(pos) >= P and len(neg) >= N, "planted formula too rare"
    traces = pos[:P] + neg[:N]
    lens = np.array([t[0] for t in traces], dtype=np.uint8)
    atoms = np.stack([t[1] for t in traces], axis=1).reshape(A, (P + N) * K)
    return ltl.Instance(P, N, lens, atoms, K=K)


def random_instance(ltl, seed, P, N, A, lo, hi):
    rng = random.Random(seed)
    K = (hi + 63) // 64
    lens = np.array([rng.randint(lo, hi) for _ in range(P + N)], dtype=np.uint8)
    atoms = np.zeros((A, P + N, K), dtype=np.uint64)
    for a in range(A):
        for i, n in enumerate(lens):
            x = rand_trace(rng, int(n))
            for k in range(K):
                atoms[a, i, k] = (x >> (64 * k)) & M64
    return ltl.Instance(P, N, lens, atoms.reshape(A, (P + N) * K), K=K)


def to_oracle(O, inst):
    return O.Instance(inst.num_pos, inst.num_neg, inst.lens, inst.atoms, inst.names, K=inst.K)


# --- CPU ----------------------------------------------------------------------
FORMULAS = ["a", "X(a)", "F(b)", "G(a)", "~(F(X(a)))", "(a)U(b)", "(b)U((a)&(X(b)))",
            "G((~(a))|(F(b)))", "X(X(X((a)U(G(b)))))", "(F(G(a)))|((~(b))U(X(a)))"]


@pytest.mark.parametrize("hi", [100, 128, 129, 190, 255])
def test_host_evaluator_matches_textbook(ltl, oracle, hi):
    inst = random_instance(ltl, hi, 9, 11, 2, 1, hi)
    for f in FORMULAS:
        got = ltl.eval_formula(inst, f)
        want = oracle.eval_formula(to_oracle(oracle, inst), f)
        assert np.array_equal(got, want), f
        assert ltl.mismatches(inst, got) == oracle.mismatches(to_oracle(oracle, inst), want)


def write_trace_file(path, pos, neg, alphabet):
    def line(tr):
        return ";".join(",".join("1" if c in tok else "0" for c in alphabet) for tok in tr)
    with open(path, "w") as f:
        for t in pos:
            f.write(line(t) + "\n")
        f.write("---\n")
        for t in neg:
            f.write(line(t) + "\n")
        f.write("---\nF,G,X,!,&,|\n---\n" + ",".join(alphabet) + "\n")


def rand_tokens(rng, n, alphabet):
    return ["".join(c for c in alphabet if rng.random() < 0.5) for _ in range(n)]


@pytest.mark.parametrize("hi", [60, 64, 65, 200, 255])
def test_reader_long_traces(ltl, tmp_path, hi):
    rng = random.Random(hi)
    alphabet = "pqr"
    pos = [rand_tokens(rng, rng.randint(1, hi), alphabet) for _ in range(5)] + [rand_tokens(rng, hi, alphabet)]
    neg = [rand_tokens(rng, rng.randint(1, hi), alphabet) for _ in range(4)]
    p = str(tmp_path / "t.trace")
    write_trace_file(p, pos, neg, alphabet)
    inst = ltl.read_trace_file(p)
    assert inst.K == (hi + 63) // 64
    ref = ltl._instance_from_tokens(pos, neg, alphabet)     # per-position encoding
    assert inst.K == ref.K and inst.names == ref.names
    assert np.array_equal(inst.lens, ref.lens) and np.array_equal(inst.atoms, ref.atoms)


def test_reader_rejects_256_positions(ltl, tmp_path):
    p = str(tmp_path / "t.trace")
    write_trace_file(p, [["p"] * 256], [["q"]], "pq")
    with pytest.raises(ltl.LtlError):
        ltl.read_trace_file(p)


# --- GPU ----------------------------------------------------------------------
def check_parity(O, learner, inst, **opts):
    gpu = learner.learn(inst, **opts)
    okw = {k: v for k, v in opts.items() if k in ("cost", "max_cost", "neg_type", "hamming_k", "capacity")}
    ref = O.enumerate_(to_oracle(O, inst), keep_store=True, **okw)
    ctx = f"{inst.num_pos}+{inst.num_neg} K={inst.K} opts={opts}"
    for k in ("status", "cost", "all_ltls", "found_id", "formula", "on_the_fly"):
        assert gpu[k] == ref[k], (ctx, k, gpu[k], ref[k])
    if gpu["status"] == 1 and gpu["found_id"] < inst.atoms.shape[0]:
        return gpu, ref
    s, e = learner.levels()
    c1 = opts.get("cost", (1,) * 8)[0]
    last = gpu["cost"] - (1 if gpu["status"] == 1 else 0)
    n_cmp = 0
    for c in range(c1, last + 1):
        assert (s[c], e[c]) == (ref["level_start"][c], ref["level_end"][c]), (ctx, c)
        n_cmp = max(n_cmp, int(e[c]))
    if n_cmp:
        words, op, left, right = learner.store(0, n_cmp, inst.W * inst.K)
        assert np.array_equal(words, ref["words"][:n_cmp]), ctx
        assert np.array_equal(op, ref["op"][:n_cmp]), ctx
        assert np.array_equal(left, ref["left"][:n_cmp]), ctx
        assert np.array_equal(right, ref["right"][:n_cmp]), ctx
    return gpu, ref


# (P, N, atoms, len_min, len_max, max_cost): K = 2, 3, 4 words per trace
LONG = [(9, 11, 2, 60, 128, 7), (20, 20, 3, 65, 190, 8), (33, 31, 3, 100, 255, 6), (70, 60, 2, 1, 200, 7),
        (100, 100, 4, 64, 130, 6), (50, 50, 4, 64, 250, 7)]


@pytest.mark.gpu
@pytest.mark.parametrize("shape", LONG, ids=lambda s: "x".join(map(str, s)))
def test_gpu_long_full_store(ltl, oracle, learner, shape):
    P, N, A, lo, hi, mc = shape
    inst = random_instance(ltl, 3 + P, P, N, A, lo, hi)
    check_parity(oracle, learner, inst, max_cost=mc)


# Each holds on 35-70% of random traces of 40-255 positions; the first and
# last depend on positions in the top word of a trace.
PLANTED = ["G((~(a))|(F(b)))", "X(X((a)U(b)))", "(a)U((b)&(X(a)))", "F((a)&(G(b)))"]


@pytest.mark.gpu
@pytest.mark.parametrize("planted", PLANTED)
@pytest.mark.parametrize("hi", [128, 255])
def test_gpu_long_planted(ltl, oracle, learner, planted, hi):
    inst = planted_instance(ltl, oracle, len(planted) + hi, 24, 24, 2, 40, hi, planted)
    gpu, _ = check_parity(oracle, learner, inst, max_cost=7)
    assert gpu["status"] == 1
    assert ltl.mismatches(inst, ltl.eval_formula(inst, gpu["formula"])) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("k", [1, 3])
def test_gpu_long_hamming(ltl, oracle, learner, k):
    inst = random_instance(ltl, 40 + k, 30, 30, 3, 70, 180)
    check_parity(oracle, learner, inst, max_cost=6, hamming_k=k)


@pytest.mark.gpu
@pytest.mark.parametrize("capacity", [80, 300])
def test_gpu_long_overflow(ltl, oracle, learner, capacity):
    inst = random_instance(ltl, 50, 16, 16, 3, 70, 150)
    check_parity(oracle, learner, inst, max_cost=7, capacity=capacity)


@pytest.mark.gpu
@pytest.mark.parametrize("neg_type", [1, 2])
def test_gpu_long_nonunit(ltl, oracle, learner, neg_type):
    inst = random_instance(ltl, 60, 12, 12, 2, 64, 140)
    check_parity(oracle, learner, inst, max_cost=10, cost=(1, 2, 1, 1, 2, 3, 1, 2), neg_type=neg_type)


@pytest.mark.gpu
def test_gpu_padding_word_changes_nothing(ltl, learner):
    """A one-word instance run with K = 2 (second word empty) returns what
    the one-word path returns, set for set."""
    one = ltl.generate_custom(17, 40, 40, 3, 10, 64, 0.5, planted_size=7)
    A, W = one.atoms.shape[0], one.W
    two = np.zeros((A, W, 2), dtype=np.uint64)
    two[:, :, 0] = one.atoms
    padded = ltl.Instance(one.num_pos, one.num_neg, one.lens, two.reshape(A, 2 * W), one.names, K=2)
    r1 = learner.learn(one, max_cost=6)
    n1 = int(learner.levels()[1].max())
    w1 = learner.store(0, n1, W)
    r2 = learner.learn(padded, max_cost=6)
    w2 = learner.store(0, n1, 2 * W)
    for k in ("status", "cost", "all_ltls", "found_id", "formula", "n_stored"):
        assert r1[k] == r2[k], k
    words2 = w2[0].reshape(n1, W, 2)
    assert np.array_equal(words2[:, :, 0], w1[0]) and not words2[:, :, 1].any()
    assert np.array_equal(w1[2], w2[2]) and np.array_equal(w1[3], w2[3])


@pytest.mark.gpu
def test_gpu_long_learned_order(ltl, oracle):
    """The learned order on long traces returns a separating formula of the
    plain order's cost."""
    L = ltl.Learner(budget=2 << 30)
    try:
        L.set_scorer(ltl.Scorer(seed=3))
        for planted in PLANTED[:3]:
            inst = planted_instance(ltl, oracle, 900 + len(planted), 30, 30, 2, 50, 200, planted)
            plain = L.learn(inst, max_cost=7, order=ltl.ORDER_PLAIN)
            learned = L.learn(inst, max_cost=7, order=ltl.ORDER_LEARNED)
            assert plain["status"] == learned["status"] == 1
            assert learned["cost"] == plain["cost"]
            assert ltl.mismatches(inst, ltl.eval_formula(inst, learned["formula"])) == 0
    finally:
        L.close()


@pytest.mark.gpu
def test_gpu_batch_with_long_traces(ltl, oracle, learner):
    insts = [random_instance(ltl, 70 + i, 10, 10, 2, 30, 60) for i in range(3)]
    insts += [planted_instance(ltl, oracle, 80 + i, 10, 10, 2, 65, 150, PLANTED[i]) for i in range(3)]
    batch = learner.learn_batch(insts, max_cost=6)
    for inst, b in zip(insts, batch):
        one = learner.learn(inst, max_cost=6)
        for k in ("status", "cost", "all_ltls", "found_id", "formula"):
            assert b[k] == one[k], k
    st = learner.batch_stats()
    assert st["deferred"] + st["host"] >= 3     # long traces go to ltl_learn (or the atom check)


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["det", "rand"])
def test_gpu_dc_long_traces(ltl, oracle, algo):
    inst = planted_instance(ltl, oracle, 5, 40, 40, 2, 70, 200, "(a)U((b)&(X(a)))")
    L = ltl.Learner(budget=2 << 30)
    try:
        r = L.learn_dc(inst, window=16, algo=ltl.DC_DET if algo == "det" else ltl.DC_RAND, seed=1,
                       max_cost=5)
        assert r["status"] == 1
        cs = oracle.eval_formula(to_oracle(oracle, inst), r["formula"])
        assert oracle.mismatches(to_oracle(oracle, inst), cs) == 0
    finally:
        L.close()


@pytest.mark.gpu
def test_gpu_cli_long_traces(ltl, oracle, tmp_path):
    rng = random.Random(4)
    alphabet = "ab"
    pos = [rand_tokens(rng, rng.randint(70, 150), alphabet) for _ in range(6)]
    neg = [rand_tokens(rng, rng.randint(70, 150), alphabet) for _ in range(6)]
    p = str(tmp_path / "long.trace")
    write_trace_file(p, pos, neg, alphabet)
    out = subprocess.run([os.path.join(PKG_DIR, "ltli"), p, "1", "1", "1", "1", "1", "1", "1", "1", "6", "3", "1"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    formula = out.stdout.split("\n")[-2].split()[-1][1:-1]     # how code/dc.py:194 parses it
    inst = ltl.read_trace_file(p)
    ref = oracle.enumerate_(to_oracle(oracle, inst), max_cost=6)
    assert formula == ref["formula"]
