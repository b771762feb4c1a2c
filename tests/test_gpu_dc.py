"""Native divide-and-conquer driver (ltl_learn_dc, the reference's DetSplit /
RandSplit, code/dc.py:188-321) against its CPU restatement oracle/dc.py over
the oracle enumerator: the same formula string and the same number of window
calls; the formula separates every trace (oracle evaluator) and its cost is
the returned cost."""
import numpy as np
import pytest

import dc as DC

from test_gpu_parity import to_oracle

pytestmark = pytest.mark.gpu

CASES = [(64, 64, 4, 6, 12, 32, 4), (100, 100, 3, 4, 8, 24, 4), (200, 150, 3, 5, 12, 40, 4),
         (80, 90, 2, 4, 10, 20, 4), (30, 34, 3, 20, 40, 16, 4)]


@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c)))
def test_dc_matches_restatement(ltl, oracle, learner, algo, case):
    P, N, A, lo, hi, window, mc = case
    inst = ltl.generate_custom(3 * P + N, P, N, A, lo, hi, 0.5, planted_size=9)
    oi = to_oracle(oracle, inst)
    want, calls = DC.learn_dc(oi, window, algo=algo, seed=7, max_cost=mc)
    got = learner.learn_dc(inst, window=window, algo=algo, seed=7, max_cost=mc)
    assert got["formula"] == want
    assert got["calls"] == calls
    cs = oracle.eval_formula(oi, got["formula"])
    assert oracle.mismatches(oi, cs) == 0
    assert oracle.formula_cost(got["formula"]) == got["cost"]


@pytest.mark.parametrize("algo", [0, 1])
def test_dc_unsplittable_is_an_error(ltl, oracle, learner, algo):
    P, N, A, lo, hi, window, mc = (100, 100, 3, 4, 8, 24, 3)
    inst = ltl.generate_custom(3 * P + N, P, N, A, lo, hi, 0.5, planted_size=9)
    with pytest.raises(DC.Limit):
        DC.learn_dc(to_oracle(oracle, inst), window, algo=algo, seed=7, max_cost=mc)
    with pytest.raises(ltl.LtlError, match="-2"):
        learner.learn_dc(inst, window=window, algo=algo, seed=7, max_cost=mc)


def test_dc_bad_arguments(ltl, learner):
    inst = ltl.generate(1, seed=1)
    with pytest.raises(ltl.LtlError):
        learner.learn_dc(inst, window=1)
    with pytest.raises(ltl.LtlError):
        learner.learn_dc(inst, window=8, algo=5)


def test_dc_hand_traced(ltl, learner):
    # the DetSplit trace written out in tests/test_oracle_dc.py
    from test_oracle_dc import hand_traced_instance
    o = hand_traced_instance()
    inst = ltl.Instance(o.num_pos, o.num_neg, o.lens, o.atoms, o.names)
    got = learner.learn_dc(inst, window=4, algo=0, max_cost=6)
    assert (got["formula"], got["calls"]) == ("((a)&(~(c)))|(b)", 3)


def test_dc_batch_equals_one_by_one(ltl, oracle, learner):
    """ltl_learn_dc_batch: every problem's result is what ltl_learn_dc gives
    it alone (formula, cost, calls, allLTLs) and what the CPU restatement
    gives, with the windows of all problems learned together."""
    insts, want = [], []
    for k, case in enumerate(CASES):
        P, N, A, lo, hi, window, mc = case
        insts.append(ltl.generate_custom(3 * P + N + k, P, N, A, lo, hi, 0.5, planted_size=9))
    n_ok = 0
    for algo in (0, 1):
        got = learner.learn_dc_batch(insts, window=24, algo=algo, seed=7, max_cost=4)
        for inst, a in zip(insts, got):
            oi = to_oracle(oracle, inst)
            try:
                b = learner.learn_dc(inst, window=24, algo=algo, seed=7, max_cost=4)
            except ltl.LtlError as e:
                # (a window that cannot be split further: the same error,
                # the other problems unaffected)
                assert str(e).endswith(str(a["status"])) and a["status"] < 0
                with pytest.raises(DC.Limit):
                    DC.learn_dc(oi, 24, algo=algo, seed=7, max_cost=4)
                continue
            assert (a["formula"], a["cost"], a["calls"], a["all_ltls"]) == \
                   (b["formula"], b["cost"], b["calls"], b["all_ltls"])
            f, calls = DC.learn_dc(oi, 24, algo=algo, seed=7, max_cost=4)
            assert (a["formula"], a["calls"]) == (f, calls)
            n_ok += 1
    assert n_ok >= 4


def test_dc_batch_reports_bad_problems(ltl, learner):
    good = ltl.generate_custom(5, 20, 20, 3, 4, 8, 0.5, planted_size=5)
    bad = ltl.Instance(1, 1, [3, 3], [[1, 2]] * 70)    # 70 atoms: beyond the limit
    got = learner.learn_dc_batch([good, bad], window=16, max_cost=4)
    assert got[1]["status"] == -1
    assert got[0] == {**learner.learn_dc(good, window=16, max_cost=4), "seconds": got[0]["seconds"]}
    assert learner.learn_dc_batch([], window=16) == []


def _rand_formula(rng, size, names):
    from test_host import random_formula
    return random_formula(rng, size, names)


@pytest.mark.parametrize("K", [1, 2, 4])
def test_device_match_equals_host_evaluator(ltl, oracle, learner, K):
    """ltl_match (the reference's LTLmatch on the device) against the
    product's host evaluator and the oracle's textbook evaluator: random
    formulas up to 40 nodes, traces of 1 .. 64K positions."""
    import random
    from test_long_traces import random_instance, to_oracle as to_oracle_k
    rng = random.Random(K)
    for trial in range(6):
        W = rng.choice([1, 63, 64, 65, 300, 1000])
        A = rng.randint(1, 5)
        P = rng.randint(0, W)
        inst = random_instance(ltl, 100 * K + trial, P, W - P, A, 1, 64 * K)
        fs = [_rand_formula(rng, rng.randint(1, 40), inst.names) for _ in range(25)]
        got = learner.match(inst, fs)
        oi = to_oracle_k(oracle, inst)
        for f, row in zip(fs, got):
            host = ltl.eval_formula(inst, f).reshape(W, K)[:, 0] & 1
            assert np.array_equal(row, host.astype(np.uint8)), f
            ref = oracle.eval_formula(oi, f).reshape(W, K)[:, 0] & 1
            assert np.array_equal(row, ref.astype(np.uint8)), f


def test_device_match_deep_and_bad_formulas(ltl, learner):
    inst = ltl.generate(1, seed=3)
    # a left-deep chain of 600 operators evaluates in a stack of two entries
    f = "a"
    for k in range(600):
        f = f"({f}){'&|U'[k % 3]}(b)" if k % 2 else f"X({f})"
    got = learner.match(inst, [f, "G(a)"])
    host = ltl.eval_formula(inst, f) & 1
    assert np.array_equal(got[0], host.astype(np.uint8))
    for bad in ("", "(a)&(b", "X(q)"):
        with pytest.raises(ltl.LtlError, match="-1"):
            learner.match(inst, [bad])
    assert learner.match(inst, []).shape == (0, inst.W)
