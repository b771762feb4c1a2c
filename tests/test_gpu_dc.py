"""Native divide-and-conquer driver (ltl_learn_dc, the reference's DetSplit /
RandSplit, code/dc.py:188-321) against its CPU restatement oracle/dc.py over
the oracle enumerator: the same formula string and the same number of window
calls; the formula separates every trace (oracle evaluator) and its cost is
the returned cost."""
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
