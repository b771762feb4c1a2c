"""The benchmark's own workload, validated: config 2 (500 + 500 traces, 5
atoms, planted G(a -> F b) & ~(c U d)) at max cost 14 with a dedup table of
2^28 slots, exactly as bench.py runs it, must find a formula of minimal cost
11 that separates every trace under the oracle's textbook evaluator and
whose cost under the oracle's cost function is 11.  Minimality in part: the
oracle's exhaustive enumeration up to cost 8 (about 4 s on 8 host threads)
finds nothing; every stored set of the device's levels up to cost 6 equals
the oracle's (tests/test_gpu_parity.py::test_config2_levels)."""
import pytest

pytestmark = pytest.mark.gpu


def test_config2_headline_run(ltl, oracle):
    inst = ltl.generate(2, seed=1)
    L = ltl.Learner(budget=0, hash_slots_log2=28, max_cost=14, neg_type=1)
    try:
        r = L.learn(inst)
        r2 = L.learn(inst)
    finally:
        L.close()
    assert r["status"] == 1 and r["cost"] == 11
    assert (r2["formula"], r2["cost"], r2["n_stored"]) == (r["formula"], r["cost"], r["n_stored"])
    oi = oracle.Instance(inst.num_pos, inst.num_neg, inst.lens, inst.atoms, inst.names)
    cs = oracle.eval_formula(oi, r["formula"])
    assert cs is not None and oracle.mismatches(oi, cs) == 0
    assert oracle.formula_cost(r["formula"]) == 11
    assert not r["on_the_fly"]
    assert oracle.enumerate_(oi, max_cost=8, threads=8)["status"] == 0
