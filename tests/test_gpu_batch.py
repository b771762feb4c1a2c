"""The persistent batched launch (ltl_learn_batch) returns, for every
instance, exactly what ltl_learn returns, whatever the tile, word-block and
deferral settings; a sample is also checked against the oracle."""
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("status", "cost", "all_ltls", "found_id", "formula")


def config4(ltl, n, seed=1):
    return [ltl.generate(4, seed=seed, index=i) for i in range(n)]


def single(learner, insts, **opts):
    return [learner.learn(i, max_cost=i.max_cost if opts.get("max_cost") is None else opts["max_cost"],
                          **{k: v for k, v in opts.items() if k != "max_cost"}) for i in insts]


def test_batch_equals_single_config4(ltl, learner):
    insts = config4(ltl, 48)
    # config 4 instances carry their own max cost (planted size); use the max
    mc = max(i.max_cost for i in insts)
    ref = [learner.learn(i, max_cost=mc) for i in insts]
    learner.set_batch_params(tile=512, word_block=4, defer_limit=1 << 22)
    got = learner.learn_batch(insts, max_cost=mc)
    for k, (a, b) in enumerate(zip(got, ref)):
        assert {x: a[x] for x in KEYS} == {x: b[x] for x in KEYS}, (k, a, b)
    st = learner.batch_stats()
    assert st["on_device"] > 0 and st["n"] == 48


@pytest.mark.parametrize("tile", [256, 512, 1024])
@pytest.mark.parametrize("word_block", [4, 8])
def test_batch_params_do_not_change_results(ltl, learner, tile, word_block):
    insts = config4(ltl, 16, seed=2)
    ref = [learner.learn(i, max_cost=8) for i in insts]
    learner.set_batch_params(tile=tile, word_block=word_block)
    got = learner.learn_batch(insts, max_cost=8)
    for a, b in zip(got, ref):
        assert {x: a[x] for x in KEYS} == {x: b[x] for x in KEYS}


def test_batch_deferral(ltl, learner):
    insts = config4(ltl, 12, seed=3)
    ref = [learner.learn(i, max_cost=7) for i in insts]
    learner.set_batch_params(defer_limit=50)   # nearly every level is deferred
    got = learner.learn_batch(insts, max_cost=7)
    assert learner.batch_stats()["deferred"] > 0
    learner.set_batch_params()
    for a, b in zip(got, ref):
        assert {x: a[x] for x in KEYS} == {x: b[x] for x in KEYS}


def test_batch_mixed_widths_and_options(ltl, learner):
    shapes = [(7, 9, 2, 2, 8), (20, 13, 3, 9, 16), (33, 40, 3, 17, 32), (25, 31, 3, 33, 64), (30, 30, 3, 5, 20)]
    insts = [ltl.generate_custom(100 + k, P, N, A, lo, hi, 0.5, planted_size=6)
             for k, (P, N, A, lo, hi) in enumerate(shapes * 2)]
    for opts in (dict(max_cost=7), dict(max_cost=7, neg_type=2), dict(max_cost=7, hamming_k=2),
                 dict(max_cost=12, cost=(1, 2, 1, 1, 2, 3, 1, 2))):
        ref = [learner.learn(i, **opts) for i in insts]
        got = learner.learn_batch(insts, **opts)
        for a, b in zip(got, ref):
            assert {x: a[x] for x in KEYS} == {x: b[x] for x in KEYS}, opts


def test_batch_vs_oracle(ltl, oracle, learner):
    insts = config4(ltl, 10, seed=5)
    got = learner.learn_batch(insts, max_cost=6)
    for inst, g in zip(insts, got):
        oi = oracle.Instance(inst.num_pos, inst.num_neg, inst.lens, inst.atoms, inst.names)
        r = oracle.enumerate_(oi, max_cost=6)
        assert {x: g[x] for x in KEYS} == {x: r[x] for x in KEYS}


def test_batch_whole_device_calls_in_parallel(ltl, learner):
    """Instances left to the whole-device path run several at once, each on a
    context of its own: nearly every instance here is left (a tiny deferral
    limit), long-trace instances among them; results equal one ltl_learn each,
    on two calls in a row (the helper contexts are kept and reused), and the
    calling context's own single calls are unaffected afterwards."""
    from test_long_traces import random_instance
    insts = config4(ltl, 24, seed=6)
    insts += [random_instance(ltl, 500 + k, 10 + k, 9 + k, 3, 60, 150) for k in range(4)]
    ref = [learner.learn(i, max_cost=7) for i in insts]
    learner.set_batch_params(defer_limit=50)
    try:
        for _ in range(2):
            got = learner.learn_batch(insts, max_cost=7)
            st = learner.batch_stats()
            assert st["deferred"] >= 20 and st["deferred"] + st["on_device"] + st["host"] == len(insts)
            for k, (a, b) in enumerate(zip(got, ref)):
                assert {x: a[x] for x in KEYS} == {x: b[x] for x in KEYS}, k
    finally:
        learner.set_batch_params()
    again = learner.learn(insts[0], max_cost=7)
    assert {x: again[x] for x in KEYS} == {x: ref[0][x] for x in KEYS}
