"""ltl_learn_noalloc (the allocation-free learning call, through the C ABI)
against ltl_learn on a context whose device budget is the scratch region's
size: the same result, level boundaries and stored sets, on two calls in a
row, the second with a smaller instance; the oracle checks the small one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(ltl, na, ref, inst, **opts):
    a = na.learn_noalloc(inst, **opts)
    b = ref.learn(inst, **opts)
    for k in ("status", "cost", "all_ltls", "found_id", "formula", "n_stored", "on_the_fly",
              "candidates", "words_deferred", "nodes"):
        assert a[k] == b[k], k
    sa, ea = na.levels()
    sb, eb = ref.levels()
    assert np.array_equal(sa, sb) and np.array_equal(ea, eb)
    n = int(a["n_stored"])
    for first, count in ((0, min(n, 4096)), (max(0, n - 4096), min(n, 4096))):
        x = na.store(first, count, inst.W)
        y = ref.store(first, count, inst.W)
        for u, v in zip(x, y):
            assert np.array_equal(u, v)
    return a


def test_noalloc_equals_learn(ltl, oracle):
    S = ltl.NOALLOC_SCRATCH_BYTES
    na = ltl.Learner(budget=S)
    ref = ltl.Learner(budget=S)   # pool of the scratch region's size: same capacity, table, chunks
    try:
        # no formula up to cost 8: levels of tens of thousands of sets
        big = ltl.generate_custom(17, 200, 200, 3, 10, 30, 0.5, planted_size=13)
        a = _same(ltl, na, ref, big, max_cost=8, hash_slots_log2=0)
        assert a["n_stored"] > 20000
        small = ltl.generate(1, seed=2)
        a = _same(ltl, na, ref, small, max_cost=10)
        oi = oracle.Instance(small.num_pos, small.num_neg, small.lens, small.atoms, small.names)
        r = oracle.enumerate_(oi, max_cost=10)
        for k in ("status", "cost", "all_ltls", "found_id", "formula"):
            assert a[k] == r[k], k
    finally:
        na.close()
        ref.close()


def test_noalloc_limits(ltl):
    na = ltl.Learner(budget=ltl.NOALLOC_SCRATCH_BYTES)
    try:
        # more traces than the scratch region holds a store of: a clear error
        W = 260_000
        rng = np.random.default_rng(1)
        atoms = rng.integers(0, 1 << 63, size=(3, W), dtype=np.uint64)
        huge = ltl.Instance(W // 2, W - W // 2, np.full(W, 64, dtype=np.uint8), atoms)
        with pytest.raises(ltl.LtlError, match=f"{ltl.E_LIMIT}"):
            na.learn_noalloc(huge, max_cost=3)
        small = ltl.generate(1, seed=3)
        with pytest.raises(ltl.LtlError, match=f"{ltl.E_ARG}"):
            na.learn_noalloc(small, max_cost=10, profile=1)
        # the context still works afterwards
        assert na.learn_noalloc(small, max_cost=10)["status"] == 1
    finally:
        na.close()
