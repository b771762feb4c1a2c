"""Every launch shape of the packed generation kernels gives the oracle's
store.  The launchers pick, per chunk, how many waves fold one task's words
(1, or 2 or 4 with the partial sums added in LDS, for chunks of few tasks),
and a tested level with equal costs of &, | and U generates each split's &
and | blocks (and its U block, when the level's candidates fit the chunk
temporaries) in one pass, deciding each block later at its canonical
position; with unit unary costs the emits of a level also fingerprint the
next level's ~, X, F and G of every stored set, so that the next level's
unary blocks need only their table inserts, and those emits leave the words
unwritten (the check of until recomputes the words it reads from the
parents; the words are written before the next level is generated in full,
or by the first store read).  Test sizes take only some of these, so each is
forced here through the environment (LTL_GEN_SPLIT = 1 / 2 / 4, LTL_GEN_FUSE
= 0 / 1 / 2, LTL_GEN_NX = 0 / 2 -- never / on every level, not only the large
ones --, LTL_GEN_LZW = 0 -- the fingerprinting emits write the words; read at
every launch, level or call) and compared set for set with the oracle:
results, counts, level boundaries, every stored set with its parents.  The
instances cover generation with and without the target check (precheck 0 /
1), the fused ~/X/F/G pass, non-unit costs, deferred levels whose X / F / G
successors are tested in generation (the successor-test kernel variants),
and traces of 2-4 words each (the multi-word fused binary pass)."""
import os

import pytest

from test_gpu_parity import check_parity
from test_long_traces import check_parity as long_parity
from test_long_traces import random_instance

pytestmark = pytest.mark.gpu


MODES = [("1", "2", "2", "1"), ("2", "2", "2", "1"), ("4", "2", "2", "1"), ("1", "1", "2", "1"), ("4", "1", "2", "1"),
         ("1", "0", "2", "1"), ("4", "0", "2", "1"), ("1", "2", "2", "0"), ("4", "0", "2", "0"), ("1", "2", "0", "1"),
         ("4", "0", "0", "1")]


@pytest.fixture(params=MODES, ids=lambda m: f"split{m[0]}-fuse{m[1]}-nx{m[2]}-lzw{m[3]}")
def gen_mode(request):
    keys = ("LTL_GEN_SPLIT", "LTL_GEN_FUSE", "LTL_GEN_NX", "LTL_GEN_LZW")
    old = {k: os.environ.get(k) for k in keys}
    for k, v in zip(keys, request.param):
        os.environ[k] = v
    yield request.param
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


# (seed, P, N, atoms, len_min, len_max, max_cost): no formula up to max_cost,
# so every level is stored; atoms 2-5 put left operands of 2-5 rows (tasks of
# one stripe shared by a workgroup with the next stripe's) against levels of
# hundreds to thousands of sets
NOT_FOUND = [(31, 20, 13, 3, 9, 16, 5), (43, 30, 30, 3, 17, 32, 5), (75, 64, 64, 4, 5, 40, 4),
             (18, 7, 9, 2, 2, 8, 5), (79, 64, 64, 5, 5, 40, 5)]


@pytest.mark.parametrize("precheck", [0, 1])
@pytest.mark.parametrize("shape", NOT_FOUND, ids=lambda s: "x".join(map(str, s)))
def test_modes_full_store(ltl, oracle, learner, gen_mode, shape, precheck):
    seed, P, N, A, lo, hi, mc = shape
    inst = ltl.generate_custom(seed, P, N, A, lo, hi, 0.5, planted_size=9)
    gpu, _ = check_parity(ltl, oracle, learner, inst, max_cost=mc, precheck=precheck)
    assert gpu["status"] == 0
    # Level mc - 1 is fingerprinted by its emits (forced); with the precheck
    # level mc is deferred, so mc - 1's words stay unwritten until the store
    # is read; without it level mc is generated in full and they are written
    # before that.
    s, e = learner.levels()
    want = e[mc - 1] - s[mc - 1] if gen_mode[2] == "2" and gen_mode[3] == "1" and precheck else 0
    assert gpu["words_unwritten"] == want


def test_modes_found_and_costs(ltl, oracle, learner, gen_mode):
    for seed in range(8):
        inst = ltl.generate_custom(500 + seed, 12 + seed % 9, 10 + seed % 7, 3, 3, 10 + seed % 12, 0.5,
                                   planted_size=5 + seed % 4)
        check_parity(ltl, oracle, learner, inst, max_cost=9)
    inst = ltl.generate_custom(77, 20, 20, 3, 4, 20, 0.5, planted_size=9)
    check_parity(ltl, oracle, learner, inst, max_cost=8, cost=(1, 2, 1, 2, 1, 1, 2, 2))


# the first binary hit of the last level is a U candidate: the deferred
# level's generation tests the X / F / G successors (successor kernels)
U_HIT = [("(F(a))U(b)", 1584, 17, 17, 5, 16, 0), ("X((a)U(b))", 1388, 15, 15, 3, 10, 2)]


@pytest.mark.parametrize("case", U_HIT, ids=lambda c: f"{c[0]}-{c[1]}")
def test_modes_successors_in_generation(ltl, oracle, learner, gen_mode, case):
    planted, seed, P, N, lo, hi, k = case
    inst = ltl.generate_custom(seed, P, N, 3, lo, hi, 0.5, planted=planted)
    gpu, _ = check_parity(ltl, oracle, learner, inst, max_cost=9, hamming_k=k)
    assert gpu["status"] == 1 and gpu["deferred_succ"] & 14


# K words per trace (2-4): the fused binary pass of multi-word levels
# (k_gen2fm) against one pass per block (LTL_GEN_FUSE = 0) and the & / | pass
# without U (1); the other knobs concern one-word levels only
LONG = [(20, 20, 3, 65, 190, 8), (50, 50, 4, 64, 250, 7), (9, 11, 2, 60, 128, 7)]


@pytest.mark.parametrize("shape", LONG, ids=lambda s: "x".join(map(str, s)))
def test_modes_long_traces(ltl, oracle, learner, gen_mode, shape):
    P, N, A, lo, hi, mc = shape
    inst = random_instance(ltl, 3 + P, P, N, A, lo, hi)
    assert inst.K > 1
    long_parity(oracle, learner, inst, max_cost=mc)


# Stores that overflow into check-only levels: a level whose candidates could
# overflow the store keeps its words (the check-only levels read them), the
# levels before it may not
# (levels of 5, 20, 81, 521 and 2936 sets up to cost 5: the store overflows in
# level 4, early in level 5 and late in level 5)
@pytest.mark.parametrize("capacity", [400, 1500, 3000])
def test_modes_overflow(ltl, oracle, learner, gen_mode, capacity):
    inst = ltl.generate_custom(79, 64, 64, 5, 5, 40, 0.5, planted_size=9)
    gpu, _ = check_parity(ltl, oracle, learner, inst, max_cost=5, capacity=capacity)
    assert gpu["on_the_fly"]
