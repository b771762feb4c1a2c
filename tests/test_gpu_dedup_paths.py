"""The three ways a generated chunk is deduplicated and compacted give the
same store: one workgroup for chunks of at most 64 tiles (4096 candidates),
one decoupled look-back launch for up to 1024 tiles, decide + device scan +
finish + compact above.  The pool budget sets the chunk size (per-candidate
temporaries within 1/32 of it), so a 4 MiB pool takes the first path only, a
64 MiB pool the first two, and a large pool whichever a block's size calls
for; each run equals the oracle set for set (results, counts, every stored
set with its parents)."""
import pytest

from test_gpu_parity import check_parity

pytestmark = pytest.mark.gpu

# (seed, P, N, atoms, len_min, len_max, max_cost); the oracle's stores: about
# 4.8 K, 0.7 K and 1.1 K sets for SMALL, 63 K for MID, 812 K for LARGE (no
# formula up to their max cost; LARGE's level 8 has a block of 5 x 17186
# pairs, 1350 tiles in one chunk of the large pool)
SMALL = [(31, 20, 13, 3, 9, 16, 6), (18, 7, 9, 2, 2, 8, 7), (43, 30, 30, 3, 17, 32, 5)]
MID = [(75, 64, 64, 4, 5, 40, 7)]
LARGE = [(79, 64, 64, 5, 5, 40, 8)]


@pytest.mark.parametrize("budget_mib,shapes,paths",
                         [(4, SMALL, "small"), (64, SMALL + MID, "small+lookback"),
                          (2048, SMALL + MID + LARGE, "small+lookback+four")],
                         ids=["pool4M", "pool64M", "pool2G"])
def test_dedup_paths_equal_oracle(ltl, oracle, budget_mib, shapes, paths):
    L = ltl.Learner(budget=budget_mib << 20)
    try:
        count = {"small": 0, "lookback": 0, "four": 0}
        for seed, P, N, A, lo, hi, mc in shapes:
            inst = ltl.generate_custom(seed, P, N, A, lo, hi, 0.5, planted_size=9)
            gpu, _ = check_parity(ltl, oracle, L, inst, max_cost=mc, profile=1)
            assert not gpu["on_the_fly"]
            ks = L.kernel_stats()
            for k in count:
                count[k] += ks["dedup_path_" + k]["launches"]
            if count["four"] == 0:
                assert ks["scan_finish"]["launches"] == 0
        # every path the pool allows is taken, and no other
        want = set(paths.split("+"))
        assert {k for k, v in count.items() if v > 0} == want, count
    finally:
        L.close()
