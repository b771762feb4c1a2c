"""The per-trace summaries the device keeps beside every stored set
(position 0 holds, position 1 holds, some position holds, every position
holds) against the same four facts recomputed from the stored words.

The check passes decide acceptance from these summaries, and the emit
kernels derive some of them from the operands' summaries (~: all four; &
and |: three of four), so they are pinned set by set here, for one-word
instances of several shapes and for K words per trace."""
import numpy as np
import pytest

from test_long_traces import random_instance

pytestmark = pytest.mark.gpu


def expected(words, lens, K):
    """(n, 4, W) booleans from the stored words (trace order)."""
    n, W = words.shape[0], len(lens)
    out = np.zeros((n, 4, W), dtype=bool)
    w = words.reshape(n, W, K)
    for i, L in enumerate(int(x) for x in lens):
        masks = [((1 << min(64, L - 64 * k)) - 1) if L > 64 * k else 0 for k in range(K)]
        x0 = w[:, i, 0]
        out[:, 0, i] = (x0 & np.uint64(1)) != 0
        out[:, 1, i] = (x0 & np.uint64(2)) != 0
        out[:, 2, i] = (w[:, i, :] != 0).any(axis=1)
        out[:, 3, i] = np.all(np.stack([w[:, i, k] == np.uint64(masks[k]) for k in range(K)], axis=1), axis=1)
    return out


def unpack(sums, W):
    n = sums.shape[0]
    bits = np.zeros((n, 4, W), dtype=bool)
    for i in range(W):
        bits[:, :, i] = ((sums[:, :, i // 64] >> np.uint64(i % 64)) & np.uint64(1)) != 0
    return bits


def check(learner, inst, **opts):
    r = learner.learn(inst, **opts)
    n = int(learner.levels()[1].max())
    assert n > inst.atoms.shape[0]
    words = learner.store(0, n, inst.W * inst.K)[0]
    got = unpack(learner.summaries(0, n, inst.W), inst.W)
    want = expected(words, inst.lens, inst.K)
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} summary bits differ, first (set, vector, trace) {bad[:3].tolist()}"
    return r


# (P, N, atoms, len_min, len_max, max_cost): lengths 1..64, one and two
# traces of length 1, full words, many traces per word
SHAPES = [(7, 9, 2, 1, 8, 6), (20, 13, 3, 9, 16, 5), (33, 40, 3, 17, 32, 5), (25, 31, 3, 33, 64, 5),
          (64, 64, 4, 1, 40, 4), (100, 90, 3, 60, 64, 4)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("neg_type", [1, 2])
def test_summaries_one_word(ltl, learner, shape, neg_type):
    P, N, A, lo, hi, mc = shape
    inst = ltl.generate_custom(31 + P, P, N, A, lo, hi, 0.5, planted_size=9)
    check(learner, inst, max_cost=mc, neg_type=neg_type)


def test_summaries_config2_levels(ltl, learner):
    check(learner, ltl.generate(2, seed=1), max_cost=6)


@pytest.mark.parametrize("hi", [128, 200, 255])
def test_summaries_multiword(ltl, learner, hi):
    inst = random_instance(ltl, hi, 20, 20, 3, 1, hi)
    check(learner, inst, max_cost=6)
