"""Pins of the CPU restatement of the divide-and-conquer drivers
(oracle/dc.py): the combined formula separates every trace (checked with the
oracle's textbook evaluator), a window that holds every trace gives the
enumerator's own answer, and the seeded sampler draws k distinct indices."""
import numpy as np
import pytest

import dc as DC
import oracle as O

CASES = [(20, 20, 3, 4, 8, 16), (40, 30, 3, 5, 10, 16), (64, 64, 4, 6, 12, 32)]


def _inst(ltl, P, N, A, lo, hi):
    g = ltl.generate_custom(3 * P + N, P, N, A, lo, hi, 0.5, planted_size=9)
    return O.Instance(g.num_pos, g.num_neg, g.lens, g.atoms, g.names)


@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c)))
def test_dc_formula_separates(ltl, algo, case):
    *shape, window = case
    inst = _inst(ltl, *shape)
    f, calls = DC.learn_dc(inst, window, algo=algo, seed=7, max_cost=6)
    cs = O.eval_formula(inst, f)
    assert cs is not None and O.mismatches(inst, cs) == 0, f
    assert calls >= 1


def test_dc_whole_window_is_direct(ltl):
    inst = _inst(ltl, 10, 10, 3, 4, 8)
    direct = O.enumerate_(inst, max_cost=8)
    assert direct["status"] == 1
    for algo in (0, 1):
        f, calls = DC.learn_dc(inst, 64, algo=algo, max_cost=8)
        assert (f, calls) == (direct["formula"], 1)


def test_sampler_draws_distinct():
    d = DC.Driver(None, 8, seed=5)
    d.rng = 5
    s = d.sample(list(range(100)), 30)
    assert len(set(s)) == 30 and all(0 <= x < 100 for x in s)


def hand_traced_instance():
    """Eight one-position traces over a, b, c (values at position 0):

        P0 P1: a     P2 P3: b     N0 N2: none     N1: c     N3: a, c

    DetSplit with window 4 (code/dc.py:188-249), traced by hand:
      - 8 traces > 4: p1 = P0 P1, p2 = P2 P3, n1 = N0 N1, n2 = N2 N3;
      - phi11 = learn(P0 P1 | N0 N1) = "a" (the first separating atom);
        it fails P2 (no a) and accepts N3, so n2AndPhi11 = [N3];
      - phi12 = learn(P0 P1 | N3): no atom separates; at cost 2 the
        negations come first, ~(a) fails P0, ~(b) accepts N3, ~(c) separates;
        ~(c) accepts N0, so it cannot replace phi11: left = (a)&(~(c)), which
        fails P2: P2MinusLeft = P2 P3;
      - phi21 = learn(P2 P3 | N0 N1) = "b"; it fails P0, rejects N2 and N3:
        right = b;
      - result (left)|(right) after three window calls."""
    a = [1, 1, 0, 0, 0, 0, 0, 1]
    b = [0, 0, 1, 1, 0, 0, 0, 0]
    c = [0, 0, 0, 0, 0, 1, 0, 1]
    return O.Instance(4, 4, np.ones(8, dtype=np.uint8), np.array([a, b, c], dtype=np.uint64), "abc")


def test_detsplit_hand_traced():
    inst = hand_traced_instance()
    f, calls = DC.learn_dc(inst, 4, algo=0, max_cost=6)
    assert (f, calls) == ("((a)&(~(c)))|(b)", 3)
    assert O.mismatches(inst, O.eval_formula(inst, f)) == 0
