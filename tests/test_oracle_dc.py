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
