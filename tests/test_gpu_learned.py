"""Learned operand order (any minimal formula: ranked tests with early exit,
binary blocks of the next level and complements tested before a level is
generated).

Log rows: every row of the enumeration log (features and label of one tile
pair) is recomputed on the CPU from the stored sets, with the textbook
recursion for U (holds at k iff r holds at k, or l holds at k and U holds at
k + 1; false past the end of the trace).

Learned vs plain: same status and cost on every instance, with an untrained
and a trained scorer; the returned formula separates the traces (CPU
evaluator of the oracle) and its cost (oracle) is the returned cost."""
import numpy as np
import pytest

from test_gpu_parity import to_oracle

pytestmark = pytest.mark.gpu

AND, OR, UNTIL = 2, 3, 7


def _b(cs, k):
    return ((cs >> np.uint64(k)) & np.uint64(1)).astype(bool)


def until0(l, r, lens):
    u = np.zeros(np.broadcast_shapes(l.shape, r.shape), dtype=bool)
    for k in range(int(lens.max()) - 1, -1, -1):
        u = (k < lens) & (_b(r, k) | (_b(l, k) & u))
    return u


def tile_stats(cs, P):
    pos, neg = cs[:, :P], cs[:, P:]
    N = cs.shape[1] - P
    p0 = _b(pos, 0).sum(1) / max(P, 1)
    n0 = _b(neg, 0).sum(1) / max(N, 1)
    pz = (pos != 0).sum(1) / max(P, 1)
    nz = (neg != 0).sum(1) / max(N, 1)
    return [p0.max(), p0.mean(), n0.min(), n0.mean(), pz.mean(), nz.mean()]


def expected_rows(L, inst, final_cost):
    """Rows for per_block = 1: tile pair (0, 0) of every binary block of the
    check passes of levels 2 .. final_cost (unit costs), in level order."""
    s, e = L.levels()
    P, W = inst.num_pos, inst.W
    rows = []
    for c in range(2, final_cost + 1):
        for op in (AND, OR, UNTIL):
            for i in range(1, (c - 1) // 2 + 1):
                j = c - 1 - i
                nL, nR = e[i] - s[i], e[j] - s[j]
                if nL <= 0 or nR <= 0:
                    continue
                l = L.store(int(s[i]), int(min(64, nL)), W)[0]
                r = L.store(int(s[j]), int(min(64, nR)), W)[0]
                x = np.zeros(24)
                x[:6], x[6:12] = tile_stats(l, P), tile_stats(r, P)
                x[12 + op - 1] = 1
                x[19], x[20], x[21], x[22], x[23] = c / 32, i / c, j / c, P / W, 1
                lv, rv = l[:, None, :], r[None, :, :]
                if op == UNTIL:
                    vals = [until0(lv, rv, inst.lens), until0(rv, lv, inst.lens)]
                elif op == AND:
                    vals = [_b(lv, 0) & _b(rv, 0)]
                else:
                    vals = [_b(lv, 0) | _b(rv, 0)]
                best = min(int(((~v[..., :P]).sum(-1) + v[..., P:].sum(-1)).min()) for v in vals)
                rows.append((x, 1 - best / W))
    return rows


LOG_CASES = [("cfg1", 1), ("cfg1", 2), ("cfg1", 3), ("custom", (20, 13, 3, 9, 16)),
             ("custom", (33, 40, 3, 17, 32)), ("custom", (25, 30, 3, 40, 64))]


def make(ltl, case):
    kind, arg = case
    if kind == "cfg1":
        return ltl.generate(1, seed=arg)
    P, N, A, lo, hi = arg
    return ltl.generate_custom(P + N, P, N, A, lo, hi, rate=0.5)


@pytest.mark.parametrize("case", LOG_CASES, ids=lambda c: f"{c[0]}-{c[1]}")
def test_log_rows_match_cpu(ltl, case):
    inst = make(ltl, case)
    L = ltl.Learner(budget=1 << 30, max_cost=5)
    try:
        L.set_log(100000, per_block=1)
        r = L.learn(inst)
        rows = L.get_log()
        final = r["cost"] if r["status"] == 1 else 5
        want = expected_rows(L, inst, final)
    finally:
        L.close()
    assert len(rows) == len(want) > 0
    for k, (x, lab) in enumerate(want):
        assert np.allclose(rows[k, :24], x, atol=1e-5), (k, rows[k, :24], x)
        assert abs(rows[k, 24] - lab) <= 1e-6, (k, rows[k, 24], lab)


@pytest.fixture(scope="module")
def scorers(ltl):
    untrained = ltl.Scorer(seed=5)
    rows = ltl.collect_logs(config1_seeds=range(1, 9), config2_max_cost=6, config3_max_cost=5)
    trained, loss, n = ltl.train_scorer(rows, seed=2, epochs=10)
    assert n > 100 and np.isfinite(loss)
    yield {"untrained": untrained, "trained": trained}
    untrained.close()
    trained.close()


SAME_CASES = ([("cfg1", s, {}) for s in range(1, 7)] +
              [("custom", (20, 13, 3, 9, 16), {}), ("custom", (33, 40, 3, 17, 32), {}),
               ("custom", (25, 30, 3, 40, 64), {}),
               ("cfg1", 7, dict(capacity=64)), ("cfg1", 8, dict(cost=(1, 2, 1, 3, 1, 2, 1, 2))),
               ("cfg1", 9, dict(neg_type=2)), ("cfg3", 1, dict(max_cost=5)), ("cfg5", 1, dict(max_cost=6)),
               ("cfg5", 2, dict(max_cost=6))])


def make_same(ltl, case):
    kind, arg, opts = case
    if kind == "cfg1":
        return ltl.generate(1, seed=arg), dict(max_cost=8, **opts)
    if kind == "cfg3":
        inst = ltl.generate(3, seed=arg)
        return inst, dict(hamming_k=inst.hamming_k, **opts)
    if kind == "cfg5":
        return ltl.generate(5, seed=1, index=arg), dict(opts)
    P, N, A, lo, hi = arg
    return ltl.generate_custom(P * 7 + N, P, N, A, lo, hi, rate=0.5), dict(max_cost=6, **opts)


@pytest.mark.parametrize("case", SAME_CASES, ids=lambda c: f"{c[0]}-{c[1]}-{'_'.join(c[2])}")
def test_learned_same_cost(ltl, oracle, learner, scorers, case):
    inst, opts = make_same(ltl, case)
    plain = learner.learn(inst, **opts)
    O = oracle
    oi = to_oracle(O, inst)
    cost_vec = opts.get("cost", (1,) * 8)
    k = opts.get("hamming_k", 0)
    try:
        for name, sc in scorers.items():
            learner.set_scorer(sc)
            r = learner.learn(inst, order=1, **opts)
            if plain["on_the_fly"]:
                # After an overflow the plain order only combines the stored
                # sets; the learned order can stop before the overflowing
                # level and return a cheaper formula, never a dearer one.
                assert r["status"] >= plain["status"], name
                assert r["status"] == 0 or plain["status"] == 0 or r["cost"] <= plain["cost"], name
            else:
                assert (r["status"], r["cost"]) == (plain["status"], plain["cost"]), name
            if r["status"] == 1:
                cs = O.eval_formula(oi, r["formula"])
                assert O.mismatches(oi, cs) <= k, (name, r["formula"])
                assert O.formula_cost(r["formula"], cost_vec) == r["cost"], (name, r["formula"])
            else:
                assert r["formula"] == "not_found"
    finally:
        learner.set_scorer(None)


def test_learned_needs_scorer(ltl, learner):
    with pytest.raises(ltl.LtlError):
        learner.learn(ltl.generate(1, seed=1), order=1)
