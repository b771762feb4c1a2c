"""CPU restatement of the reference's divide-and-conquer drivers (test
infrastructure only): DetSplit (code/dc.py:188-249) and RandSplit
(code/dc.py:252-321), with the oracle enumerator (oracle.enumerate_) as the
window learner and the oracle's textbook evaluator as LTLmatch
(code/dc.py:185).  It mirrors the device driver's documented choices so the
two can be compared formula for formula: traces are handled by index (the
reference's set differences act on trace strings, which selects the same
traces up to duplicates), windows are sampled with a seeded SplitMix64
stream reseeded on every RandSplit call (the reference reseeds Python's
random module with 0, dc.py:255), and an empty window, a split that cannot
shrink or a recursion deeper than 64 raises instead of recursing forever."""
import numpy as np

import oracle as O

MASK = (1 << 64) - 1


class Limit(Exception):
    pass


class Driver:
    def __init__(self, inst, window, seed=0, **opts):
        self.inst, self.window, self.seed, self.opts = inst, window, seed, opts
        self.rng = 0
        self.cache = {}
        self.calls = 0

    def match(self, f):
        if f not in self.cache:
            cs = O.eval_formula(self.inst, f)
            if cs is None:
                raise ValueError(f"cannot evaluate {f!r}")
            self.cache[f] = (cs & np.uint64(1)).astype(bool)
        return self.cache[f]

    def all_of(self, f, s):
        m = self.match(f)
        return all(m[i] for i in s)

    def none_of(self, f, s):
        m = self.match(f)
        return not any(m[i] for i in s)

    def filt(self, f, s, want):
        m = self.match(f)
        return [i for i in s if bool(m[i]) == want]

    @staticmethod
    def minus(a, b):
        bs = set(b)
        return [i for i in a if i not in bs]

    def solve(self, pos, neg):
        idx = list(pos) + list(neg)
        sub = O.Instance(len(pos), len(neg), self.inst.lens[idx], self.inst.atoms[:, idx], self.inst.names)
        self.calls += 1
        r = O.enumerate_(sub, **self.opts)
        return r["formula"] if r["status"] == 1 else ""

    def sample(self, s, k):
        a = list(s)
        for i in range(k):
            self.rng = (self.rng + 0x9E3779B97F4A7C15) & MASK
            z = self.rng
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
            z ^= z >> 31
            j = i + z % (len(a) - i)
            a[i], a[j] = a[j], a[i]
        return a[:k]

    @staticmethod
    def both(a, b):
        return "(" + a + ")&(" + b + ")"

    def det(self, pos, neg, depth=0):
        if depth > 64 or not (pos or neg):
            raise Limit()
        if len(pos) + len(neg) <= self.window:
            f = self.solve(pos, neg)
            if f:
                return f
            if len(pos) + len(neg) <= 1:
                raise Limit()
        p1, p2 = pos[:len(pos) // 2], pos[len(pos) // 2:]
        n1, n2 = neg[:len(neg) // 2], neg[len(neg) // 2:]
        phi11 = self.det(p1, n1, depth + 1)
        if self.all_of(phi11, p2) and self.none_of(phi11, n2):
            return phi11
        if self.none_of(phi11, n2):
            left = phi11
        else:
            n2a = self.filt(phi11, n2, True)
            phi12 = self.det(p1, n2a, depth + 1)
            left = phi12 if self.none_of(phi12, self.minus(neg, n2a)) else self.both(phi11, phi12)
            if self.all_of(left, p2):
                return left
        p2m = self.filt(left, p2, False)
        phi21 = self.det(p2m, n1, depth + 1)
        rest = self.minus(pos, p2m)
        if self.all_of(phi21, rest) and self.none_of(phi21, n2):
            return phi21
        if self.none_of(phi21, n2):
            right = phi21
        else:
            n2b = self.filt(phi21, n2, True)
            phi22 = self.det(p2m, n2b, depth + 1)
            right = phi22 if self.none_of(phi22, self.minus(neg, n2b)) else self.both(phi21, phi22)
            if self.all_of(right, rest):
                return right
        return "(" + left + ")|(" + right + ")"

    def rnd(self, pos, neg, depth=0):
        if depth > 64 or not (pos or neg):
            raise Limit()
        self.rng = self.seed
        win = self.window
        while True:
            if win < 1:
                raise Limit()
            if len(pos) + len(neg) <= win:
                p1, n1 = pos, neg
            elif len(pos) <= win // 2:
                p1, n1 = pos, self.sample(neg, win - len(pos))
            elif len(neg) <= win // 2:
                n1 = neg
                p1 = self.sample(pos, win - len(neg))
            else:
                p1 = self.sample(pos, win // 2)
                n1 = self.sample(neg, win - win // 2)
            phi11 = self.solve(p1, n1)
            if phi11:
                break
            win //= 2
        p2, n2 = self.filt(phi11, pos, False), self.filt(phi11, neg, True)
        p1, n1 = self.minus(pos, p2), self.minus(neg, n2)
        if not p2 and not n2:
            return phi11
        if not n2:
            left = phi11
        else:
            phi12 = self.rnd(p1, n2, depth + 1)
            left = phi12 if self.none_of(phi12, n1) else self.both(phi11, phi12)
            if self.all_of(left, p2):
                return left
        phi21 = self.rnd(p2, n1, depth + 1)
        if self.all_of(phi21, p1) and self.none_of(phi21, n2):
            return phi21
        if self.none_of(phi21, n2):
            right = phi21
        else:
            phi22 = self.rnd(p2, n2, depth + 1)
            right = phi22 if self.none_of(phi22, n1) else self.both(phi21, phi22)
            if self.all_of(right, p1):
                return right
        return "(" + left + ")|(" + right + ")"


def learn_dc(inst, window, algo=0, seed=0, **opts):
    """(formula, number of window learner calls)."""
    d = Driver(inst, window, seed, **opts)
    pos = list(range(inst.num_pos))
    neg = list(range(inst.num_pos, inst.num_pos + inst.num_neg))
    f = d.det(pos, neg) if algo == 0 else d.rnd(pos, neg)
    return f, d.calls
