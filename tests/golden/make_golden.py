"""Write tests/golden/expected.json: oracle results for the reference trace
fixtures in tests/golden/traces (copied from the reference's benchmarks/),
with the brute-force minimal cost where the search is small enough.

    python tests/golden/make_golden.py
"""
import glob
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

MAX_COST = 7


def main():
    cases = []
    for path in sorted(glob.glob(os.path.join(HERE, "traces", "*.trace"))):
        inst = O.read_trace_file(path)
        for neg_type in (1, 2):
            r = O.enumerate_(inst, max_cost=MAX_COST, neg_type=neg_type)
            bf = None
            if inst.W <= 20:
                # found: the minimal cost; not found: -1 (nothing up to MAX_COST)
                bf = O.bruteforce_min_cost(inst, max_cost=r["cost"], neg_type=neg_type,
                                           max_formulas=20_000_000)
                bf = None if bf == -2 else bf
            cases.append(dict(file=os.path.basename(path), neg_type=neg_type, max_cost=MAX_COST,
                              W=inst.W, status=r["status"], cost=r["cost"], all_ltls=r["all_ltls"],
                              found_id=r["found_id"], formula=r["formula"], n_stored=r["n_stored"],
                              bruteforce=bf))
            print(cases[-1])
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
