"""Per-launch log (ltl_set_launch_log; the reference's "Cost N | (op) |
AllLTLs | StoredLTLs | ToBeChecked" lines, code/ltli6463.cu:1096, 1225-1423).

The generation records of a plain-order call follow the reference's block
order within a level (code/ltli6463.cu:1209-1440: ~, &, |, X, F, G, U; the
binary operators over the level pairs (i, c - 1 - i), i <= c - 1 - i), with
sizes from the stored level boundaries, and their AllLTLs are consecutive
canonical ids from the atoms on; the found id lies in a record of the final
level with the found formula's operator; the CLI prints the lines in the
reference's format and still ends with the line code/dc.py:194 parses."""
import os
import subprocess

import pytest

from conftest import GOLDEN, PKG_DIR

pytestmark = pytest.mark.gpu

ATOM, NOT, AND, OR, NEXT, FIN, GLOB, UNTIL = range(8)


def expected_blocks(s, e, c):
    """(op, candidates) of level c's generation with unit costs, from the
    level boundaries s, e of levels below c."""
    n = lambda lv: int(e[lv] - s[lv]) if 1 <= lv < len(s) else 0
    out = []
    if n(c - 1):
        out.append((NOT, n(c - 1)))
    for op in (AND, OR):
        for i in range(1, c):
            j = c - 1 - i
            if i <= j and n(i) and n(j):
                out.append((op, n(i) * n(j)))
    if n(c - 1):
        out += [(NEXT, n(c - 1)), (FIN, n(c - 1)), (GLOB, n(c - 1))]
    for i in range(1, c):
        j = c - 1 - i
        if i <= j and n(i) and n(j):
            out.append((UNTIL, 2 * n(i) * n(j)))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_launch_log_follows_reference_order(ltl, seed):
    inst = ltl.generate(1, seed=seed)
    L = ltl.Learner(budget=1 << 30, max_cost=9)
    try:
        L.launch_log(True)
        r = L.learn(inst)
        recs = list(L.launches)
        s, e = L.levels()
        L.launch_log(False)
        L.learn(inst)
        assert L.launches == []
    finally:
        L.close()
    A = inst.atoms.shape[0]
    assert recs[0] == dict(cost=1, op=ATOM, pass_=1, launched=1, all_ltls=0, stored=0, to_check=A)
    gen = [x for x in recs[1:] if x["pass_"] == 1]
    # consecutive canonical ids over the generation passes
    nxt = A
    for x in gen:
        assert x["all_ltls"] == nxt
        nxt += x["to_check"]
    # levels generated in full, block by block, in the reference's order
    final = r["cost"] if r["status"] == 1 else 9
    for c in range(2, final):
        got = [(x["op"], x["to_check"]) for x in gen if x["cost"] == c]
        assert got == expected_blocks(s, e, c), c
        assert all(x["stored"] == s[c] for x in recs if x["cost"] == c and x["pass_"] == 1)
    if r["status"] == 1 and r["cost"] > 1:
        op = r["nodes"][0][0]
        hits = [x for x in recs if x["cost"] == r["cost"] and x["op"] == op
                and x["all_ltls"] <= r["found_id"] < x["all_ltls"] + x["to_check"]]
        assert hits, (r["found_id"], op)


def test_cli_launch_log(ltl, tmp_path):
    path = sorted(p for p in os.listdir(os.path.join(GOLDEN, "traces")) if p.endswith(".trace"))[0]
    path = os.path.join(GOLDEN, "traces", path)
    cli = os.path.join(PKG_DIR, "ltli")
    args = [path] + ["1"] * 8 + ["7", "3", "2"]
    plain = subprocess.run([cli] + args, capture_output=True, text=True, timeout=120)
    logged = subprocess.run([cli, "--launch-log"] + args, capture_output=True, text=True, timeout=120)
    assert plain.returncode == 0 and logged.returncode == 0
    lines = [ln for ln in logged.stdout.split("\n") if ln.startswith("Cost ") and " | (" in ln]
    assert lines and lines[0].startswith("Cost 1  | (A) | AllLTLs: 0 ")
    assert all(" | StoredLTLs: " in ln and " | ToBeChecked: " in ln for ln in lines)
    assert not any(ln.startswith("Cost ") and "| (" in ln for ln in plain.stdout.split("\n"))
    last = lambda out: out.split("\n")[-2].split()[-1][1:-1]    # code/dc.py:194
    assert last(logged.stdout) == last(plain.stdout)
