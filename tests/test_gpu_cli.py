"""The reference-compatible CLI on the GPU: a driver that parses its stdout
the way code/dc.py:192-194 does gets the oracle's formula."""
import glob
import os
import subprocess

import pytest

from conftest import GOLDEN, PKG_DIR

pytestmark = pytest.mark.gpu

CLI = os.path.join(PKG_DIR, "ltli")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "traces", "*.trace")))[:6],
                         ids=os.path.basename)
def test_cli_like_dc_py(oracle, path):
    out = subprocess.run([CLI, path, "1", "1", "1", "1", "1", "1", "1", "1", "7", "3", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    formula = out.stdout.split("\n")[-2].split()[-1][1:-1]     # code/dc.py:194
    ref = oracle.enumerate_(oracle.read_trace_file(path), max_cost=7, neg_type=2)
    assert formula == ref["formula"]
    assert f"Cost of Final LTL: {ref['cost']}" in out.stdout
    assert f"Number of All LTLs: {ref['all_ltls']}" in out.stdout
    assert "Number of Traces:" in out.stdout and "Positive:" in out.stdout


def test_cli_argument_errors():
    out = subprocess.run([CLI, "x", "1", "1", "1", "1", "1", "1", "1", "0", "7", "4", "2"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode != 0
    assert "should be a positive short integer" in out.stdout
    assert "RlxUnqChkTyp" in out.stdout


def _write_trace_file(path, inst):
    """An instance in the reference format (positions ';', bits ',')."""
    lines = []
    for t in range(inst.W):
        n = int(inst.lens[t])
        lines.append(";".join(",".join(str((int(inst.atoms[a, t]) >> k) & 1) for a in range(inst.atoms.shape[0]))
                              for k in range(n)))
    P = inst.num_pos
    with open(path, "w") as f:
        f.write("\n".join(lines[:P]) + "\n---\n" + "\n".join(lines[P:]) + "\n---\nF,G,X,!,&,|\n---\n"
                + ",".join(inst.names))


def _run(args, env=None):
    out = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, **(env or {})))
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout.split("\n")[-2].split()[-1][1:-1]     # code/dc.py:194


def test_cli_reference_keys(ltl, oracle, tmp_path):
    # RlxUnqChkTyp (argument 11) selects the reference's relaxed key when
    # asked to: the mirrored 4-trace instance answers X((a)U(b)) in mode 3,
    # X(b) with exact dedup (the default)
    from test_gpu_parity import _twin4, to_oracle
    inst = _twin4(ltl, 3)
    path = str(tmp_path / "twin.trace")
    _write_trace_file(path, inst)
    args = [path] + ["1"] * 8 + ["6", "3", "1"]
    assert _run(args) == "X(b)"
    assert _run(["--reference-keys"] + args) == "X((a)U(b))"
    assert _run(args, {"LTLI_REFERENCE_KEYS": "1"}) == "X((a)U(b))"
    ref = oracle.enumerate_(to_oracle(oracle, inst), max_cost=6, dedup_mode=3)
    assert ref["formula"] == "X((a)U(b))"


def test_cli_more_than_26_propositions(ltl, oracle, tmp_path):
    from test_host import write_named_traces
    path = str(tmp_path / "named.trace")
    pos, neg, names = write_named_traces(path)
    f = _run([path] + ["1"] * 8 + ["6", "3", "1"])
    inst = ltl.read_trace_file(path)
    # the oracle spells atom i with one character (a-z, then capitals other
    # than the operators X, F, G, U): its formula, re-spelled with the names
    chars = [chr(ord("a") + i) for i in range(26)] + [c for c in "ABCDEHIJKLMNOPQRSTVWYZ"][:4]
    oi = oracle.Instance(inst.num_pos, inst.num_neg, inst.lens, inst.atoms, "".join(chars))
    ref = oracle.enumerate_(oi, max_cost=6)
    assert ref["status"] == 1
    respelled = "".join(names[chars.index(c)] if c in chars else c for c in ref["formula"])
    assert f == respelled
    assert ltl.mismatches(inst, ltl.eval_formula(inst, f)) == 0


def _envelope():
    import json
    with open(os.path.join(GOLDEN, "envelope.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _envelope(), ids=lambda c: c["file"][-40:])
def test_cli_envelope_files(case):
    """Every reference benchmark file within the reference's single-run
    limits (69 of 611), through the CLI as the reference's example runs it:
    the oracle's formula, cost and candidate count, exact and with
    --reference-keys (RlxUnqChkTyp 3, the notebook's default)."""
    path = os.path.join(GOLDEN, "envelope", case["file"])
    out = subprocess.run([CLI, path] + ["1"] * 8 + [str(case["max_cost"]), "3", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.split("\n")[-2].split()[-1][1:-1] == case["formula"]
    assert f"Cost of Final LTL: {case['cost']}\n" in out.stdout
    assert f"Number of All LTLs: {case['all_ltls']}\n" in out.stdout
    r3 = case["ref3"]
    out = subprocess.run([CLI, "--reference-keys", path] + ["1"] * 8 + [str(r3["max_cost"]), "3", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.split("\n")[-2].split()[-1][1:-1] == r3["formula"]
    assert f"Cost of Final LTL: {r3['cost']}\n" in out.stdout
    assert f"Number of All LTLs: {r3['all_ltls']}\n" in out.stdout
