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
