"""The library's host-only code (trace file readers, formula parser and
evaluator, instance generator) built with AddressSanitizer and
UndefinedBehaviorSanitizer and driven over the golden trace files, as they
are and mutated, plus random and damaged formulas (make sanitize in the
package's csrc/)."""
import os
import shutil
import subprocess

import pytest

from conftest import PKG_DIR


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_host_code_under_sanitizers():
    out = subprocess.run(["make", "-s", "-C", os.path.join(PKG_DIR, "csrc"), "sanitize"],
                         capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-3000:]
    assert "0 failures" in out.stdout
    assert "ERROR: AddressSanitizer" not in out.stderr and "runtime error" not in out.stderr
