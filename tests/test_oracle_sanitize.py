"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: a gcc
build of every oracle source with a driver that calls each entry point on
small inputs (all five models, the three capture modes, split cells, both
moment methods, errors, acceptance, model probabilities, data statistics).
The build stops at the first report, so a clean exit means none."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_oracle_is_clean_under_sanitizers():
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr
    assert "runtime error" not in r.stderr
    assert "data stats" in r.stdout
