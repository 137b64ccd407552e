"""The C++ drop-in layer (the reference's API over the C-ABI): the reference's
MathTest / PatternTest cases that need no device."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "cassandra_amd", "_build", "dropin_test")


def test_dropin_host_cases():
    if not os.path.exists(RUNNER):
        pytest.fail(f"{RUNNER} missing: run __graft_entry__.build()")
    r = subprocess.run([RUNNER, "host"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
