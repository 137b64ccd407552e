"""The C++ drop-in layer on the GPU: MatrixCSR products (the reference's
MathTest known answers), FitWeights against the reference-structured host PCG
of Solver.h, ScoreEstimator::EvalBatch against Eval, device LSQR."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "cassandra_amd", "_build", "dropin_test")


def test_dropin_gpu_cases():
    print("dropin gpu cases", flush=True)
    r = subprocess.run([RUNNER, "gpu"], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
