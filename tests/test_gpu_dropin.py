"""The C++ drop-in layer on the GPU: MatrixCSR products (the reference's
MathTest known answers), FitWeights against the reference-structured host PCG
of Solver.h, ScoreEstimator::EvalBatch, device LSQR -- and, through the
runner's file modes, FitStages and EvalBatch against the oracle."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from cassandra_amd import names as N

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "cassandra_amd", "_build", "dropin_test")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(*args):
    r = subprocess.run([RUNNER, *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def run_u64(a):
    a = np.ascontiguousarray(a)
    return np.array([a.size], np.uint64).tobytes() + a.tobytes()


def read_runs(data, dtype, count):
    out, at = [], 0
    for _ in range(count):
        n = int(np.frombuffer(data, np.uint64, 1, at)[0])
        at += 8
        out.append(np.frombuffer(data, dtype, n, at).copy())
        at += n * np.dtype(dtype).itemsize
    return out


def test_dropin_gpu_cases():
    say("dropin gpu cases")
    r = run("gpu")
    print(r.stdout)
    assert "0 failures" in r.stdout


def test_fit_stages_against_oracle(oracle):
    # PatternFit/PatternFit.cpp:72-83: every stage of the estimator is fitted by
    # FitWeights on the positions of its empty-count range; FitStages fits them
    # all in one batch.  Each stage's weights are checked against the oracle's
    # FP64 Jacobi-PCG on that stage's rows alone.  FitWeights iterates in
    # float; how far float rounding carries PCG's iterates on these singular
    # systems is measured on the oracle itself (its float PCG against its FP64
    # one, 1e-7 to 4e-5 here after 5 iterations), and the GPU's predictions
    # must lie within 4 times that distance (at least 1e-5) of the FP64 ones.
    say("FitStages against the oracle")
    masks = np.array(N.FIT4, np.uint64)
    g = oracle.Group(N.FIT4)
    pos = oracle.random_play(77, 24000, 20, 44)
    idx = g.indices(pos)
    _, b = oracle.make_targets(idx, g.cols, 2077, 3077)
    score = (b / 2).astype(np.int8)  # Score: half the disc difference
    stage_size, iterations = 5, 5
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([stage_size, iterations], np.int32).tobytes())
            f.write(run_u64(masks) + run_u64(pos.reshape(-1)) + run_u64(score))
        run("fit-stages", fin, fout)
        data = open(fout, "rb").read()
    stages = int(np.frombuffer(data, np.uint64, 1, 0)[0])
    assert stages == 13  # ceil(65 / 5)
    w = read_runs(data[8:], np.float32, stages)
    empties = np.array([64 - bin(int(p) | int(o)).count("1") for p, o in pos])
    fitted = 0
    for st in range(stages):
        rows = np.flatnonzero(empties // stage_size == st)
        if rows.size == 0:
            assert not np.any(w[st])
            continue
        bs = score[rows].astype(np.float64)
        xo, _ = oracle.pcg(idx[rows], g.cols, bs, iterations)
        x32, _ = oracle.pcg(idx[rows], g.cols, bs, iterations, dtype=np.float32)
        po = oracle.spmv(idx[rows], xo)
        norm = np.linalg.norm(po)
        d32 = np.linalg.norm(oracle.spmv(idx[rows], x32.astype(np.float64)) - po) / norm
        err = np.linalg.norm(oracle.spmv(idx[rows], w[st].astype(np.float64)) - po) / norm
        say(f"  stage {st}: {rows.size} rows, prediction error {err:.2e} (oracle float vs FP64 {d32:.2e})")
        assert err <= max(1e-5, 4 * d32)
        fitted += 1
    assert fitted == 5


def test_eval_batch_against_oracle(oracle):
    # ScoreEstimator::EvalBatch (the batched Pattern/Evaluator.cpp:48-54) must
    # equal the oracle's expanded-table Eval bit for bit
    say("EvalBatch against the oracle")
    g = oracle.Group(N.FIT4)
    w = np.random.default_rng(11).normal(size=g.cols).astype(np.float32)
    pos = oracle.random_play(78, 20000, 10, 50)
    want = g.eval_expanded(g.expand(w), pos)
    assert np.array_equal(want, oracle.spmv(g.indices(pos), w))
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(run_u64(np.array(N.FIT4, np.uint64)) + run_u64(w) + run_u64(pos.reshape(-1)))
        run("eval-batch", fin, fout)
        got = read_runs(open(fout, "rb").read(), np.float32, 1)[0]
    assert np.array_equal(got, want)
