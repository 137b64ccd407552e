"""The batched-bucket workload at its full size: 50 independent systems, one
per empty count 10..59, rows drawn log-uniform in [20k, 800k] (seed 1003,
about 10.7M rows), pattern set {L02X, B5, Q3, D8}; all solved in the same
launches by the mixed-precision refined LSQR to 1e-10 (PatternFit/
PatternFit.cpp:41-84 fits the stages one FitWeights call at a time).  Every
system's FP64 metric |A_s^T(b_s - A_s x_s)| / |A_s^T b_s| is recomputed by
the oracle from its own indices of the same positions."""
import sys

import numpy as np
import pytest
import torch

import cassandra_amd.capi as capi
from cassandra_amd import names as N
from cassandra_amd.fit import Patterns

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bucket_rows():
    rng = np.random.default_rng(1003)
    empties = list(range(10, 60))
    rows = np.rint(np.exp(rng.uniform(np.log(20000), np.log(800000), size=len(empties)))).astype(np.int64)
    return empties, rows


def test_fifty_buckets_reach_tolerance_against_oracle(ctx, oracle):
    empties, rows = bucket_rows()
    say(f"buckets workload: 50 systems, {int(rows.sum())} rows")
    P = Patterns(ctx, N.FIT4)
    parts = [ctx.random_play(1003 + 7919 * i, int(n), e) for i, (e, n) in enumerate(zip(empties, rows))]
    rb = np.concatenate([[0], np.cumsum(rows)])
    pos = torch.cat(parts)
    del parts
    A = P.matrix_systems(pos, rb)
    _, b = A.make_targets(2003, 3003)
    say("buckets workload: solving")
    x, rep, rel, its = A.solve_systems(b, precision=capi.PREC_MIXED, tol=1e-10, max_iterations=20000)
    say(f"buckets workload: {rep}, iterations per system {its.tolist()}")
    assert rep["converged"] == 1, rep
    assert np.all(rel <= 1e-10), {int(s): float(rel[s]) for s in np.flatnonzero(rel > 1e-10)}
    say("buckets workload: oracle indices and metrics")
    pos_h = pos.cpu().numpy().view(np.uint64)
    # the device's random play is the oracle's (checked on a sample here)
    for s in (0, 25, 49):
        r0 = int(rb[s])
        assert np.array_equal(pos_h[r0:r0 + 500], oracle.random_play(1003 + 7919 * s, 500, empties[s]))
    g = oracle.Group(N.FIT4)
    cols = g.cols
    assert A.cols_per_system == cols
    bh, xh = b.cpu().numpy(), x.cpu().numpy()
    for s in range(len(empties)):
        r0, r1 = int(rb[s]), int(rb[s + 1])
        idx = g.indices(pos_h[r0:r1])
        rel_o = oracle.rel_normal_residual(idx, cols, bh[r0:r1], xh[s * cols:(s + 1) * cols])[0]
        assert rel_o <= 1e-10, (s, empties[s], rel_o)
        # the device's FP64 metric and the oracle's differ only by summation order
        assert abs(rel_o - rel[s]) <= 1e-2 * rel_o + 1e-14, (s, rel_o, rel[s])
