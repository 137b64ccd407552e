"""Expanded tables on the device against the oracle: expansion of compact
weights (ScoreEstimator's constructor, Pattern/Evaluator.cpp:11-29),
compaction back (ScoreEstimator::Weights, Pattern/Evaluator.cpp:56-73) and
the resident evaluator (ScoreEstimator::Eval, Pattern/Evaluator.cpp:48-54),
all bit for bit, over every symmetry class."""
import sys

import numpy as np
import pytest
import torch

from cassandra_amd import names as N
from cassandra_amd.fit import Patterns

pytestmark = pytest.mark.gpu

SETS = {"fit4": N.FIT4, "classes": N.test_mix, "fit12": N.FIT12, "b6": [N.B6, N.D8]}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", list(SETS))
def test_expand_and_compact_match_oracle(ctx, oracle, name):
    say("tables", name)
    masks = SETS[name]
    P = Patterns(ctx, masks)
    g = oracle.Group(masks)
    w = np.random.default_rng(3).normal(size=P.cols).astype(np.float32)
    want = g.expand(w)
    assert P.expanded_size() == want.size
    tables = P.expand(torch.from_numpy(w).cuda())
    assert np.array_equal(tables.cpu().numpy(), want)
    # the round trip gives the weights back, as the oracle's does
    back = P.compact(tables).cpu().numpy()
    assert np.array_equal(back, g.compact(want))
    assert np.array_equal(back, w)
    # tables that are not an expansion: the compaction keeps, for each column,
    # the last configuration in enumeration order, as the reference does
    t = np.random.default_rng(4).normal(size=want.size).astype(np.float32)
    assert np.array_equal(P.compact(torch.from_numpy(t).cuda()).cpu().numpy(), g.compact(t))


@pytest.mark.parametrize("name", ["fit4", "classes", "fit12"])
def test_resident_evaluator_matches_eval(ctx, oracle, name):
    say("resident evaluator", name)
    masks = SETS[name]
    P = Patterns(ctx, masks)
    g = oracle.Group(masks)
    rng = np.random.default_rng(5)
    w = rng.normal(size=P.cols).astype(np.float32)
    pos = oracle.random_play(91, 30000, 5, 55)
    E = P.evaluator(torch.from_numpy(w).cuda())
    want = g.eval_expanded(g.expand(w), pos)
    assert np.array_equal(want, oracle.spmv(g.indices(pos), w))  # Eval(pos) == (A x)[row]
    got = E(torch.from_numpy(pos.view(np.int64)).cuda()).cpu().numpy()
    assert np.array_equal(got, want)
    # host batches through the evaluator's staging, a larger then a smaller one
    assert np.array_equal(E.eval_host(pos), want)
    assert np.array_equal(E.eval_host(pos[:777]), want[:777])
    # new weights replace the tables
    w2 = rng.normal(size=P.cols).astype(np.float32)
    E.set_weights(torch.from_numpy(w2).cuda())
    assert np.array_equal(E.eval_host(pos), g.eval_expanded(g.expand(w2), pos))
