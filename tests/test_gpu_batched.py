"""Batched buckets: several independent systems (one per empty-count stage, of
uneven sizes) in one block-diagonal matrix, solved in the same launches with
per-system scalars.  Each system is checked against the oracle run on that
system alone (PatternFit/PatternFit.cpp:41-84 fits the stages one by one)."""
import sys

import numpy as np
import pytest
import torch

import cassandra_amd.capi as capi
from cassandra_amd import names as N
from cassandra_amd.fit import Patterns

pytestmark = pytest.mark.gpu

# (empties, rows): uneven, including a tiny system
SYSTEMS = [(12, 3000), (20, 777), (33, 12000), (47, 5), (50, 5000), (24, 2500)]
# Truncated iterates are compared on the systems of many distinct rows: the
# 5-row system is solved exactly within 5 steps, after which both solvers
# iterate on rounding noise (so would a system near the start position, which
# has only a few distinct positions).
LARGE = [s for s, (_, n) in enumerate(SYSTEMS) if n >= 100]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch(ctx, oracle):
    g = oracle.Group(N.FIT4)
    pos, idx, rb = [], [], [0]
    for s, (e, n) in enumerate(SYSTEMS):
        p = oracle.random_play(500 + s, n, e)
        pos.append(p)
        idx.append(g.indices(p))
        rb.append(rb[-1] + n)
    allpos = torch.from_numpy(np.concatenate(pos).view(np.int64)).cuda()
    P = Patterns(ctx, N.FIT4)
    A = P.matrix_systems(allpos, rb)
    xt, b = A.make_targets(2100, 3100)
    return g.cols, idx, rb, A, b


def blocks(cols, idx):
    """Global column indices of the block-diagonal matrix."""
    return np.concatenate([i + np.uint32(s * cols) for s, i in enumerate(idx)])


def test_layout_and_products(ctx, oracle, batch):
    cols, idx, rb, A, b = batch
    say("batched layout")
    assert (A.nsys, A.cols_per_system, A.cols, A.rows) == (len(SYSTEMS), cols, cols * len(SYSTEMS), rb[-1])
    assert np.array_equal(A.row_begin, rb)
    G = blocks(cols, idx)
    assert np.array_equal(A.download(), G)
    x = torch.randn(A.cols, dtype=torch.float32, device="cuda")
    assert np.array_equal(A.spmv(x).cpu().numpy(), oracle.spmv(G, x.cpu().numpy()))
    y = torch.randint(-64, 65, (A.rows,), device="cuda").float()
    assert np.array_equal(A.spmtv(y).cpu().numpy(), oracle.spmtv(G, y.cpu().numpy(), A.cols))
    jo = oracle.jacobi(G, A.cols)
    assert np.array_equal(A.jacobi().cpu().numpy(), jo)


def colscale(oracle, idx, cols):
    d = oracle.jacobi(idx, cols).astype(np.float64)
    ok = np.isfinite(d) & (d > 0)
    cnt = np.rint(1.0 / np.where(ok, d, 1.0))
    return np.where(ok, 1.0 / np.sqrt(cnt), 1.0)


@pytest.mark.parametrize("iters", [5, 10])
def test_fp64_lsqr_iterates_per_system(ctx, oracle, batch, iters):
    cols, idx, rb, A, b = batch
    say("batched lsqr fp64", iters)
    x, rep, rel, its = A.solve_systems(b, precision=capi.PREC_FP64, tol=0.0, max_iterations=iters, check_every=5)
    assert rep["iterations"] == iters and np.all(its == iters)
    xh, bh = x.cpu().numpy(), b.cpu().numpy()
    for s in LARGE:
        bs = bh[rb[s]:rb[s + 1]]
        xs = xh[s * cols:(s + 1) * cols]
        xo, _, _ = oracle.lsqr(idx[s], cols, bs, iters, colscale=colscale(oracle, idx[s], cols))
        pg, po = oracle.spmv(idx[s], xs), oracle.spmv(idx[s], xo)
        assert np.linalg.norm(pg - po) <= 1e-7 * np.linalg.norm(po), s
        assert rel[s] == pytest.approx(oracle.rel_normal_residual(idx[s], cols, bs, xs)[0], rel=1e-6)


def test_mixed_reaches_tolerance_per_system(ctx, oracle, batch):
    cols, idx, rb, A, b = batch
    say("batched mixed")
    x, rep, rel, its = A.solve_systems(b, precision=capi.PREC_MIXED, tol=1e-10)
    say(rep, rel, its)
    assert rep["converged"] == 1
    xh, bh = x.cpu().numpy(), b.cpu().numpy()
    for s in range(len(SYSTEMS)):
        bs = bh[rb[s]:rb[s + 1]]
        r = oracle.rel_normal_residual(idx[s], cols, bs, xh[s * cols:(s + 1) * cols])[0]
        assert r <= 1.5e-10, (s, r)
        assert rel[s] <= 1e-10
    assert np.all(its <= rep["iterations"]) and its.max() == rep["iterations"]


def test_batched_equals_separate_fits(ctx, oracle, batch):
    # the batch gives each system the fit it gets alone (FP64 LSQR to 1e-12)
    cols, idx, rb, A, b = batch
    say("batched vs separate")
    x, rep, rel, its = A.solve_systems(b, precision=capi.PREC_FP64, tol=1e-12, max_iterations=20000)
    assert rep["converged"] == 1
    P = Patterns(ctx, N.FIT4)
    for s in (0, 2, 3):
        pos_s = torch.from_numpy(np.ascontiguousarray(
            oracle.random_play(500 + s, SYSTEMS[s][1], SYSTEMS[s][0])).view(np.int64)).cuda()
        As = P.matrix(pos_s)
        xs, reps = As.solve(b[rb[s]:rb[s + 1]].contiguous(), precision=capi.PREC_FP64, tol=1e-12,
                            max_iterations=20000)
        pb = oracle.spmv(idx[s], x[s * cols:(s + 1) * cols].cpu().numpy())
        ps = oracle.spmv(idx[s], xs.cpu().numpy())
        assert np.linalg.norm(pb - ps) <= 1e-8 * np.linalg.norm(ps), s


def test_pcg_per_system(ctx, oracle, batch):
    cols, idx, rb, A, b = batch
    say("batched pcg")
    x, rep, rel, its = A.solve_systems(b, solver=capi.SOLVER_PCG, precision=capi.PREC_FP64, tol=0.0,
                                       max_iterations=10, check_every=10)
    xh, bh = x.cpu().numpy(), b.cpu().numpy()
    for s in LARGE:
        xo, _ = oracle.pcg(idx[s], cols, bh[rb[s]:rb[s + 1]], 10)
        pg, po = oracle.spmv(idx[s], xh[s * cols:(s + 1) * cols]), oracle.spmv(idx[s], xo)
        assert np.linalg.norm(pg - po) <= 1e-9 * np.linalg.norm(po), s


def test_argument_errors(ctx):
    P = Patterns(ctx, N.FIT4)
    pos = ctx.random_play(1, 100, 30)
    with pytest.raises(capi.CadError):
        P.matrix_systems(pos, [0, 50, 50, 100])  # an empty system
    with pytest.raises(ValueError):
        P.matrix_systems(pos, [0, 50])


def test_mixed_finishes_degenerate_systems_in_fp64(ctx, oracle):
    # positions one or two plies from the start (4 and 12 distinct positions,
    # thousands of copies): FP32 products cannot resolve their residuals, so
    # the refined solve must hand them to FP64 products and still meet tol
    say("batched mixed, degenerate systems")
    systems = [(58, 5000), (59, 3000), (30, 4000)]
    g = oracle.Group(N.FIT4)
    pos, idx, rb = [], [], [0]
    for s, (e, n) in enumerate(systems):
        p = oracle.random_play(600 + s, n, e)
        pos.append(p)
        idx.append(g.indices(p))
        rb.append(rb[-1] + n)
    P = Patterns(ctx, N.FIT4)
    A = P.matrix_systems(torch.from_numpy(np.concatenate(pos).view(np.int64)).cuda(), rb)
    _, b = A.make_targets(2200, 3200)
    x, rep, rel, its = A.solve_systems(b, precision=capi.PREC_MIXED, tol=1e-10)
    say(rep, rel, its)
    assert rep["converged"] == 1
    xh, bh = x.cpu().numpy(), b.cpu().numpy()
    for s in range(len(systems)):
        r = oracle.rel_normal_residual(idx[s], g.cols, bh[rb[s]:rb[s + 1]], xh[s * g.cols:(s + 1) * g.cols])[0]
        assert r <= 1.5e-10, (s, r)
