"""Differentiable evaluator: forward gather-sum and hand-written A^T g backward
against the oracle, central finite differences (h = 1e-3, FP64) and the CSR
A^T product; Huber/Adam training against the least-squares fit."""
import sys

import numpy as np
import pytest
import torch

import cassandra_amd.capi as capi
from cassandra_amd import names as N
from cassandra_amd.evaluator import PatternEvaluator, huber_fit, pattern_eval
from cassandra_amd.fit import Patterns

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c1(ctx, oracle):
    pos = oracle.random_play(1004, 30001, 30)
    P = Patterns(ctx, N.FIT4)
    d = torch.from_numpy(pos.view(np.int64)).cuda()
    idx = P.extract(d, capi.LAYOUT_ROW_MAJOR)
    return P, pos, d, idx, oracle.Group(N.FIT4).indices(pos)


def test_forward_bit_exact(ctx, oracle, c1):
    P, pos, d, idx, oidx = c1
    say("eval forward")
    assert np.array_equal(idx.cpu().numpy().view(np.uint32), oidx)
    w = torch.randn(P.cols, dtype=torch.float32, device="cuda")
    out = pattern_eval(w, idx, P)
    assert np.array_equal(out.cpu().numpy(), oracle.spmv(oidx, w.cpu().numpy()))
    # the expanded-table ScoreEstimator::Eval gives the same float (same terms, same order)
    g = oracle.Group(N.FIT4)
    tables = g.expand(w.cpu().numpy())
    assert np.array_equal(out.cpu().numpy(), g.eval_expanded(tables, pos))
    w64 = w.double()
    assert np.array_equal(pattern_eval(w64, idx, P).cpu().numpy(), oracle.spmv(oidx, w64.cpu().numpy()))


def test_backward_against_oracle_and_csr(ctx, oracle, c1):
    P, pos, d, idx, oidx = c1
    say("eval backward")
    w = torch.zeros(P.cols, dtype=torch.float32, device="cuda", requires_grad=True)
    g = torch.randint(-64, 65, (idx.shape[0],), device="cuda").float()
    pattern_eval(w, idx, P).backward(g)
    want = oracle.spmtv(oidx, g.cpu().numpy(), P.cols)
    assert np.array_equal(w.grad.cpu().numpy(), want)  # integer sums: exact
    A = P.matrix(d)
    assert torch.equal(w.grad, A.spmtv(g))  # the CSR A^T product
    w64 = torch.zeros(P.cols, dtype=torch.float64, device="cuda", requires_grad=True)
    gr = torch.randn(idx.shape[0], dtype=torch.float64, device="cuda")
    pattern_eval(w64, idx, P).backward(gr)
    ref = oracle.spmtv(oidx, gr.cpu().numpy(), P.cols)
    bound = oracle.spmtv(oidx, np.abs(gr.cpu().numpy()), P.cols)
    assert np.all(np.abs(w64.grad.cpu().numpy() - ref) <= 1e-13 * bound + 1e-300)


def test_backward_full_minibatch_against_oracle(ctx, oracle):
    # the training workload's minibatch: 65,536 rows (the backward's scatter
    # plan for this size splits the rows into chunks that give every CU a tile)
    say("eval backward, 65,536-row minibatch")
    P = Patterns(ctx, N.FIT4)
    pos = ctx.random_play(1005, 65536, 30)
    idx = P.extract(pos, capi.LAYOUT_ROW_MAJOR)
    oidx = oracle.Group(N.FIT4).indices(pos.cpu().numpy().view(np.uint64))
    assert np.array_equal(idx.cpu().numpy().view(np.uint32), oidx)
    gen = torch.Generator(device="cuda").manual_seed(5)
    gi = torch.randint(-64, 65, (idx.shape[0],), device="cuda", generator=gen).float()
    w = torch.zeros(P.cols, dtype=torch.float32, device="cuda", requires_grad=True)
    pattern_eval(w, idx, P).backward(gi)
    assert np.array_equal(w.grad.cpu().numpy(), oracle.spmtv(oidx, gi.cpu().numpy(), P.cols))  # exact
    gr = torch.randn(idx.shape[0], dtype=torch.float32, device="cuda", generator=gen)
    w.grad = None
    pattern_eval(w, idx, P).backward(gr)
    g64 = gr.cpu().numpy().astype(np.float64)
    ref = oracle.spmtv(oidx, g64, P.cols)
    bound = oracle.spmtv(oidx, np.abs(g64), P.cols)
    # FP64 window sums rounded to float per row chunk, the chunks summed in FP64
    assert np.all(np.abs(w.grad.cpu().numpy().astype(np.float64) - ref) <= 2.0 ** -23 * bound + 1e-30)
    w64 = torch.zeros(P.cols, dtype=torch.float64, device="cuda", requires_grad=True)
    pattern_eval(w64, idx, P).backward(gr.double())
    assert np.all(np.abs(w64.grad.cpu().numpy() - ref) <= 1e-13 * bound + 1e-300)


def test_gradcheck_small(ctx, oracle):
    say("gradcheck")
    masks = [N.test_a, N.test_d, N.test_vh, N.D4]
    P = Patterns(ctx, masks)
    pos = torch.from_numpy(oracle.random_play(5, 48, 40).view(np.int64)).cuda()
    idx = P.extract(pos, capi.LAYOUT_ROW_MAJOR)
    w = torch.randn(P.cols, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda v: pattern_eval(v, idx, P), (w,), eps=1e-3, atol=1e-7, rtol=1e-7)


def test_central_differences_directional(ctx, oracle, c1):
    # <grad, v> against (L(w + h v) - L(w - h v)) / 2h for a nonlinear loss, FP64
    P, pos, d, idx, oidx = c1
    say("central differences")
    torch.manual_seed(0)
    w = torch.randn(P.cols, dtype=torch.float64, device="cuda", requires_grad=True)
    t = torch.randn(idx.shape[0], dtype=torch.float64, device="cuda") * 8
    loss = lambda v: torch.nn.functional.huber_loss(pattern_eval(v, idx, P), t, delta=4.0)  # noqa: E731
    loss(w).backward()
    h = 1e-3
    for _ in range(3):
        v = torch.randn(P.cols, dtype=torch.float64, device="cuda")
        with torch.no_grad():
            fd = (loss(w + h * v) - loss(w - h * v)) / (2 * h)
        an = torch.dot(w.grad, v)
        assert abs(float(fd - an)) <= 1e-6 * max(1.0, abs(float(an)))


def test_huber_fit_against_least_squares(ctx, oracle):
    say("huber fit")
    P = Patterns(ctx, N.FIT4)
    pos = ctx.random_play(1004, 200000, 30)
    A = P.matrix(pos)
    _, b = A.make_targets(2004, 3004)
    idx = P.extract(pos, capi.LAYOUT_ROW_MAJOR)
    model, losses = huber_fit(P, idx, b.float(), epochs=6, batch=65536, lr=0.05)
    # Huber(4) of zero weights is ~16 here; the noise (sd 4) alone leaves ~5
    assert all(b_ < a_ for a_, b_ in zip(losses, losses[1:]))
    assert losses[-1] < 0.75 * losses[0]
    x_ls, rep = A.solve(b, precision=capi.PREC_MIXED, tol=1e-8)
    with torch.no_grad():
        p_h = model(idx).double()
    p_ls = A.spmv(x_ls)
    # both fits explain the targets far better than zero weights
    r0 = float(torch.mean(b ** 2))
    assert float(torch.mean((b - p_ls) ** 2)) < 0.5 * r0
    assert float(torch.mean((b - p_h) ** 2)) < 0.7 * r0
