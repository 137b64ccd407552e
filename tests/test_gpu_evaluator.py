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


def test_fused_train_step_closed_form(ctx, oracle, c1):
    """One fused step from zero weights over all rows (no shuffle): pred = 0,
    so the loss is mean huber(-t) and the gradient g = A^T clamp(-t, -4, 4) / n
    (the oracle's FP64 A^T); Adam's first step is w = -lr g / (|g| + eps)
    (bias corrections cancel)."""
    from cassandra_amd.evaluator import huber_fit_fused

    say("fused step, closed form")
    P, pos, d, idx, ih = c1
    n = idx.shape[0]
    rng = np.random.default_rng(11)
    t = (2.0 * rng.integers(-32, 33, size=n)).astype(np.float32)  # even integers in [-64, 64]
    model, losses = huber_fit_fused(P, idx, torch.from_numpy(t).cuda(), epochs=1, batch=n, lr=0.05,
                                    shuffle=False)
    w = model.weight.detach().cpu().numpy().astype(np.float64)
    dd = -t.astype(np.float64)
    loss = np.where(np.abs(dd) < 4.0, 0.5 * dd * dd, 4.0 * (np.abs(dd) - 2.0)).mean()
    assert abs(losses[0] - loss) <= 1e-12 * loss
    g = oracle.spmtv(ih, np.clip(dd, -4.0, 4.0) / n, P.cols)
    want = -0.05 * g / (np.abs(g) + 1e-8)
    # the exact gradients are integer multiples of 2/n (>= 6.7e-5) or 0 (the
    # oracle's FP64 sum leaves ~1e-19 where the terms cancel)
    nz = np.abs(g) > 1e-9
    err = np.max(np.abs(w - want)[nz])
    say(f"max |w - closed form| where g != 0: {err:.2e}")
    assert err <= 1e-6
    # where the exact gradient is 0, the device's float partial sums may leave
    # a residual up to 2^-22 of the column's sum of |terms|, and Adam's first
    # step moves the weight by lr |g| / (|g| + eps)
    bound = oracle.spmtv(ih, np.abs(np.clip(dd, -4.0, 4.0)) / n, P.cols) * 2.0 ** -22
    lim = 0.05 * bound / (bound + 1e-8)
    assert np.all(np.abs(w[~nz]) <= lim[~nz] + 1e-12)


def test_fused_fit_matches_torch_path(ctx, oracle):
    """The fused steps against huber_fit's torch path (autograd Function,
    torch's Huber loss, torch's fused Adam) on the same shuffled minibatches.
    The two compute the same gradient up to one rounding of 1/B per row and
    Adam's update in another float order; Adam normalises each column's
    gradient, so where a column's gradient nearly cancels those roundings move
    its step by a fraction of lr.  Bounds: the weights' relative L2 distance
    1e-5, any weight within 1 % of lr (5e-4), the epoch losses within 1e-5."""
    from cassandra_amd.evaluator import huber_fit_fused

    say("fused fit vs torch path")
    P = Patterns(ctx, N.FIT4)
    pos = ctx.random_play(1014, 20000, 30)
    A = P.matrix(pos)
    _, b = A.make_targets(2014, 3014)
    idx = P.extract(pos, capi.LAYOUT_ROW_MAJOR)
    mt, lt = huber_fit(P, idx, b.float(), epochs=2, batch=4096, lr=0.05, seed=5)
    mf, lf = huber_fit_fused(P, idx, b.float(), epochs=2, batch=4096, lr=0.05, seed=5)
    torch.cuda.synchronize()
    wt, wf = mt.weight.detach().double(), mf.weight.detach().double()
    err = float(torch.max(torch.abs(wt - wf)))
    rel = float(torch.linalg.norm(wt - wf) / torch.linalg.norm(wt))
    say(f"max |w_torch - w_fused| = {err:.2e}, relative L2 {rel:.2e}, losses {lt} {lf}")
    assert err <= 5e-4
    assert rel <= 1e-5
    for a_, b_ in zip(lt, lf):
        assert abs(a_ - b_) <= 1e-5 * abs(a_)


def test_train_step_calls_equal_epoch_call(ctx, oracle, c1):
    """cad_eval_train_step_f32 called once per minibatch gives bit for bit the
    weights, moments and loss of one cad_eval_train_epoch_f32 call over the
    same order (the same kernels in the same order)."""
    import ctypes as C

    say("train step calls vs epoch call")
    P, pos, d, idx, ih = c1
    n, B = idx.shape[0], 4096
    rng = np.random.default_rng(12)
    tgt = torch.from_numpy((2.0 * rng.integers(-32, 33, size=n)).astype(np.float32)).cuda()
    order = torch.from_numpy(rng.permutation(n).astype(np.int64)).cuda()
    L = capi.lib()
    ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    outs = []
    for per_step in (True, False):
        w = torch.zeros(P.cols, dtype=torch.float32, device="cuda")
        m = torch.zeros_like(w)
        v = torch.zeros_like(w)
        loss = torch.zeros((), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        if per_step:
            for i, b0 in enumerate(range(0, n, B)):
                sel = order[b0:b0 + B]
                capi.check(L.cad_eval_train_step_f32(P.h, ptr(idx), ptr(tgt), ptr(sel), sel.shape[0], 4.0, 0.05, 0.9,
                                                     0.999, 1e-8, i + 1, ptr(w), ptr(m), ptr(v), ptr(loss)))
        else:
            capi.check(L.cad_eval_train_epoch_f32(P.h, ptr(idx), ptr(tgt), ptr(order), n, B, 4.0, 0.05, 0.9, 0.999,
                                                  1e-8, 1, ptr(w), ptr(m), ptr(v), ptr(loss)))
        ctx.synchronize()
        outs.append((w.clone(), m.clone(), v.clone(), float(loss)))
    (w0, m0, v0, l0), (w1, m1, v1, l1) = outs
    assert torch.equal(w0, w1) and torch.equal(m0, m1) and torch.equal(v0, v1)
    assert l0 == l1 and l0 > 0


def test_fused_train_step_wide_rows_closed_form(ctx, oracle):
    """The one-step closed form of the fused step on a second pattern group
    with wider rows (FIT12 with L02X and B5 once more: k = 66)."""
    from cassandra_amd.evaluator import huber_fit_fused

    say("fused step, 66 slots a row")
    masks = list(N.FIT12) + [N.L02X, N.B5]
    P = Patterns(ctx, masks)
    assert P.k == 66
    pos = oracle.random_play(1024, 9001, 30)
    d = torch.from_numpy(pos.view(np.int64)).cuda()
    idx = P.extract(d, capi.LAYOUT_ROW_MAJOR)
    ih = oracle.Group(masks).indices(pos)
    n = idx.shape[0]
    rng = np.random.default_rng(13)
    t = (2.0 * rng.integers(-32, 33, size=n)).astype(np.float32)
    model, losses = huber_fit_fused(P, idx, torch.from_numpy(t).cuda(), epochs=1, batch=n, lr=0.05,
                                    shuffle=False)
    w = model.weight.detach().cpu().numpy().astype(np.float64)
    dd = -t.astype(np.float64)
    loss = np.where(np.abs(dd) < 4.0, 0.5 * dd * dd, 4.0 * (np.abs(dd) - 2.0)).mean()
    assert abs(losses[0] - loss) <= 1e-12 * loss
    g = oracle.spmtv(ih, np.clip(dd, -4.0, 4.0) / n, P.cols)
    nz = np.abs(g) > 1e-9
    err = np.max(np.abs(w - (-0.05 * g / (np.abs(g) + 1e-8)))[nz])
    say(f"max |w - closed form| where g != 0: {err:.2e}")
    assert err <= 1e-6
