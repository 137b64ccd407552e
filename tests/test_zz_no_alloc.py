"""The no-allocation evaluator backward (cad_eval_backward_f32_noalloc)
against cad_eval_backward_f32 and the oracle, on two calls in a row (the
second with fewer rows, so stale scratch contents of the first must not leak
into it)."""
import ctypes as C
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


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _backward(name, P, idx, g):
    out = torch.full((P.cols,), float("nan"), dtype=torch.float32, device="cuda")
    capi.call(name, P.h, _ptr(idx), idx.shape[0], _ptr(g), _ptr(out))
    P.ctx.synchronize()
    return out.cpu().numpy()


def test_noalloc_backward_matches_allocating_entry_point(ctx, oracle):
    P = Patterns(ctx, N.FIT4)
    go = oracle.Group(N.FIT4)
    for seed, n in ((11, 70000), (12, 30001)):
        say(f"no-alloc backward: FIT4, {n} rows at 30 empties, seed {seed}")
        pos = oracle.random_play(seed, n, 30)
        idx = P.extract(torch.from_numpy(pos.view(np.int64)).cuda(), capi.LAYOUT_ROW_MAJOR)
        oidx = go.indices(pos)
        assert np.array_equal(idx.cpu().numpy().view(np.uint32), oidx)
        # integer g: every sum is exact, so all three agree bit for bit
        gen = torch.Generator(device="cuda").manual_seed(seed)
        gi = torch.randint(-64, 65, (n,), device="cuda", generator=gen).float()
        a = _backward("cad_eval_backward_f32_noalloc", P, idx, gi)
        b = _backward("cad_eval_backward_f32", P, idx, gi)
        assert np.array_equal(a, b)
        assert np.array_equal(a, oracle.spmtv(oidx, gi.cpu().numpy(), P.cols))
        # real g: both sum in FP64 and round to float (the allocating one also rounds each chunk slice): within
        # 2^-22 of the sum of |g| over the column's entries
        gr = torch.randn(n, device="cuda", generator=gen)
        a = _backward("cad_eval_backward_f32_noalloc", P, idx, gr)
        b = _backward("cad_eval_backward_f32", P, idx, gr)
        bound = oracle.spmtv(oidx, np.abs(gr.cpu().numpy()).astype(np.float64), P.cols)
        assert np.all(np.isfinite(a))
        assert np.all(np.abs(a.astype(np.float64) - b) <= 2.0 ** -22 * bound + 1e-30)
        ref = oracle.spmtv(oidx, gr.cpu().numpy().astype(np.float64), P.cols)
        assert np.all(np.abs(a.astype(np.float64) - ref) <= 2.0 ** -24 * np.abs(ref) + 1e-12 * bound + 1e-30)


def test_noalloc_backward_empty_batch(ctx):
    P = Patterns(ctx, N.FIT4)
    idx = torch.zeros((0, P.k), dtype=torch.int32, device="cuda")
    g = torch.zeros(1, dtype=torch.float32, device="cuda")
    out = _backward("cad_eval_backward_f32_noalloc", P, idx, g)
    assert np.array_equal(out, np.zeros(P.cols, np.float32))
