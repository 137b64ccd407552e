"""Every product schedule computes the same products: A x bit-identical to the
oracle's slot-order sums (Math/MatrixCSR.cpp:49-65) in FP32 and FP64, A^T y
exact on integer data (Math/MatrixCSR.cpp:67-87), whichever tile shape, block
size, scatter order and chunking is set -- the autotuner only ever picks among
these."""
import itertools
import sys

import numpy as np
import pytest
import torch

import cassandra_amd.capi as capi
from cassandra_amd import names as N
from cassandra_amd.fit import Patterns

pytestmark = pytest.mark.gpu

# (rows per thread, threads per block, sub-tiles per staging)
GATHER32 = [(8, 1024, 1), (4, 1024, 1), (2, 1024, 1), (8, 512, 1), (4, 512, 1), (2, 512, 1), (0, 1024, 1),
            (2, 1024, 2), (2, 1024, 3)]
GATHER64 = [(2, 1024, 1), (1, 1024, 1), (2, 512, 1), (1, 512, 1), (0, 1024, 1), (1, 1024, 2), (1, 1024, 3)]
# (rows per step, order, tiles per CU, partial cap, balance)
SCATTER = [(4, 0, 2.0, 0.25, 1.0), (8, 1, 4.0, 0.25, 1.0), (4, 1, 1.0, 1.0, 1.0), (8, 0, 4.0, 0.1, 0.25),
           (4, 0, 1.0, 1.0, 0.5)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fit8(ctx, oracle):
    # 30011 rows: ragged tiles and chunks at every block size
    pos = oracle.random_play(93, 30011, 20)
    idx = oracle.Group(N.FIT8).indices(pos)
    A = Patterns(ctx, N.FIT8).matrix(torch.from_numpy(pos.view(np.int64)).cuda())
    return idx, A


def schedules():
    n = max(len(GATHER32), len(GATHER64), len(SCATTER))
    cyc = lambda L: itertools.islice(itertools.cycle(L), n)  # noqa: E731
    return list(zip(cyc(GATHER32), cyc(GATHER64), cyc(SCATTER)))


@pytest.mark.parametrize("g32,g64,sc", schedules())
def test_schedule_products(ctx, oracle, fit8, g32, g64, sc):
    idx, A = fit8
    say("schedule", g32, g64, sc)
    A.set_schedule(gather_f32_rpt=g32[0], gather_f32_block=g32[1], gather_f32_sub=g32[2], gather_f64_rpt=g64[0],
                   gather_f64_block=g64[1], gather_f64_sub=g64[2], scatter_rpt=sc[0], scatter_order=sc[1],
                   scatter_tiles_per_cu=sc[2], scatter_partial_cap=sc[3], scatter_balance=sc[4])
    got = A.schedule()
    assert (got["gather_f32_rpt"], got["gather_f32_block"], got["gather_f32_sub"]) == g32
    assert (got["gather_f64_rpt"], got["gather_f64_block"], got["gather_f64_sub"]) == g64
    gen = torch.Generator(device="cuda").manual_seed(7)
    x = torch.randn(A.cols, dtype=torch.float32, device="cuda", generator=gen)
    assert np.array_equal(A.spmv(x).cpu().numpy(), oracle.spmv(idx, x.cpu().numpy()))
    x64 = torch.randn(A.cols, dtype=torch.float64, device="cuda", generator=gen)
    assert np.array_equal(A.spmv(x64).cpu().numpy(), oracle.spmv(idx, x64.cpu().numpy()))
    g = torch.randint(-64, 65, (A.rows,), device="cuda", generator=gen).float()
    assert np.array_equal(A.spmtv(g).cpu().numpy(), oracle.spmtv(idx, g.cpu().numpy(), A.cols))
    assert np.array_equal(A.spmtv(g.double()).cpu().numpy(), oracle.spmtv(idx, g.cpu().numpy(), A.cols))


def test_schedule_solver_and_errors(ctx, oracle, fit8):
    idx, A = fit8
    say("schedule solver + errors")
    _, b = A.make_targets(2007, 3007)
    A.set_schedule(gather_f32_rpt=2, gather_f32_block=1024, gather_f32_sub=3, gather_f64_rpt=1,
                   gather_f64_block=1024, gather_f64_sub=2)
    xm, rep = A.solve(b, precision=capi.PREC_MIXED, tol=1e-10)
    assert rep["converged"] == 1
    assert oracle.rel_normal_residual(idx, A.cols, b.cpu().numpy(), xm.cpu().numpy())[0] <= 1.5e-10
    A.set_schedule(gather_f32_rpt=4, gather_f32_block=512, gather_f32_sub=1, gather_f64_rpt=2, gather_f64_block=512,
                   gather_f64_sub=1)
    xm, rep = A.solve(b, precision=capi.PREC_MIXED, tol=1e-10)
    assert rep["converged"] == 1
    assert oracle.rel_normal_residual(idx, A.cols, b.cpu().numpy(), xm.cpu().numpy())[0] <= 1.5e-10
    # both ways an LSQR iteration forms beta: in the window tiles (4 rows a
    # step), or in the gather's finish launch (8 rows a step, whose window
    # kernel has no registers left for that prologue)
    for srpt, order in ((8, 1), (4, 0)):
        A.set_schedule(scatter_rpt=srpt, scatter_order=order)
        xm, rep = A.solve(b, precision=capi.PREC_MIXED, tol=1e-10)
        assert rep["converged"] == 1, srpt
        assert oracle.rel_normal_residual(idx, A.cols, b.cpu().numpy(), xm.cpu().numpy())[0] <= 1.5e-10, srpt
    before = A.schedule()
    for bad in (dict(gather_f32_rpt=3), dict(gather_f32_block=256), dict(gather_f64_rpt=4),
                dict(gather_f32_rpt=8, gather_f32_sub=2), dict(gather_f32_rpt=2, gather_f32_sub=4),
                dict(gather_f32_rpt=2, gather_f32_block=512, gather_f32_sub=2),
                dict(scatter_rpt=2), dict(scatter_order=2), dict(scatter_tiles_per_cu=0.0),
                dict(scatter_partial_cap=1.5), dict(scatter_balance=1.5)):
        with pytest.raises(capi.CadError) as e:
            A.set_schedule(**bad)
        assert e.value.code == capi.ERR_ARGUMENT
        assert A.schedule() == before


def test_matrix_free_keeps_full_blocks(ctx, oracle):
    say("schedule matrix-free")
    pos = oracle.random_play(94, 5000, 20)
    F = Patterns(ctx, N.FIT8).matrix_free(torch.from_numpy(pos.view(np.int64)).cuda())
    with pytest.raises(capi.CadError):
        F.set_schedule(gather_f32_block=512)
    with pytest.raises(capi.CadError):
        F.set_schedule(gather_f32_rpt=8)
    with pytest.raises(capi.CadError):
        F.set_schedule(gather_f32_rpt=2, gather_f32_sub=2)
    F.set_schedule(gather_f32_rpt=2, scatter_rpt=4)
    idx = oracle.Group(N.FIT8).indices(pos)
    x = torch.randn(F.cols, dtype=torch.float32, device="cuda")
    assert np.array_equal(F.spmv(x).cpu().numpy(), oracle.spmv(idx, x.cpu().numpy()))
