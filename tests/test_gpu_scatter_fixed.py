"""The FP32 A^T y of a stored matrix accumulates its LDS windows in 64-bit
fixed point (exact integer sums of the terms rounded to a per-system power of
two): bit-reproducible from call to call, as accurate as FP32 allows against
the oracle's FP64 sums (Math/MatrixCSR.cpp:67-87 computes the same sums in
FP32), scaled per system in a batch, and falling back to the FP64 window for
a system with a non-finite value (NaN / inf reach only its columns)."""
import sys

import numpy as np
import pytest
import torch

from cassandra_amd import names as N
from cassandra_amd.fit import Patterns

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_close(z, idx, y64, cols, tol=2e-7):
    """Column by column against FP64 sums: the FP32 rounding of the result
    (tol of |A|^T |y|) plus the fixed-point rounding of each term (at most
    2^-40 of max |y| a term; the scale keeps about 42 bits of max |y|)."""
    import oracle
    exact = oracle.spmtv(idx, y64, cols)
    bound = oracle.spmtv(idx, np.abs(y64), cols)
    terms = oracle.spmtv(idx, np.ones_like(y64), cols)
    ymax = float(np.max(np.abs(y64))) if y64.size else 0.0
    err = np.abs(z.astype(np.float64) - exact)
    assert np.all(err <= tol * bound + terms * ymax * 2.0 ** -40), float(np.max(err - tol * bound))


@pytest.fixture(scope="module")
def mat(ctx, oracle):
    pos = oracle.random_play(95, 50000, 20)
    idx = oracle.Group(N.FIT8).indices(pos)
    A = Patterns(ctx, N.FIT8).matrix(torch.from_numpy(pos.view(np.int64)).cuda())
    return pos, idx, A


def test_fp32_scatter_reproducible_and_accurate(ctx, oracle, mat):
    pos, idx, A = mat
    say("fixed-point scatter: reproducible")
    gen = torch.Generator(device="cuda").manual_seed(11)
    for scale in (1.0, 1e-30, 3e25):
        y = (torch.randn(A.rows, dtype=torch.float32, device="cuda", generator=gen) * scale).float()
        z1, z2 = A.spmtv(y), A.spmtv(y)
        assert torch.equal(z1, z2)
        check_close(z1.cpu().numpy(), idx, y.cpu().numpy().astype(np.float64), A.cols)
    # under the other scatter schedules too (other chunkings, other bounds)
    y = torch.randn(A.rows, dtype=torch.float32, device="cuda", generator=gen)
    for sc in (dict(scatter_rpt=8, scatter_order=1, scatter_tiles_per_cu=4.0, scatter_partial_cap=0.25),
               dict(scatter_rpt=4, scatter_order=0, scatter_tiles_per_cu=1.0, scatter_partial_cap=1.0)):
        A.set_schedule(**sc)
        z = A.spmtv(y)
        check_close(z.cpu().numpy(), idx, y.cpu().numpy().astype(np.float64), A.cols)
        assert torch.equal(A.spmtv(y), z)
    A.set_schedule(scatter_rpt=4, scatter_order=0, scatter_tiles_per_cu=2.0, scatter_partial_cap=0.25)


def test_fp32_scatter_nonfinite_falls_back(ctx, oracle, mat):
    pos, idx, A = mat
    say("fixed-point scatter: NaN / inf")
    y = torch.randn(A.rows, dtype=torch.float32, device="cuda")
    for bad in (float("nan"), float("inf")):
        yb = y.clone()
        yb[123] = bad
        z = A.spmtv(yb).cpu().numpy()
        hit = np.zeros(A.cols, bool)
        hit[np.unique(idx[123])] = True
        assert np.all(~np.isfinite(z[hit]))
        y64 = yb.cpu().numpy().astype(np.float64)
        y64[123] = 0.0
        zc = z.copy()
        zc[hit] = 0.0
        ok = ~hit
        # the system falls back to the FP64 window: FP32 rounding only
        exact = oracle.spmtv(idx, y64, A.cols)
        bound = oracle.spmtv(idx, np.abs(y64), A.cols)
        assert np.all(np.abs(zc[ok] - exact[ok]) <= 2e-7 * bound[ok])


def test_fp32_scatter_scaled_per_system(ctx, oracle):
    say("fixed-point scatter: per-system scale")
    pos = oracle.random_play(96, 30000, 20)
    rb = [0, 9000, 21000, 30000]
    P = Patterns(ctx, N.FIT8)
    A = P.matrix_systems(torch.from_numpy(pos.view(np.int64)).cuda(), rb)
    cols = P.cols
    y = torch.randn(A.rows, dtype=torch.float32, device="cuda")
    scales = [1e-12, 1.0, 1e12]
    for s_, (a, b) in enumerate(zip(rb[:-1], rb[1:])):
        y[a:b] *= scales[s_]
    z = A.spmtv(y).cpu().numpy()
    g = oracle.Group(N.FIT8)
    for s_, (a, b) in enumerate(zip(rb[:-1], rb[1:])):
        idx = g.indices(pos[a:b])
        check_close(z[s_ * cols:(s_ + 1) * cols], idx, y[a:b].cpu().numpy().astype(np.float64), cols)
