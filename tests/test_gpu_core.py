"""GPU parity of the data generator, index extraction, CSR products and Jacobi
against the oracle (bit-exact for integer work and for the FP32 / FP64 row
sums, whose order matches the reference)."""
import sys

import numpy as np
import pytest
import torch

import cassandra_amd.capi as capi
from cassandra_amd import names as N
from cassandra_amd.fit import Matrix, Patterns

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(PO):
    return torch.from_numpy(np.ascontiguousarray(PO).view(np.int64)).cuda()


def positions(oracle, n, seed=1337):
    a = oracle.random_positions_mt(seed, n // 2)
    b = oracle.random_play(seed + 1, n - n // 2, 5, 58)
    return np.concatenate([a, b])


@pytest.mark.parametrize("seed,n,lo,hi", [(1, 4097, 24, 24), (7, 3000, 10, 59), (11, 1, 15, 45), (3, 500, 60, 60)])
def test_random_play_matches_oracle(ctx, oracle, seed, n, lo, hi):
    say("random_play", seed, n, lo, hi)
    got = ctx.random_play(seed, n, lo, hi).cpu().numpy().view(np.uint64)
    assert np.array_equal(got, oracle.random_play(seed, n, lo, hi))


SETS = {"fit4": N.FIT4, "fit8": N.FIT8, "fit12": N.FIT12, "mix": N.test_mix, "edax": N.edax,
        "cassandra": N.cassandra, "b6": [N.B6, N.D8]}


@pytest.mark.parametrize("name", list(SETS))
@pytest.mark.parametrize("n", [1, 1000, 4097])
def test_index_extract_matches_oracle(ctx, oracle, name, n):
    say("index_extract", name, n)
    masks = SETS[name]
    pos = positions(oracle, n)
    want = oracle.Group(masks).indices(pos)
    P = Patterns(ctx, masks)
    d = dev(pos)
    row = P.extract(d, capi.LAYOUT_ROW_MAJOR).cpu().numpy().view(np.uint32)
    assert np.array_equal(row, want)
    slot = P.extract(d, capi.LAYOUT_SLOT_MAJOR).cpu().numpy().view(np.uint32)
    assert np.array_equal(slot.T, want)
    A = P.matrix(d)
    assert A.index_bytes == (4 if name == "b6" else 2)
    assert np.array_equal(A.download(), want)


def test_known_answer_matrix(ctx):
    # MathTest/MatrixCSR.cpp:28-54
    say("kat")
    A = Matrix.from_host(ctx, np.array([[0, 1], [1, 2], [3, 3]], np.uint32), 4)
    assert A.spmv(torch.tensor([1, 2, 3, 4], dtype=torch.float32, device="cuda")).tolist() == [3, 5, 8]
    assert A.spmtv(torch.tensor([1, 2, 3], dtype=torch.float32, device="cuda")).tolist() == [1, 3, 2, 6]
    assert A.spmv(torch.tensor([1, 2, 3, 4], dtype=torch.float64, device="cuda")).tolist() == [3, 5, 8]
    assert A.spmtv(torch.tensor([1, 2, 3], dtype=torch.float64, device="cuda")).tolist() == [1, 3, 2, 6]
    assert A.jacobi().tolist() == [1, 0.5, 1, 0.25]
    with pytest.raises(ValueError):
        A.spmv(torch.zeros(3, dtype=torch.float32, device="cuda"))
    with pytest.raises(capi.CadError):
        Matrix.from_host(ctx, np.array([[0, 5]], np.uint32), 4)  # column out of range


@pytest.fixture(scope="module")
def fit8(ctx, oracle):
    pos = oracle.random_play(1002, 20000, 24)
    g = oracle.Group(N.FIT8)
    P = Patterns(ctx, N.FIT8)
    return g.indices(pos), g.cols, P.matrix(dev(pos))


def test_spmv_bit_exact(ctx, oracle, fit8):
    idx, cols, A = fit8
    say("spmv", idx.shape)
    rng = np.random.default_rng(1)
    x = rng.normal(size=cols)
    for dt, tdt in ((np.float32, torch.float32), (np.float64, torch.float64)):
        xs = x.astype(dt)
        y = A.spmv(torch.from_numpy(xs).cuda()).cpu().numpy()
        assert np.array_equal(y, oracle.spmv(idx, xs))


def test_spmtv(ctx, oracle, fit8):
    idx, cols, A = fit8
    say("spmtv", idx.shape)
    rng = np.random.default_rng(2)
    yi = rng.integers(-64, 65, size=idx.shape[0]).astype(np.float32)
    z = A.spmtv(torch.from_numpy(yi).cuda()).cpu().numpy()
    assert np.array_equal(z, oracle.spmtv(idx, yi, cols))  # integer sums: exact in any order
    y = rng.normal(size=idx.shape[0])
    z64 = A.spmtv(torch.from_numpy(y).cuda()).cpu().numpy()
    ref = oracle.spmtv(idx, y, cols)
    absy = oracle.spmtv(idx, np.abs(y), cols)
    assert np.all(np.abs(z64 - ref) <= 1e-13 * absy + 1e-300)
    z32 = A.spmtv(torch.from_numpy(y.astype(np.float32)).cuda()).cpu().numpy().astype(np.float64)
    ref32 = oracle.spmtv(idx, y.astype(np.float32).astype(np.float64), cols)
    assert np.all(np.abs(z32 - ref32) <= 2e-6 * absy + 1e-30)


def test_jacobi_exact(ctx, oracle, fit8):
    idx, cols, A = fit8
    say("jacobi")
    assert np.array_equal(A.jacobi().cpu().numpy(), oracle.jacobi(idx, cols))


def test_generic_wide_matrix(ctx, oracle):
    # columns spread over 300k: u32 storage and several LDS windows per slot
    say("generic wide")
    rng = np.random.default_rng(3)
    rows, k, cols = 30000, 5, 300000
    idx = rng.integers(0, cols, size=(rows, k)).astype(np.uint32)
    idx[::3, 2] = idx[::3, 1]
    A = Matrix.from_host(ctx, idx, cols)
    assert A.index_bytes == 4
    x = rng.normal(size=cols).astype(np.float32)
    assert np.array_equal(A.spmv(torch.from_numpy(x).cuda()).cpu().numpy(), oracle.spmv(idx, x))
    yi = rng.integers(-9, 10, size=rows).astype(np.float32)
    assert np.array_equal(A.spmtv(torch.from_numpy(yi).cuda()).cpu().numpy(), oracle.spmtv(idx, yi, cols))
    assert np.array_equal(A.jacobi().cpu().numpy(), oracle.jacobi(idx, cols))


def test_targets_bit_exact(ctx, oracle, fit8):
    idx, cols, A = fit8
    say("targets")
    xt, b = A.make_targets(2002, 3002)
    oxt, ob = oracle.make_targets(idx, cols, 2002, 3002)
    assert np.array_equal(xt.cpu().numpy(), oxt)
    assert np.array_equal(b.cpu().numpy(), ob)


def test_full_size_properties(ctx):
    # 5M rows at 24 empties with the 8-pattern set: size-independent checks
    say("full size properties")
    P = Patterns(ctx, N.FIT8)
    pos = ctx.random_play(1002, 5_000_000, 24)
    A = P.matrix(pos)
    ones = torch.ones(A.cols, dtype=torch.float32, device="cuda")
    y = A.spmv(ones)
    assert torch.all(y == A.k)  # every row holds k entries
    z = A.spmtv(torch.ones(A.rows, dtype=torch.float32, device="cuda"))
    assert float(z.double().sum()) == float(A.rows * A.k)  # integer counts, exact in float
    # per pattern block the column counts add up to rows * order
    for i in range(P.npatterns):
        info = P.pattern_info(i)
        blk = z[info["offset"]:info["offset"] + info["size"]].double().sum()
        assert float(blk) == float(A.rows * info["order"])
    # rows of the transposed view: every row's entries are inside its slot blocks
    idx = P.extract(pos[:100000], capi.LAYOUT_ROW_MAJOR).cpu().numpy()
    lo = np.repeat([P.pattern_info(i)["offset"] for i in range(P.npatterns)],
                   [P.pattern_info(i)["order"] for i in range(P.npatterns)])
    hi = lo + np.repeat([P.pattern_info(i)["size"] for i in range(P.npatterns)],
                        [P.pattern_info(i)["order"] for i in range(P.npatterns)])
    assert np.all((idx >= lo) & (idx < hi))


def test_full_size_products_bit_exact(ctx, oracle):
    # the bench's 5M x 139,644 matrix (8 patterns, u16 entries) with the default
    # schedule: A x in FP32 and FP64 bit-identical to the oracle's slot-order
    # sums, A^T y exact on integer y
    say("full size products")
    P = Patterns(ctx, N.FIT8)
    pos = ctx.random_play(1002, 5_000_000, 24)
    A = P.matrix(pos)
    idx = A.download()
    gen = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn(A.cols, dtype=torch.float32, device="cuda", generator=gen)
    assert np.array_equal(A.spmv(x).cpu().numpy(), oracle.spmv(idx, x.cpu().numpy()))
    x64 = torch.randn(A.cols, dtype=torch.float64, device="cuda", generator=gen)
    assert np.array_equal(A.spmv(x64).cpu().numpy(), oracle.spmv(idx, x64.cpu().numpy()))
    y = torch.randint(-64, 65, (A.rows,), device="cuda", generator=gen).float()
    assert np.array_equal(A.spmtv(y).cpu().numpy(), oracle.spmtv(idx, y.cpu().numpy(), A.cols))


def test_reference_golden_vectors(ctx):
    # outputs of the reference's compiled Math sources (tests/golden/make_reference_math.py)
    import os

    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_math.npz"),
                allow_pickle=False)
    idx, cols = G["idx"], int(G["cols"])
    P = Patterns(ctx, [N.D8, N.L1, N.D6])
    for A in (P.matrix(torch.from_numpy(G["positions"].view(np.int64)).cuda()), Matrix.from_host(ctx, idx, cols)):
        assert np.array_equal(A.download(), idx)
        assert np.array_equal(A.spmv(torch.from_numpy(G["x"]).cuda()).cpu().numpy(), G["ax"])
        assert np.array_equal(A.spmtv(torch.from_numpy(G["y"]).cuda()).cpu().numpy(), G["ya"])
        assert np.array_equal(A.jacobi().cpu().numpy(), G["jacobi"])
