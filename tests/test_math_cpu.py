"""CSR products, Jacobi and solvers of the oracle, pinned by the reference's
MathTest known answers, by the reference's own compiled Math sources
(oracle/_ref, where they were built), by dense brute force and by scipy's
LSQR / LSMR / CG iterates."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O

# MathTest/MatrixCSR.cpp:3-17 (row 2 holds column 3 twice: value 2)
KAT_IDX = np.array([[0, 1], [1, 2], [3, 3]], dtype=np.uint32)


def dense(idx, cols):
    A = np.zeros((idx.shape[0], cols))
    for r in range(idx.shape[0]):
        for c in idx[r]:
            A[r, c] += 1
    return A


def csr(idx, cols):
    rows, k = idx.shape
    return sp.csr_matrix((np.ones(rows * k), idx.reshape(-1).astype(np.int64),
                          np.arange(0, rows * k + 1, k)), shape=(rows, cols))


def small_pattern_problem(n=3000, masks=None, seed=31, E=20):
    from cassandra_amd import names as N

    g = O.Group(masks or N.FIT4)
    pos = O.random_play(seed, n, E)
    idx = g.indices(pos)
    xt, b = O.make_targets(idx, g.cols, 2000 + seed, 3000 + seed)
    return idx, g.cols, b


def test_matrix_csr_known_answers():
    # MathTest/MatrixCSR.cpp:28-54
    assert np.array_equal(O.spmv(KAT_IDX, np.array([1, 2, 3, 4], np.float32)), [3, 5, 8])
    assert np.array_equal(O.spmtv(KAT_IDX, np.array([1, 2, 3], np.float32), 4), [1, 3, 2, 6])
    assert np.array_equal(O.jacobi(KAT_IDX, 4), np.array([1, 0.5, 1, 0.25], np.float32))


def test_reference_math_known_answers():
    R = O.ref()
    if R is None:
        pytest.skip("reference Math sources were not built (oracle/_ref)")
    import ctypes as C

    m = R.ref_csr_create(KAT_IDX.ctypes.data_as(O.u32p), 3, 2, 4)
    try:
        y = np.zeros(3, np.float32)
        R.ref_spmv(m, np.array([1, 2, 3, 4], np.float32).ctypes.data_as(O.f32p), 4, y.ctypes.data_as(O.f32p))
        assert np.array_equal(y, [3, 5, 8])
        z = np.zeros(4, np.float32)
        R.ref_spmtv(m, np.array([1, 2, 3], np.float32).ctypes.data_as(O.f32p), 3, z.ctypes.data_as(O.f32p))
        assert np.array_equal(z, [1, 3, 2, 6])
        d = np.zeros(4, np.float32)
        R.ref_jacobi(m, d.ctypes.data_as(O.f32p))
        assert np.array_equal(d, [1, 0.5, 1, 0.25])
        # size mismatch throws std::runtime_error (MatrixCSR.cpp:51-52)
        assert R.ref_spmv(m, np.zeros(3, np.float32).ctypes.data_as(O.f32p), 3, y.ctypes.data_as(O.f32p)) == -1
    finally:
        R.ref_csr_free(m)
    # MathTest/Vector.cpp:112-147
    two, three = np.full(3, 2, np.float32), np.full(3, 3, np.float32)
    assert R.ref_dot(two.ctypes.data_as(O.f32p), three.ctypes.data_as(O.f32p), 3) == 18
    assert R.ref_l1(np.full(3, -2, np.float32).ctypes.data_as(O.f32p), 3) == 6
    assert R.ref_norm(np.array([3, 4], np.float32).ctypes.data_as(O.f32p), 2) == 5
    assert R.ref_sum(two.ctypes.data_as(O.f32p), 3) == 6
    assert R.ref_dot_size_mismatch_throws() == 1
    del C


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_products_against_dense(seed):
    rng = np.random.default_rng(seed)
    rows, k, cols = 400, 7, 50
    idx = rng.integers(0, cols, size=(rows, k)).astype(np.uint32)
    idx[::5, 1] = idx[::5, 0]  # repeated columns inside rows
    A = dense(idx, cols)
    x = rng.normal(size=cols)
    y = rng.normal(size=rows)
    assert np.allclose(O.spmv(idx, x), A @ x, rtol=1e-14, atol=1e-12)
    assert np.allclose(O.spmtv(idx, y, cols), A.T @ y, rtol=1e-13, atol=1e-12)
    assert np.allclose(O.spmv(idx, x.astype(np.float32)), A @ x, rtol=1e-5, atol=1e-5)
    diag = (A * A).sum(axis=0)
    with np.errstate(divide="ignore"):
        assert np.array_equal(O.jacobi(idx, cols), (1.0 / diag).astype(np.float32))


def test_reference_products_match_oracle():
    R = O.ref()
    if R is None:
        pytest.skip("reference Math sources were not built (oracle/_ref)")
    idx, cols, b = small_pattern_problem(5000)
    rows, k = idx.shape
    m = R.ref_csr_create(idx.ctypes.data_as(O.u32p), rows, k, cols)
    try:
        rng = np.random.default_rng(7)
        x = rng.normal(size=cols).astype(np.float32)
        y = np.zeros(rows, np.float32)
        R.ref_spmv(m, x.ctypes.data_as(O.f32p), cols, y.ctypes.data_as(O.f32p))
        assert np.array_equal(y, O.spmv(idx, x))  # same terms, same order
        yi = rng.integers(-64, 65, size=rows).astype(np.float32)  # integer: exact in any order
        z = np.zeros(cols, np.float32)
        R.ref_spmtv(m, yi.ctypes.data_as(O.f32p), rows, z.ctypes.data_as(O.f32p))
        assert np.array_equal(z, O.spmtv(idx, yi, cols))
        d = np.zeros(cols, np.float32)
        R.ref_jacobi(m, d.ctypes.data_as(O.f32p))
        assert np.array_equal(d, O.jacobi(idx, cols))
        # FitWeights' PCG in float: the reference against the oracle's float
        # restatement (reduction order differs: tolerance on predictions)
        xr = np.zeros(cols, np.float32)
        hr = np.zeros(10, np.float32)
        R.ref_fit_pcg(m, b.astype(np.float32).ctypes.data_as(O.f32p), rows, 10, xr.ctypes.data_as(O.f32p),
                      hr.ctypes.data_as(O.f32p))
        xo, ho = O.pcg(idx, cols, b, 10, dtype=np.float32)
        x64, _ = O.pcg(idx, cols, b, 10, dtype=np.float64)
        p_ref, p_or, p64 = O.spmv(idx, xr.astype(np.float64)), O.spmv(idx, xo.astype(np.float64)), O.spmv(idx, x64)
        scale = np.linalg.norm(p64)
        assert np.linalg.norm(p_ref - p64) / scale < 1e-4
        assert np.linalg.norm(p_or - p64) / scale < 1e-4
        assert np.allclose(hr, ho, rtol=1e-3)
        # LSQR in float (Solver.h:128-182): the reference against the oracle's
        # float restatement; both drift from the FP64 iterate by the same
        # float rounding (up to ~6e-3 around iteration 20 on this problem), and
        # their OpenMP float reductions combine thread sums in arrival order,
        # so the two differ from run to run by up to ~3e-4.
        for it in (5, 20, 40):
            xl = np.zeros(cols, np.float32)
            R.ref_lsqr(m, b.astype(np.float32).ctypes.data_as(O.f32p), rows, it, xl.ctypes.data_as(O.f32p), None)
            xo, _, _ = O.lsqr(idx, cols, b, it, dtype=np.float32)
            x64, _, _ = O.lsqr(idx, cols, b, it)
            p64 = O.spmv(idx, x64)
            p_ref, p_or = O.spmv(idx, xl.astype(np.float64)), O.spmv(idx, xo.astype(np.float64))
            assert np.linalg.norm(p_ref - p_or) / np.linalg.norm(p64) < 2e-3
            assert np.linalg.norm(p_ref - p64) / np.linalg.norm(p64) < 2e-2
    finally:
        R.ref_csr_free(m)


# rounding differences between two FP64 LSQR codes grow transiently on this
# ill-conditioned matrix (peak ~1e-8 near iteration 20, back to ~1e-13 by
# 60); a wrong term gives O(1) differences
@pytest.mark.parametrize("iters,tol", [(1, 1e-12), (2, 1e-12), (5, 1e-12), (12, 1e-10), (20, 1e-6), (60, 1e-10)])
def test_lsqr_iterates_match_scipy(iters, tol):
    idx, cols, b = small_pattern_problem(2000)
    A = csr(idx, cols)
    x, phibar, nres = O.lsqr(idx, cols, b, iters)
    ref = spla.lsqr(A, b, damp=0.0, atol=0, btol=0, conlim=0, iter_lim=iters)
    assert ref[2] == iters
    assert np.linalg.norm(x - ref[0]) <= tol * np.linalg.norm(ref[0])
    # Residuum() = phibar equals |b - A x_k| (LSQR's r-norm recurrence)
    assert abs(phibar[-1] - np.linalg.norm(b - A @ x)) <= 1e-8 * np.linalg.norm(b)
    # |A^T r_k| recurrence against the explicit value
    assert abs(nres[-1] - np.linalg.norm(A.T @ (b - A @ x))) <= 1e-6 * np.linalg.norm(A.T @ b)


def test_lsqr_column_scaling_is_lsqr_on_scaled_matrix():
    idx, cols, b = small_pattern_problem(2000)
    A = csr(idx, cols)
    cnt = np.asarray((A.multiply(A)).sum(axis=0)).ravel()
    D = np.where(cnt > 0, 1.0 / np.sqrt(np.maximum(cnt, 1)), 1.0)
    x, _, _ = O.lsqr(idx, cols, b, 25, colscale=D)
    ref = spla.lsqr(A @ sp.diags(D), b, atol=0, btol=0, conlim=0, iter_lim=25)
    assert np.linalg.norm(x - D * ref[0]) <= 1e-9 * np.linalg.norm(D * ref[0])


@pytest.mark.parametrize("iters", [1, 3, 10, 40])
def test_lsmr_iterates_match_scipy(iters):
    idx, cols, b = small_pattern_problem(2000)
    A = csr(idx, cols)
    x = O.lsmr(idx, cols, b, iters)
    ref = spla.lsmr(A, b, damp=0.0, atol=0, btol=0, conlim=0, maxiter=iters)
    assert np.linalg.norm(x - ref[0]) <= 1e-9 * np.linalg.norm(ref[0])


@pytest.mark.parametrize("iters", [1, 4, 15])
def test_pcg_and_cg_match_scipy(iters):
    idx, cols, b = small_pattern_problem(2000)
    A = csr(idx, cols)
    N = (A.T @ A).tocsr()
    rhs = A.T @ b
    d = O.jacobi(idx, cols).astype(np.float64)
    d[~np.isfinite(d) | (d == 0)] = 1.0
    M = sp.diags(d)
    ref = spla.cg(N, rhs, x0=np.zeros(cols), rtol=0, atol=0, maxiter=iters, M=M)[0]
    x, _ = O.pcg(idx, cols, b, iters)
    assert np.linalg.norm(x - ref) <= 1e-8 * np.linalg.norm(ref)
    ref = spla.cg(N, rhs, x0=np.zeros(cols), rtol=0, atol=0, maxiter=iters)[0]
    x, _ = O.cg(idx, cols, b, iters)
    assert np.linalg.norm(x - ref) <= 1e-8 * np.linalg.norm(ref)


def test_noise_free_full_rank_recovery():
    # A with full column rank and b = A x_true: LSQR recovers x_true
    rng = np.random.default_rng(4)
    rows, k, cols = 600, 5, 40
    idx = rng.integers(0, cols, size=(rows, k)).astype(np.uint32)
    A = dense(idx, cols)
    assert np.linalg.matrix_rank(A) == cols
    xt = rng.normal(size=cols)
    x, _, _ = O.lsqr(idx, cols, A @ xt, 200)
    assert np.allclose(x, xt, rtol=1e-8, atol=1e-8)


def test_cg_finite_termination():
    # CG on an SPD n x n system converges in at most n steps (exact arithmetic)
    rng = np.random.default_rng(8)
    rows, k, cols = 300, 4, 12
    idx = rng.integers(0, cols, size=(rows, k)).astype(np.uint32)
    A = dense(idx, cols)
    b = rng.normal(size=rows)
    x, _ = O.cg(idx, cols, b, cols)
    xs = np.linalg.lstsq(A, b, rcond=None)[0]
    assert np.allclose(A @ x, A @ xs, rtol=1e-8, atol=1e-8)


def test_rel_normal_residual_definition():
    idx, cols, b = small_pattern_problem(1500)
    A = csr(idx, cols)
    x = np.random.default_rng(2).normal(size=cols)
    v, rn, an = O.rel_normal_residual(idx, cols, b, x)
    r = b - A @ x
    assert np.isclose(v, np.linalg.norm(A.T @ r) / np.linalg.norm(A.T @ b), rtol=1e-12)
    assert np.isclose(rn, np.linalg.norm(r), rtol=1e-12)


def test_targets_construction():
    idx, cols, b = small_pattern_problem(3000)
    assert np.all(b % 2 == 0) and b.min() >= -64 and b.max() <= 64
    xt, b2 = O.make_targets(idx, cols, 2031, 3031)
    assert np.array_equal(b, b2)
    assert abs(xt.std() - 1.5) < 0.02 and abs(xt.mean()) < 0.02
    # the noise is what remains after the exact prediction (before rounding)
    resid = b - O.spmv(idx, xt)
    assert 3.5 < resid.std() < 4.7


def golden():
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_math.npz")
    return np.load(path, allow_pickle=False)


def test_oracle_against_reference_golden_vectors():
    # outputs of the reference's compiled Math sources (tests/golden/make_reference_math.py)
    from cassandra_amd import names as N

    G = golden()
    idx, cols = G["idx"], int(G["cols"])
    assert np.array_equal(O.Group([N.D8, N.L1, N.D6]).indices(G["positions"]), idx)
    assert np.array_equal(O.spmv(idx, G["x"]), G["ax"])
    assert np.array_equal(O.spmtv(idx, G["y"], cols), G["ya"])
    assert np.array_equal(O.jacobi(idx, cols), G["jacobi"])
