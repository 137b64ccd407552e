"""ctypes wrapper of the CPU oracle (liboracle.so) and of the compiled reference
Math sources (_ref/libref_math.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
The product package (cassandra_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i64 = C.c_int64


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        for n in ("or_flip_vertical", "or_flip_horizontal", "or_flip_diagonal", "or_flip_codiagonal"):
            getattr(L, n).restype = C.c_uint64
            getattr(L, n).argtypes = [C.c_uint64]
        for n in ("or_pext", "or_pdep", "or_hw_pext", "or_hw_pdep"):
            getattr(L, n).restype = C.c_uint64
            getattr(L, n).argtypes = [C.c_uint64, C.c_uint64]
        L.or_fast_index.restype = C.c_uint32
        L.or_fast_index.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.or_configurations.restype = i64
        L.or_configurations.argtypes = [C.c_uint64, u64p, i64]
        L.or_group_create.restype = C.c_void_p
        L.or_group_create.argtypes = [u64p, C.c_int]
        L.or_group_free.argtypes = [C.c_void_p]
        L.or_group_k.restype = C.c_int
        L.or_group_k.argtypes = [C.c_void_p]
        L.or_group_cols.restype = i64
        L.or_group_cols.argtypes = [C.c_void_p]
        L.or_group_npatterns.restype = C.c_int
        L.or_group_npatterns.argtypes = [C.c_void_p]
        L.or_group_pattern_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), u64p,
                                            C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        L.or_group_variations.argtypes = [C.c_void_p, u64p]
        L.or_group_indices.argtypes = [C.c_void_p, u64p, i64, u32p]
        L.or_group_pattern_index.restype = C.c_uint32
        L.or_group_pattern_index.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64]
        L.or_spmv_f32.argtypes = [u32p, i64, C.c_int, f32p, f32p]
        L.or_spmv_f64.argtypes = [u32p, i64, C.c_int, f64p, f64p]
        L.or_spmtv_f32.argtypes = [u32p, i64, C.c_int, i64, f32p, f32p, C.c_int]
        L.or_spmtv_f64.argtypes = [u32p, i64, C.c_int, i64, f64p, f64p]
        L.or_jacobi_f32.argtypes = [u32p, i64, C.c_int, i64, f32p]
        L.or_pcg_f32.argtypes = [u32p, i64, C.c_int, i64, f32p, C.c_int, f32p, f32p]
        L.or_pcg_f64.argtypes = [u32p, i64, C.c_int, i64, f64p, C.c_int, f64p, f64p]
        L.or_cg_f64.argtypes = [u32p, i64, C.c_int, i64, f64p, C.c_int, f64p, f64p]
        L.or_lsqr_f32.argtypes = [u32p, i64, C.c_int, i64, f32p, C.c_int, f32p, f32p]
        L.or_lsqr_f64.argtypes = [u32p, i64, C.c_int, i64, f64p, C.c_int, f64p, f64p, f64p, f64p]
        L.or_lsmr_f64.argtypes = [u32p, i64, C.c_int, i64, f64p, C.c_int, f64p]
        L.or_rel_normal_residual.restype = C.c_double
        L.or_rel_normal_residual.argtypes = [u32p, i64, C.c_int, i64, f64p, f64p, f64p, f64p]
        L.or_possible_moves.restype = C.c_uint64
        L.or_possible_moves.argtypes = [C.c_uint64, C.c_uint64]
        L.or_flips.restype = C.c_uint64
        L.or_flips.argtypes = [C.c_uint64, C.c_uint64, C.c_int]
        L.or_perft.restype = C.c_uint64
        L.or_perft.argtypes = [C.c_int]
        L.or_random_positions_mt.argtypes = [C.c_uint32, i64, u64p]
        L.or_random_positions_mt_dist.argtypes = [C.c_uint32, i64, u64p]
        L.or_random_play.argtypes = [C.c_uint64, i64, C.c_int, C.c_int, u64p]
        L.or_make_targets.argtypes = [u32p, i64, C.c_int, i64, C.c_uint64, C.c_uint64,
                                      C.c_double, C.c_double, f64p, f64p]
        L.or_expand_weights.restype = i64
        L.or_expand_weights.argtypes = [C.c_void_p, f32p, f32p]
        L.or_eval_expanded.argtypes = [C.c_void_p, f32p, u64p, i64, f32p]
        L.or_compact_weights.argtypes = [C.c_void_p, f32p, f32p]
        _lib = L
    return _lib


def ref():
    """The reference's own Math sources (MatrixCSR, Vector, Solver.h), or None."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libref_math.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_csr_create.restype = C.c_void_p
        R.ref_csr_create.argtypes = [u32p, i64, C.c_int, i64]
        R.ref_csr_free.argtypes = [C.c_void_p]
        R.ref_spmv.restype = C.c_int
        R.ref_spmv.argtypes = [C.c_void_p, f32p, i64, f32p]
        R.ref_spmtv.restype = C.c_int
        R.ref_spmtv.argtypes = [C.c_void_p, f32p, i64, f32p]
        R.ref_jacobi.argtypes = [C.c_void_p, f32p]
        R.ref_fit_pcg.argtypes = [C.c_void_p, f32p, i64, C.c_int, f32p, f32p]
        R.ref_lsqr.argtypes = [C.c_void_p, f32p, i64, C.c_int, f32p, f32p]
        R.ref_lsmr.argtypes = [C.c_void_p, f32p, i64, C.c_int, f32p]
        for n in ("ref_dot",):
            getattr(R, n).restype = C.c_float
            getattr(R, n).argtypes = [f32p, f32p, i64]
        for n in ("ref_norm", "ref_sum", "ref_l1"):
            getattr(R, n).restype = C.c_float
            getattr(R, n).argtypes = [f32p, i64]
        R.ref_dot_size_mismatch_throws.restype = C.c_int
        _ref = R
    return _ref


# ------------------------------------------------------------------ helpers --
class Group:
    """CreateIndexer/GroupIndexer restatement (Pattern/Indexer.cpp:319-422)."""

    def __init__(self, masks):
        self.masks = np.ascontiguousarray(np.array(masks, dtype=np.uint64))
        self.h = lib().or_group_create(_p(self.masks, u64p), len(self.masks))
        self.k = lib().or_group_k(self.h)
        self.cols = lib().or_group_cols(self.h)
        self.npat = lib().or_group_npatterns(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_group_free(self.h)
            self.h = None

    def info(self, i):
        cls, mask, size, order = C.c_int(), C.c_uint64(), C.c_int64(), C.c_int()
        lib().or_group_pattern_info(self.h, i, C.byref(cls), C.byref(mask), C.byref(size), C.byref(order))
        return dict(cls=cls.value, mask=mask.value, size=size.value, order=order.value)

    def variations(self):
        out = np.zeros(self.k, dtype=np.uint64)
        lib().or_group_variations(self.h, _p(out, u64p))
        return out

    def indices(self, PO):
        PO = np.ascontiguousarray(PO, dtype=np.uint64).reshape(-1, 2)
        out = np.zeros((PO.shape[0], self.k), dtype=np.uint32)
        lib().or_group_indices(self.h, _p(PO, u64p), PO.shape[0], _p(out, u32p))
        return out

    def pattern_index(self, i, P, O):
        return lib().or_group_pattern_index(self.h, i, P, O)

    def expand(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        n = lib().or_expand_weights(self.h, None, None)
        out = np.zeros(n, dtype=np.float32)
        lib().or_expand_weights(self.h, _p(w, f32p), _p(out, f32p))
        return out

    def eval_expanded(self, tables, PO):
        PO = np.ascontiguousarray(PO, dtype=np.uint64).reshape(-1, 2)
        out = np.zeros(PO.shape[0], dtype=np.float32)
        lib().or_eval_expanded(self.h, _p(tables, f32p), _p(PO, u64p), PO.shape[0], _p(out, f32p))
        return out

    def compact(self, tables):
        out = np.zeros(self.cols, dtype=np.float32)
        lib().or_compact_weights(self.h, _p(np.ascontiguousarray(tables, dtype=np.float32), f32p),
                                 _p(out, f32p))
        return out


def configurations(mask):
    n = lib().or_configurations(mask, None, 0)
    out = np.zeros(2 * n, dtype=np.uint64)
    lib().or_configurations(mask, _p(out, u64p), n)
    return out.reshape(-1, 2)


def _csr(idx):
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    return idx, idx.shape[0], idx.shape[1]


def spmv(idx, x):
    idx, rows, k = _csr(idx)
    if x.dtype == np.float64:
        y = np.zeros(rows, np.float64)
        lib().or_spmv_f64(_p(idx, u32p), rows, k, _p(np.ascontiguousarray(x), f64p), _p(y, f64p))
    else:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(rows, np.float32)
        lib().or_spmv_f32(_p(idx, u32p), rows, k, _p(x, f32p), _p(y, f32p))
    return y


def spmtv(idx, y, cols, nthreads=0):
    idx, rows, k = _csr(idx)
    if y.dtype == np.float64:
        z = np.zeros(cols, np.float64)
        lib().or_spmtv_f64(_p(idx, u32p), rows, k, cols, _p(np.ascontiguousarray(y), f64p), _p(z, f64p))
    else:
        y = np.ascontiguousarray(y, dtype=np.float32)
        z = np.zeros(cols, np.float32)
        lib().or_spmtv_f32(_p(idx, u32p), rows, k, cols, _p(y, f32p), _p(z, f32p), nthreads)
    return z


def jacobi(idx, cols):
    idx, rows, k = _csr(idx)
    d = np.zeros(cols, np.float32)
    lib().or_jacobi_f32(_p(idx, u32p), rows, k, cols, _p(d, f32p))
    return d


def pcg(idx, cols, b, iterations, dtype=np.float64):
    idx, rows, k = _csr(idx)
    b = np.ascontiguousarray(b, dtype=dtype)
    x = np.zeros(cols, dtype)
    hist = np.zeros(max(iterations, 1), dtype)
    if dtype == np.float32:
        lib().or_pcg_f32(_p(idx, u32p), rows, k, cols, _p(b, f32p), iterations, _p(x, f32p), _p(hist, f32p))
    else:
        lib().or_pcg_f64(_p(idx, u32p), rows, k, cols, _p(b, f64p), iterations, _p(x, f64p), _p(hist, f64p))
    return x, hist[:iterations]


def cg(idx, cols, b, iterations):
    idx, rows, k = _csr(idx)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(cols)
    hist = np.zeros(max(iterations, 1))
    lib().or_cg_f64(_p(idx, u32p), rows, k, cols, _p(b, f64p), iterations, _p(x, f64p), _p(hist, f64p))
    return x, hist[:iterations]


def lsqr(idx, cols, b, iterations, dtype=np.float64, colscale=None):
    idx, rows, k = _csr(idx)
    b = np.ascontiguousarray(b, dtype=dtype)
    x = np.zeros(cols, dtype)
    ph = np.zeros(max(iterations, 1), dtype)
    if dtype == np.float32:
        lib().or_lsqr_f32(_p(idx, u32p), rows, k, cols, _p(b, f32p), iterations, _p(x, f32p), _p(ph, f32p))
        return x, ph[:iterations], None
    nr = np.zeros(max(iterations, 1))
    cs = None if colscale is None else np.ascontiguousarray(colscale, dtype=np.float64)
    lib().or_lsqr_f64(_p(idx, u32p), rows, k, cols, _p(b, f64p), iterations, _p(cs, f64p),
                      _p(x, f64p), _p(ph, f64p), _p(nr, f64p))
    return x, ph[:iterations], nr[:iterations]


def lsmr(idx, cols, b, iterations):
    idx, rows, k = _csr(idx)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(cols)
    lib().or_lsmr_f64(_p(idx, u32p), rows, k, cols, _p(b, f64p), iterations, _p(x, f64p))
    return x


def rel_normal_residual(idx, cols, b, x):
    idx, rows, k = _csr(idx)
    rn, an = C.c_double(), C.c_double()
    v = lib().or_rel_normal_residual(_p(idx, u32p), rows, k, cols,
                                     _p(np.ascontiguousarray(b, dtype=np.float64), f64p),
                                     _p(np.ascontiguousarray(x, dtype=np.float64), f64p),
                                     C.byref(rn), C.byref(an))
    return v, rn.value, an.value


def random_play(seed, n, e_lo, e_hi=None):
    out = np.zeros(2 * n, dtype=np.uint64)
    lib().or_random_play(seed, n, e_lo, e_lo if e_hi is None else e_hi, _p(out, u64p))
    return out.reshape(-1, 2)


def random_positions_mt(seed, n):
    out = np.zeros(2 * n, dtype=np.uint64)
    lib().or_random_positions_mt(seed, n, _p(out, u64p))
    return out.reshape(-1, 2)


def random_positions_mt_dist(seed, n):
    out = np.zeros(2 * n, dtype=np.uint64)
    lib().or_random_positions_mt_dist(seed, n, _p(out, u64p))
    return out.reshape(-1, 2)


def make_targets(idx, cols, seed_x, seed_n, sx=1.5, sn=4.0):
    idx, rows, k = _csr(idx)
    xt = np.zeros(cols)
    b = np.zeros(rows)
    lib().or_make_targets(_p(idx, u32p), rows, k, cols, seed_x, seed_n, sx, sn, _p(xt, f64p), _p(b, f64p))
    return xt, b
