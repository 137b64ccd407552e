"""Python mirror of the reference's fit API over the C-ABI, with torch tensors
as device buffers.

    Context            device + stream                    (no reference counterpart)
    Patterns           GroupIndexer / CreateIndexer       (Pattern/Indexer.h:10-40)
    Matrix             MatrixCSR, A*x, x*A, Jacobi        (Math/MatrixCSR.h:11-33)
    Matrix.solve       PCG / LSQR / LSMR / CG             (Math/Solver.h:9-287)
    fit_weights        FitWeights                         (PatternFit/PatternFit.cpp:7-39)

Device tensors are CUDA (HIP) torch tensors; positions are int64 tensors of
shape [n, 2] holding the bits of {P, O}.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np
import torch

from . import capi
from .capi import call


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _dev(t, dtype):
    if not (t.is_cuda and t.dtype == dtype and t.is_contiguous()):
        raise ValueError(f"expected a contiguous {dtype} device tensor, got {t.dtype} on {t.device}")
    return t


class Context:
    """A device and the stream the library enqueues on (its own by default)."""

    def __init__(self, device: int = 0):
        self.device = device
        h = C.c_void_p()
        call("cad_context_create", device, C.byref(h))
        self.h = h
        self.torch_device = torch.device("cuda", device)
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        if not self.h:
            raise RuntimeError("context is closed")
        self._children.add(obj)

    def close(self):
        """Destroys every live pattern group / matrix of the context, then the
        context itself."""
        if getattr(self, "h", None):
            for child in list(self._children):
                child._destroy()
            capi.lib().cad_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream_handle(self) -> int:
        s = C.c_void_p()
        call("cad_context_get_stream", self.h, C.byref(s))
        return s.value or 0

    def torch_stream(self):
        """The library stream as a torch stream (for events / ordering)."""
        return torch.cuda.ExternalStream(self.stream_handle(), device=self.torch_device)

    def synchronize(self):
        call("cad_context_synchronize", self.h)

    def random_play(self, seed: int, n: int, e_lo: int, e_hi: int | None = None) -> torch.Tensor:
        """n positions of uniformly random legal play (int64 [n, 2])."""
        out = torch.empty((n, 2), dtype=torch.int64, device=self.torch_device)
        torch.cuda.synchronize(self.torch_device)
        call("cad_random_play", self.h, seed, n, e_lo, e_lo if e_hi is None else e_hi, _ptr(out))
        self.synchronize()
        return out


class Patterns:
    """GroupIndexer over a list of pattern masks."""

    def __init__(self, ctx: Context, masks):
        self.ctx = ctx
        arr = np.ascontiguousarray(np.array(masks, dtype=np.uint64))
        h = C.c_void_p()
        call("cad_patterns_create", ctx.h, arr.ctypes.data_as(C.c_void_p), len(arr), C.byref(h))
        self.h = h
        ctx._adopt(self)
        k, cols, npat = C.c_int(), C.c_int64(), C.c_int()
        call("cad_patterns_info", h, C.byref(k), C.byref(cols), C.byref(npat))
        self.k, self.cols, self.npatterns = k.value, cols.value, npat.value

    def _destroy(self):
        if getattr(self, "h", None):
            capi.lib().cad_patterns_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass

    def pattern_info(self, i):
        cls, mask, size, order, off = C.c_int(), C.c_uint64(), C.c_int64(), C.c_int(), C.c_int64()
        call("cad_patterns_pattern_info", self.h, i, C.byref(cls), C.byref(mask), C.byref(size), C.byref(order),
             C.byref(off))
        return dict(cls=cls.value, mask=mask.value, size=size.value, order=order.value, offset=off.value)

    def variations(self):
        out = np.zeros(self.k, dtype=np.uint64)
        call("cad_patterns_variations", self.h, out.ctypes.data_as(C.c_void_p))
        return out

    def indices_host(self, PO) -> np.ndarray:
        PO = np.ascontiguousarray(np.asarray(PO, dtype=np.uint64).reshape(-1, 2))
        out = np.zeros((PO.shape[0], self.k), dtype=np.uint32)
        call("cad_patterns_indices_host", self.h, PO.ctypes.data_as(C.c_void_p), PO.shape[0],
             out.ctypes.data_as(C.c_void_p))
        return out

    def extract(self, pos: torch.Tensor, layout: int = capi.LAYOUT_ROW_MAJOR) -> torch.Tensor:
        """Column indices of every row on the device (int32; [n, k] row-major
        or [k, n] slot-major)."""
        _dev(pos, torch.int64)
        n = pos.shape[0]
        shape = (n, self.k) if layout == capi.LAYOUT_ROW_MAJOR else (self.k, n)
        out = torch.empty(shape, dtype=torch.int32, device=pos.device)
        torch.cuda.synchronize(pos.device)
        call("cad_index_extract", self.h, _ptr(pos), n, _ptr(out), layout)
        self.ctx.synchronize()
        return out

    def expanded_size(self) -> int:
        n = C.c_int64()
        call("cad_patterns_expanded_size", self.h, C.byref(n))
        return n.value

    def expand(self, w: torch.Tensor) -> torch.Tensor:
        """Compact weights -> the expanded tables, one of 3^n floats per
        variation (ScoreEstimator's constructor, Pattern/Evaluator.cpp:11-29)."""
        _dev(w, torch.float32)
        out = torch.empty(self.expanded_size(), dtype=torch.float32, device=w.device)
        torch.cuda.synchronize(w.device)
        call("cad_weights_expand", self.h, _ptr(w), _ptr(out))
        self.ctx.synchronize()
        return out

    def compact(self, tables: torch.Tensor) -> torch.Tensor:
        """Expanded tables -> compact weights (ScoreEstimator::Weights,
        Pattern/Evaluator.cpp:56-73)."""
        _dev(tables, torch.float32)
        out = torch.empty(self.cols, dtype=torch.float32, device=tables.device)
        torch.cuda.synchronize(tables.device)
        call("cad_weights_compact", self.h, _ptr(tables), _ptr(out))
        return out

    def evaluator(self, w: torch.Tensor) -> "Evaluator":
        return Evaluator(self, w)

    def matrix(self, pos: torch.Tensor) -> "Matrix":
        _dev(pos, torch.int64)
        torch.cuda.synchronize(pos.device)
        h = C.c_void_p()
        call("cad_matrix_from_positions", self.h, _ptr(pos), pos.shape[0], C.byref(h))
        return Matrix(self.ctx, h)

    def matrix_free(self, pos: torch.Tensor) -> "Matrix":
        """The matrix of pos without stored indices: every product recomputes
        the rows' indices from a library-held copy of the positions."""
        _dev(pos, torch.int64)
        torch.cuda.synchronize(pos.device)
        h = C.c_void_p()
        call("cad_matrix_from_positions_mf", self.h, _ptr(pos), pos.shape[0], C.byref(h))
        return Matrix(self.ctx, h)

    def matrix_systems(self, pos: torch.Tensor, row_begin) -> "Matrix":
        """Block-diagonal batch: rows [row_begin[s], row_begin[s+1]) of pos form
        system s, which owns columns [s * cols, (s + 1) * cols) of x."""
        _dev(pos, torch.int64)
        rb = np.ascontiguousarray(np.asarray(row_begin, dtype=np.int64))
        if rb.ndim != 1 or rb.shape[0] < 2 or int(rb[-1]) != pos.shape[0]:
            raise ValueError("row_begin must hold nsys + 1 boundaries ending at the row count")
        torch.cuda.synchronize(pos.device)
        h = C.c_void_p()
        call("cad_matrix_from_positions_systems", self.h, _ptr(pos), rb.ctypes.data_as(C.c_void_p),
             rb.shape[0] - 1, C.byref(h))
        return Matrix(self.ctx, h)


class Evaluator:
    """A resident evaluator: the expanded tables of one weight vector kept on
    the device (ScoreEstimator::Eval for batches of positions)."""

    def __init__(self, patterns: Patterns, w: torch.Tensor):
        _dev(w, torch.float32)
        self.patterns = patterns
        self.ctx = patterns.ctx
        h = C.c_void_p()
        torch.cuda.synchronize(w.device)
        call("cad_evaluator_create", patterns.h, _ptr(w), C.byref(h))
        self.ctx.synchronize()
        self.h = h
        self.ctx._adopt(self)

    def _destroy(self):
        if getattr(self, "h", None):
            capi.lib().cad_evaluator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass

    def set_weights(self, w: torch.Tensor):
        _dev(w, torch.float32)
        torch.cuda.synchronize(w.device)
        call("cad_evaluator_set_weights", self.h, _ptr(w))
        self.ctx.synchronize()

    def __call__(self, pos: torch.Tensor) -> torch.Tensor:
        _dev(pos, torch.int64)
        out = torch.empty(pos.shape[0], dtype=torch.float32, device=pos.device)
        torch.cuda.synchronize(pos.device)
        call("cad_evaluator_eval", self.h, _ptr(pos), pos.shape[0], _ptr(out))
        self.ctx.synchronize()
        return out

    def eval_host(self, PO) -> np.ndarray:
        PO = np.ascontiguousarray(np.asarray(PO, dtype=np.uint64).reshape(-1, 2))
        out = np.zeros(PO.shape[0], dtype=np.float32)
        call("cad_evaluator_eval_host", self.h, PO.ctypes.data_as(C.c_void_p), PO.shape[0],
             out.ctypes.data_as(C.c_void_p))
        return out


class Matrix:
    """Device MatrixCSR (fixed k per row, implicit value 1, repeats allowed)."""

    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self.h = h
        ctx._adopt(self)
        rows, k, cols, ib = C.c_int64(), C.c_int(), C.c_int64(), C.c_int()
        call("cad_matrix_info", h, C.byref(rows), C.byref(k), C.byref(cols), C.byref(ib))
        self.rows, self.k, self.cols, self.index_bytes = rows.value, k.value, cols.value, ib.value
        ns, cps = C.c_int(), C.c_int64()
        call("cad_matrix_systems", h, C.byref(ns), C.byref(cps), None)
        self.nsys, self.cols_per_system = ns.value, cps.value
        rb = np.zeros(self.nsys + 1, dtype=np.int64)
        call("cad_matrix_systems", h, None, None, rb.ctypes.data_as(C.c_void_p))
        self.row_begin = rb

    @classmethod
    def from_host(cls, ctx: Context, idx: np.ndarray, cols: int) -> "Matrix":
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        h = C.c_void_p()
        call("cad_matrix_create", ctx.h, idx.ctypes.data_as(C.c_void_p), idx.shape[0], idx.shape[1], cols,
             C.byref(h))
        return cls(ctx, h)

    def _destroy(self):
        if getattr(self, "h", None):
            capi.lib().cad_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass

    def autotune(self) -> str:
        """Pick the fastest product schedules for this matrix; returns the log."""
        buf = C.create_string_buffer(8192)
        call("cad_matrix_autotune", self.h, buf, len(buf))
        return buf.value.decode()

    def kernel_times(self, reps: int = 20) -> dict:
        """Average launch times (us) of the FP32 product kernels (HIP events on
        the library's stream): gather, scatter tiles, scatter reduce, absmax,
        and the Jacobi count pass."""
        out = (C.c_double * 5)()
        call("cad_matrix_kernel_times", self.h, int(reps), C.cast(out, C.c_void_p), 5)
        return dict(zip(("gather", "scatter_tiles", "scatter_reduce", "absmax", "jacobi_part"), list(out)))

    def schedule(self) -> dict:
        """The product schedule in use (cad_schedule fields)."""
        sc = capi.Schedule()
        call("cad_matrix_get_schedule", self.h, C.byref(sc))
        return {name: getattr(sc, name) for name, _ in capi.Schedule._fields_}

    def set_schedule(self, **knobs) -> None:
        """Change some fields of the product schedule (the rest stay)."""
        cur = self.schedule()
        unknown = set(knobs) - set(cur)
        if unknown:
            raise ValueError(f"unknown schedule fields {sorted(unknown)}")
        cur.update(knobs)
        sc = capi.Schedule(**cur)
        call("cad_matrix_set_schedule", self.h, C.byref(sc))

    def download(self) -> np.ndarray:
        out = np.zeros((self.rows, self.k), dtype=np.uint32)
        call("cad_matrix_download", self.h, out.ctypes.data_as(C.c_void_p))
        return out

    def _vec(self, n, dtype):
        return torch.empty(n, dtype=dtype, device=self.ctx.torch_device)

    def spmv(self, x: torch.Tensor) -> torch.Tensor:
        """A x (Math/MatrixCSR.cpp:49-65)."""
        if x.shape[0] != self.cols:
            raise ValueError("Size mismatch.")
        y = self._vec(self.rows, x.dtype)
        torch.cuda.synchronize(x.device)
        name = "cad_spmv_f32" if x.dtype == torch.float32 else "cad_spmv_f64"
        call(name, self.h, _ptr(_dev(x, x.dtype)), _ptr(y))
        self.ctx.synchronize()
        return y

    def spmtv(self, y: torch.Tensor) -> torch.Tensor:
        """y A = A^T y (Math/MatrixCSR.cpp:67-87)."""
        if y.shape[0] != self.rows:
            raise ValueError("Size mismatch.")
        z = self._vec(self.cols, y.dtype)
        torch.cuda.synchronize(y.device)
        name = "cad_spmtv_f32" if y.dtype == torch.float32 else "cad_spmtv_f64"
        call(name, self.h, _ptr(_dev(y, y.dtype)), _ptr(z))
        self.ctx.synchronize()
        return z

    def jacobi(self) -> torch.Tensor:
        """JacobiPreconditionerOfATA (Math/MatrixCSR.cpp:21-47)."""
        d = self._vec(self.cols, torch.float32)
        call("cad_jacobi", self.h, _ptr(d))
        self.ctx.synchronize()
        return d

    def make_targets(self, seed_x: int, seed_n: int, sx: float = 1.5, sn: float = 4.0):
        xt = self._vec(self.cols, torch.float64)
        b = self._vec(self.rows, torch.float64)
        call("cad_make_targets", self.h, seed_x, seed_n, sx, sn, _ptr(xt), _ptr(b))
        self.ctx.synchronize()
        return xt, b

    def solve(self, b: torch.Tensor, **options):
        """Solve from x0 = 0; returns (x [cols] float64, report dict)."""
        _dev(b, torch.float64)
        if b.shape[0] != self.rows:
            raise ValueError("Size mismatch.")
        x = self._vec(self.cols, torch.float64)
        o = capi.default_options(**options)
        rep = capi.SolveReport()
        torch.cuda.synchronize(b.device)
        call("cad_solve", self.h, _ptr(b), _ptr(x), C.byref(o), C.byref(rep))
        return x, rep.as_dict()

    def solve_systems(self, b: torch.Tensor, **options):
        """Solve every system of a batch in the same launches; returns (x,
        report dict, per-system metric, per-system iterations)."""
        _dev(b, torch.float64)
        if b.shape[0] != self.rows:
            raise ValueError("Size mismatch.")
        x = self._vec(self.cols, torch.float64)
        o = capi.default_options(**options)
        rep = capi.SolveReport()
        rel = np.zeros(self.nsys, dtype=np.float64)
        its = np.zeros(self.nsys, dtype=np.int32)
        torch.cuda.synchronize(b.device)
        call("cad_solve_systems", self.h, _ptr(b), _ptr(x), C.byref(o), C.byref(rep), rel.ctypes.data_as(C.c_void_p),
             its.ctypes.data_as(C.c_void_p))
        return x, rep.as_dict(), rel, its

    def rel_normal_residual(self, b: torch.Tensor, x: torch.Tensor):
        out, rn = C.c_double(), C.c_double()
        torch.cuda.synchronize(b.device)
        call("cad_rel_normal_residual", self.h, _ptr(_dev(b, torch.float64)), _ptr(_dev(x, torch.float64)),
             C.byref(out), C.byref(rn))
        return out.value, rn.value


def fit_weights(patterns: Patterns, pos: torch.Tensor, scores: torch.Tensor, iterations: int,
                precision: int = capi.PREC_FP32):
    """FitWeights (PatternFit/PatternFit.cpp:7-39): Jacobi-PCG on A^T A from
    x0 = 0 for `iterations` iterations; returns the weights (float64 tensor)."""
    A = patterns.matrix(pos)
    x, rep = A.solve(scores.to(torch.float64).contiguous(), solver=capi.SOLVER_PCG, precision=precision,
                     max_iterations=iterations, tol=0.0, check_every=max(1, iterations))
    return x, rep
