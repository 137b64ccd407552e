"""Writes tests/golden/reference_math.npz: outputs of the reference's own
Math/Vector.cpp + Math/MatrixCSR.cpp (compiled by oracle/Makefile into
oracle/_ref/libref_math.so) on a fixed pattern matrix, so the CPU tests can
check the oracle and the GPU against the reference where the build is absent.

Run from the repository root after `make -C oracle`:
    python tests/golden/make_reference_math.py
Only order-independent results are stored: A x in float (a sequential row
sum), x A on integer-valued y (exact in any order), the Jacobi diagonal
(integer counts), and the reference's dot/norm on short integer vectors.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from cassandra_amd import names as N  # noqa: E402


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libref_math.so is not built (make -C oracle with /root/reference present)")
    g = O.Group([N.D8, N.L1, N.D6])  # classes DC, V (from h), D; 7,020 columns
    pos = O.random_play(4242, 600, 22)
    idx = np.ascontiguousarray(g.indices(pos))
    rows, k = idx.shape
    cols = g.cols
    rng = np.random.default_rng(4242)
    x = rng.normal(size=cols).astype(np.float32)
    yi = rng.integers(-64, 65, size=rows).astype(np.float32)
    m = R.ref_csr_create(idx.ctypes.data_as(O.u32p), rows, k, cols)
    try:
        ax = np.zeros(rows, np.float32)
        R.ref_spmv(m, x.ctypes.data_as(O.f32p), cols, ax.ctypes.data_as(O.f32p))
        ya = np.zeros(cols, np.float32)
        R.ref_spmtv(m, yi.ctypes.data_as(O.f32p), rows, ya.ctypes.data_as(O.f32p))
        jac = np.zeros(cols, np.float32)
        R.ref_jacobi(m, jac.ctypes.data_as(O.f32p))
    finally:
        R.ref_csr_free(m)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_math.npz")
    np.savez_compressed(out, positions=pos, idx=idx, cols=np.int64(cols), x=x, ax=ax, y=yi, ya=ya, jacobi=jac)
    print(f"wrote {out}: {rows} rows, k={k}, cols={cols}")


if __name__ == "__main__":
    main()
