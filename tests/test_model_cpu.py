"""The weight file (IO/Binary.cpp:3-68): the Python writer and reader against
the C++ drop-in (same bytes both ways) and the layout the format fixes."""
import os
import struct
import subprocess

import numpy as np
import pytest

from cassandra_amd import model
from cassandra_amd import names as N

RUNNER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cassandra_amd", "_build",
                      "dropin_test")


def sample():
    rng = np.random.default_rng(4)
    pattern = [N.L02X, N.D8]
    cols = 29646 + 3321
    return model.Estimator(20, pattern, [rng.normal(size=cols).astype(np.float32) for _ in range(model.stages_for(20))])


def test_layout():
    e = sample()
    data = model.dumps(e)
    assert struct.unpack("<i", data[:4]) == (20,)
    assert struct.unpack("<QQQ", data[4:28]) == (2, N.L02X, N.D8)
    assert struct.unpack("<Q", data[28:36]) == (4,)
    m = e.weights[0].shape[0]
    assert len(data) == 4 + 24 + 8 + 4 * (24 + 8 + 4 * m) + 8 + 4 * 5
    back = model.loads(data)
    assert back.stage_size == 20 and back.pattern == e.pattern and back.stages() == 4
    assert all(np.array_equal(a, b) for a, b in zip(back.weights, e.weights))
    assert back.accuracy == pytest.approx(model.DEFAULT_ACCURACY)
    with pytest.raises(ValueError):
        model.loads(data[:-1])
    with pytest.raises(ValueError):
        model.loads(data + b"\0")


def test_stage_count():
    # Pattern/Evaluator.cpp:6-9: ceil(65 / stage_size)
    assert [model.stages_for(s) for s in (1, 5, 10, 20, 64, 65, 66)] == [65, 13, 7, 4, 2, 1, 1]


def transformed(mask):
    """The mask an indexer holds (Pattern/Indexer.cpp:345-364), from the library."""
    import ctypes as C

    import cassandra_amd.capi as capi

    arr = np.array([mask], dtype=np.uint64)
    h = C.c_void_p()
    capi.call("cad_patterns_create", None, arr.ctypes.data_as(C.c_void_p), 1, C.byref(h))
    cls, m, size, order, off = C.c_int(), C.c_uint64(), C.c_int64(), C.c_int(), C.c_int64()
    capi.call("cad_patterns_pattern_info", h, 0, C.byref(cls), C.byref(m), C.byref(size), C.byref(order),
              C.byref(off))
    capi.lib().cad_patterns_destroy(h)
    return m.value


@pytest.mark.skipif(not os.path.exists(RUNNER), reason="drop-in runner not built")
def test_cpp_reads_and_writes_the_same_bytes(tmp_path):
    # ScoreEstimator::Pattern() returns the indexers' (transformed) masks, so a
    # file written from the given masks comes back with those; from then on the
    # copy is byte-identical
    e = sample()
    src, dst, again = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "again.bin"
    model.save(src, e)
    for a, b in ((src, dst), (dst, again)):
        r = subprocess.run([RUNNER, "copy-model", str(a), str(b)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
    back = model.load(dst)
    assert back.pattern == [transformed(m) for m in e.pattern]
    assert all(np.array_equal(a, b) for a, b in zip(back.weights, e.weights))
    assert again.read_bytes() == dst.read_bytes()
    e.pattern, e.stage_patterns = back.pattern, None
    model.save(src, e)
    assert src.read_bytes() == dst.read_bytes()
