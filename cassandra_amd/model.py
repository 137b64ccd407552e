"""The weight file of a PatternBasedEstimator (IO/Binary.cpp:3-68), so that
weights fitted here load into the reference's search unchanged.

Little-endian, size_t = 8 bytes:
    i32 stage_size
    u64 n, n x u64 pattern masks (as ScoreEstimator::Pattern() returns them)
    u64 stages, then per stage: u64 n, n x u64 masks; u64 m, m x f32 compact weights
    u64 p, p x f32 accuracy-model parameters
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

# Pattern/Evaluator.h:45 default parameters of the AccuracyModel
DEFAULT_ACCURACY = [-0.16594754, 0.95554334, 0.27973529, -0.01616647, 1.28078789]


@dataclass
class Estimator:
    stage_size: int
    pattern: list                       # u64 masks (transformed, as Pattern() returns)
    weights: list                       # one float32 array of compact weights per stage
    stage_patterns: list = None         # per-stage masks (default: pattern)
    accuracy: list = field(default_factory=lambda: list(DEFAULT_ACCURACY))

    def stages(self) -> int:
        return len(self.weights)


def _vec(fmt, values):
    arr = np.ascontiguousarray(np.asarray(values, dtype=fmt))
    return struct.pack("<Q", arr.shape[0]) + arr.astype(arr.dtype.newbyteorder("<")).tobytes()


def dumps(e: Estimator) -> bytes:
    out = [struct.pack("<i", int(e.stage_size)), _vec(np.uint64, e.pattern), struct.pack("<Q", e.stages())]
    sp = e.stage_patterns or [e.pattern] * e.stages()
    for masks, w in zip(sp, e.weights):
        out.append(_vec(np.uint64, masks))
        out.append(_vec(np.float32, w))
    out.append(_vec(np.float32, e.accuracy))
    return b"".join(out)


def loads(data: bytes) -> Estimator:
    pos = 0

    def take(n):
        nonlocal pos
        if pos + n > len(data):
            raise ValueError("truncated weight file")
        b = data[pos:pos + n]
        pos += n
        return b

    def vec(dtype):
        (n,) = struct.unpack("<Q", take(8))
        dt = np.dtype(dtype).newbyteorder("<")
        return np.frombuffer(take(n * dt.itemsize), dtype=dt).astype(dtype)

    (stage_size,) = struct.unpack("<i", take(4))
    pattern = [int(v) for v in vec(np.uint64)]
    (stages,) = struct.unpack("<Q", take(8))
    sp, ws = [], []
    for _ in range(stages):
        sp.append([int(v) for v in vec(np.uint64)])
        ws.append(vec(np.float32))
    acc = [float(v) for v in vec(np.float32)]
    if pos != len(data):
        raise ValueError("trailing bytes after the weight file")
    return Estimator(stage_size, pattern, ws, sp, acc)


def save(path, e: Estimator) -> None:
    with open(path, "wb") as f:
        f.write(dumps(e))


def load(path) -> Estimator:
    with open(path, "rb") as f:
        return loads(f.read())


def stages_for(stage_size: int) -> int:
    """Number of stages of a MultiStageScoreEstimator (Pattern/Evaluator.cpp:6-9)."""
    return (65 + stage_size - 1) // stage_size


def from_fit(patterns, stage_size: int, weights_per_stage) -> Estimator:
    """An estimator file from compact weights fitted on the device (one
    cols-long vector per stage; the masks are the pattern group's, transformed
    as the reference's indexers hold them)."""
    masks = [patterns.pattern_info(i)["mask"] for i in range(patterns.npatterns)]
    ws = [np.asarray(w.detach().cpu() if hasattr(w, "detach") else w, dtype=np.float32) for w in weights_per_stage]
    for w in ws:
        if w.shape != (patterns.cols,):
            raise ValueError(f"each stage needs {patterns.cols} weights")
    return Estimator(stage_size, masks, ws)
