"""Differentiable pattern evaluator over the C-ABI.

    score(pos) = sum_s w[idx[pos, s]]     (Pattern/Evaluator.cpp:48-54, compact
                                           weights: the A x identity)

`pattern_eval(w, idx, patterns)` is a torch.autograd.Function: the forward is
the HIP gather-sum, the hand-written backward is the scatter-add adjoint
grad_w = A^T g.  Both run on torch's current stream.  `huber_fit` trains the
weights by Adam on a Huber loss over minibatches of precomputed indices.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import capi
from .fit import Context, Patterns


class _OnTorchStream:
    """Run library calls on torch's current stream for the duration."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def __enter__(self):
        s = torch.cuda.current_stream(self.ctx.torch_device).cuda_stream
        capi.call("cad_context_set_stream", self.ctx.h, C.c_void_p(s))
        return self

    def __exit__(self, *exc):
        capi.call("cad_context_reset_stream", self.ctx.h)
        return False


def _check(idx, w, patterns):
    if not (idx.is_cuda and idx.dtype == torch.int32 and idx.dim() == 2 and idx.shape[1] == patterns.k):
        raise ValueError(f"indices must be a CUDA int32 tensor [B, {patterns.k}]")
    if not (w.is_cuda and w.dim() == 1 and w.shape[0] == patterns.cols and w.dtype in (torch.float32,
                                                                                       torch.float64)):
        raise ValueError(f"weights must be a CUDA float32/float64 vector of {patterns.cols}")


class PatternEval(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, idx, patterns):
        _check(idx, w, patterns)
        idx = idx.contiguous()
        w = w.contiguous()
        out = torch.empty(idx.shape[0], dtype=w.dtype, device=w.device)
        name = "cad_eval_forward_f32" if w.dtype == torch.float32 else "cad_eval_forward_f64"
        with _OnTorchStream(patterns.ctx):
            capi.call(name, patterns.h, C.c_void_p(idx.data_ptr()), idx.shape[0], C.c_void_p(w.data_ptr()),
                      C.c_void_p(out.data_ptr()))
        ctx.save_for_backward(idx)
        ctx.patterns = patterns
        ctx.wdtype = w.dtype
        return out

    @staticmethod
    def backward(ctx, g):
        (idx,) = ctx.saved_tensors
        patterns = ctx.patterns
        g = g.contiguous().to(ctx.wdtype)
        grad = torch.empty(patterns.cols, dtype=ctx.wdtype, device=g.device)
        name = "cad_eval_backward_f32" if ctx.wdtype == torch.float32 else "cad_eval_backward_f64"
        with _OnTorchStream(patterns.ctx):
            capi.call(name, patterns.h, C.c_void_p(idx.data_ptr()), idx.shape[0], C.c_void_p(g.data_ptr()),
                      C.c_void_p(grad.data_ptr()))
        return grad, None, None


def pattern_eval(w: torch.Tensor, idx: torch.Tensor, patterns: Patterns) -> torch.Tensor:
    return PatternEval.apply(w, idx, patterns)


class PatternEvaluator(torch.nn.Module):
    """Compact-weight pattern evaluator (ScoreEstimator, Pattern/Evaluator.h:12-25)."""

    def __init__(self, patterns: Patterns, dtype=torch.float32):
        super().__init__()
        self.patterns = patterns
        self.weight = torch.nn.Parameter(torch.zeros(patterns.cols, dtype=dtype, device=patterns.ctx.torch_device))

    def forward(self, idx: torch.Tensor) -> torch.Tensor:
        return pattern_eval(self.weight, idx, self.patterns)


def huber_fit_fused(patterns: Patterns, idx: torch.Tensor, target: torch.Tensor, epochs: int = 1,
                    batch: int = 65536, lr: float = 0.05, delta: float = 4.0, seed: int = 0, shuffle: bool = True,
                    betas=(0.9, 0.999), eps: float = 1e-8):
    """huber_fit with every epoch one call of the library
    (cad_eval_train_epoch_f32; each step: the forward with the loss and its
    gradient, reading the minibatch's rows through the shuffled order, the
    backward, the Adam update); the same minibatches, loss and update as
    huber_fit.  Returns the trained evaluator and the per-epoch mean loss."""
    if idx.dtype != torch.int32 or not idx.is_contiguous() or idx.dim() != 2 or idx.shape[1] != patterns.k:
        raise ValueError("idx must be a contiguous int32 [N, k] tensor of the pattern group's indices")
    if target.shape[0] != idx.shape[0]:
        raise ValueError("Size mismatch.")
    n = idx.shape[0]
    gen = torch.Generator(device=idx.device).manual_seed(seed)
    L = capi.lib()
    ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    losses = []
    step = 0
    torch.cuda.current_stream(idx.device).synchronize()  # idx and target are complete
    # everything on the library's stream, in order with its kernels
    with torch.cuda.stream(patterns.ctx.torch_stream()):
        model = PatternEvaluator(patterns)
        w = model.weight.data
        mom = torch.zeros_like(w)
        var = torch.zeros_like(w)
        tgt = target.to(torch.float32).contiguous()
        for _ in range(epochs):
            order = torch.randperm(n, device=idx.device, generator=gen) if shuffle else torch.arange(n, device=idx.device)
            total = torch.zeros((), dtype=torch.float64, device=idx.device)
            nb = math.ceil(n / batch)
            # the epoch's steps, enqueued by one call (cad_eval_train_epoch_f32)
            capi.check(L.cad_eval_train_epoch_f32(patterns.h, ptr(idx), ptr(tgt), ptr(order), n, batch, delta, lr,
                                                  betas[0], betas[1], eps, step + 1, ptr(w), ptr(mom), ptr(var),
                                                  ptr(total)))
            step += nb
            losses.append(float(total) / nb)
    return model, losses


def huber_fit(patterns: Patterns, idx: torch.Tensor, target: torch.Tensor, epochs: int = 1, batch: int = 65536,
              lr: float = 0.05, delta: float = 4.0, seed: int = 0, shuffle: bool = True):
    """Adam on the Huber loss (delta) over minibatches of rows; returns the
    trained evaluator and the per-epoch mean loss."""
    model = PatternEvaluator(patterns)
    # the fused step: one kernel per minibatch instead of a chain of small ones
    opt = torch.optim.Adam(model.parameters(), lr=lr, fused=True)
    n = idx.shape[0]
    gen = torch.Generator(device=idx.device).manual_seed(seed)
    losses = []
    for _ in range(epochs):
        order = torch.randperm(n, device=idx.device, generator=gen) if shuffle else torch.arange(n, device=idx.device)
        total = torch.zeros((), dtype=torch.float64, device=idx.device)
        nb = math.ceil(n / batch)
        for i in range(nb):
            sel = order[i * batch:(i + 1) * batch]
            pred = model(idx.index_select(0, sel))
            loss = torch.nn.functional.huber_loss(pred, target.index_select(0, sel).to(pred.dtype), delta=delta)
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()
            total += loss.detach()
        losses.append(float(total) / nb)
    return model, losses
