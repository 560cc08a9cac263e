"""torch.autograd harness around the WFBP operator and its hand-written adjoint.

A   : raw fan data p -> native volume x (rebin -> filter -> W-weighted
      backprojection, the FP32 plan of wfbp.Plan).
A^T : the exact transpose, i.e. a matched helical forward projector
      (wfbp.Plan.adjoint, HIP kernels in csrc/adjoint.hip).

WFBPFunction.apply(p) = A p with backward g -> A^T g, and
ProjectFunction.apply(x) = A^T x with backward g -> A g, so torch can
differentiate any loss built on them.  `refine` runs gradient steps on
||A^T x - p||^2 starting from the WFBP volume, with the step size 1 / ||A A^T||
estimated by power iteration.
"""
from __future__ import annotations

import time

import torch


class WFBPFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, raw, plan):
        native = plan.alloc_native()
        plan.forward_native(raw.contiguous(), native)
        ctx.plan = plan
        return native

    @staticmethod
    def backward(ctx, g):
        plan = ctx.plan
        raw_bar = plan.alloc_raw()
        plan.adjoint(g.contiguous(), raw_bar)
        return raw_bar, None


class ProjectFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan):
        p = plan.alloc_raw()
        plan.adjoint(x.contiguous(), p)
        ctx.plan = plan
        return p

    @staticmethod
    def backward(ctx, g):
        plan = ctx.plan
        x_bar = plan.alloc_native()
        plan.forward_native(g.contiguous(), x_bar)
        return x_bar, None


def wfbp(raw, plan):
    return WFBPFunction.apply(raw, plan)


def project(x, plan):
    return ProjectFunction.apply(x, plan)


def power_iteration(plan, shape, iters: int = 12, seed: int = 0) -> float:
    """Largest eigenvalue of A A^T (= ||A^T||^2) by power iteration on volumes."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    v = torch.randn(shape, device="cuda", generator=gen)
    v /= v.norm()
    lam = 0.0
    p = plan.alloc_raw()
    w = plan.alloc_native()
    for _ in range(iters):
        plan.adjoint(v, p)
        plan.forward_native(p, w)
        lam = float(torch.dot(w.flatten().double(), v.flatten().double()))
        v = w / w.norm()
    return lam


def refine(plan, raw, x0, steps: int = 10, mode: str = "store", lam: float = None,
           callback=None):
    """Gradient descent on 0.5 ||A^T x - p||^2 from x0, step 1/lam.

    mode "store": autograd keeps the sinogram-sized residual A^T x - p for the
    backward pass; mode "recompute": torch.utils.checkpoint drops it and
    recomputes A^T x in the backward pass (one extra adjoint per step, one
    sinogram buffer less).  Returns (x, stats).
    """
    from torch.utils.checkpoint import checkpoint

    if lam is None:
        lam = power_iteration(plan, x0.shape)
    tau = 1.0 / lam
    x = x0.detach().clone()
    stats = {"lambda_max": lam, "step": tau, "loss": [], "ms_per_step": [], "peak_mb": []}

    def residual_loss(xx):
        r = project(xx, plan) - raw
        return 0.5 * (r * r).sum()

    for it in range(steps):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        t0 = time.perf_counter()
        xv = x.requires_grad_(True)
        if mode == "recompute":
            loss = checkpoint(residual_loss, xv, use_reentrant=False)
        else:
            loss = residual_loss(xv)
        loss.backward()
        with torch.no_grad():
            x = (xv - tau * xv.grad).detach()
        torch.cuda.synchronize()
        stats["ms_per_step"].append((time.perf_counter() - t0) * 1e3)
        stats["peak_mb"].append(torch.cuda.max_memory_allocated() / 2 ** 20)
        stats["loss"].append(float(loss))
        if callback:
            callback(it, x)
    return x, stats
