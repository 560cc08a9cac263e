"""torch.autograd harness around the WFBP operator and its hand-written adjoint.

A   : raw fan data p -> native volume x (rebin -> filter -> W-weighted
      backprojection, the FP32 plan of wfbp.Plan).
A^T : its exact transpose (wfbp.Plan.adjoint, HIP kernels in csrc/adjoint.hip).
P0  : R^T M^T, a matched helical projector: M^T projects the volume onto the
      parallel rows through the backprojection's own bilinear footprint with
      weight R / l over every view that sees the voxel (adjoint.hip, matched
      mode), R^T is the rebinning transpose; P0^T = M R (rebinning, then the
      R / l-weighted plain sum of backproject_v1.hip's matched mode).

Pr  : the ray-driven projector (Joseph's method, rayproj.hip): line
      integrals of the native volume along the raw rays; Pr^T its exact
      transpose.

WFBPFunction.apply(p) = A p with backward g -> A^T g; ProjectFunction.apply(x)
= A^T x with backward g -> A g; MatchedProjectFunction.apply(x) = P0 x with
backward g -> P0^T g; RayProjectFunction.apply(x) = Pr x with backward
g -> Pr^T g.  `refine` runs gradient steps on 0.5 ||k P x - p||^2 from the
WFBP volume, P = Pr (default, k = 1: a model of the measured line
integrals), P = P0 (k fitted by least squares at the start so that k P0 x0
matches p) or P = A^T (k = 1, the objective whose minimiser is not the
attenuation map, since A^T carries the ramp filter), with step
1 / (k^2 ||P P^T||) from power iteration.
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


def _p0(plan, x):
    """P0 x = R^T M^T x (raw fan data), M^T the matched voxel-driven projector."""
    lines = plan.alloc_lines()
    plan.matched_project(x.contiguous(), lines)   # zero-fills, then scatters
    p = plan.alloc_raw()
    plan.rebin_adjoint(lines, p)
    return p


def _p0t(plan, g):
    """P0^T g = M R g (native volume), M the R / l-weighted plain sum."""
    lines = plan.alloc_lines()
    plan.rebin_lines(g.contiguous(), lines)
    x = plan.alloc_native()
    plan.matched_backproject(lines, x)
    return x


class MatchedProjectFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan):
        ctx.plan = plan
        return _p0(plan, x)

    @staticmethod
    def backward(ctx, g):
        return _p0t(ctx.plan, g), None


def matched_project(x, plan):
    return MatchedProjectFunction.apply(x, plan)


def _pr(plan, x):
    p = plan.alloc_raw()
    plan.ray_project(x.contiguous(), p)
    return p


def _prt(plan, g):
    x = plan.alloc_native()
    plan.ray_backproject(g.contiguous(), x)
    return x


class RayProjectFunction(torch.autograd.Function):
    """Pr x: line integrals of the native volume along the raw rays (Joseph's
    method, rayproj.hip); backward g -> Pr^T g, its exact transpose."""

    @staticmethod
    def forward(ctx, x, plan):
        ctx.plan = plan
        return _pr(plan, x)

    @staticmethod
    def backward(ctx, g):
        return _prt(ctx.plan, g), None


def ray_project(x, plan):
    return RayProjectFunction.apply(x, plan)


def wfbp(raw, plan):
    return WFBPFunction.apply(raw, plan)


def project(x, plan):
    return ProjectFunction.apply(x, plan)


def power_iteration(plan, shape, iters: int = 12, seed: int = 0,
                    operator: str = "adjoint") -> float:
    """Largest eigenvalue of P^T P for P = A^T ("adjoint") or P0 ("matched"),
    by power iteration on volumes."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    v = torch.randn(shape, device="cuda", generator=gen)
    v /= v.norm()
    lam = 0.0
    for _ in range(iters):
        with torch.no_grad():
            if operator == "matched":
                w = _p0t(plan, _p0(plan, v))
            elif operator == "ray":
                w = _prt(plan, _pr(plan, v))
            else:
                p = plan.alloc_raw()
                plan.adjoint(v, p)
                w = plan.alloc_native()
                plan.forward_native(p, w)
        lam = float(torch.dot(w.flatten().double(), v.flatten().double()))
        v = w / w.norm()
    return lam


def inside_rays(plan, margin: float = 1.0):
    """[view][row][1] mask of the rays whose whole path through the
    reconstruction FOV lies inside the volume's z-range (shrunk by `margin`
    mm): only those are explained by the reconstructed slab.  The ray of
    detector row r leaves the source at the view's table z and climbs
    (r - cr) drow per Dsd of in-plane travel; inside the FOV the in-plane
    distance from the source is within R -+ FOV/2."""
    sc, rc = plan.scan, plan.recon
    if sc.ffs_mode:
        raise ValueError("inside_rays: FFS scans are not supported")
    v = torch.arange(sc.n_views, device="cuda", dtype=torch.float64)
    zsrc = sc.z0 + v * sc.feed / sc.views_per_rotation
    w = (torch.arange(sc.n_rows, device="cuda", dtype=torch.float64) - sc.central_row) * \
        sc.row_spacing
    half = 0.5 * rc.recon_fov * 2 ** 0.5
    lo_l, hi_l = max(sc.r - half, 0.0), sc.r + half
    za = zsrc[:, None] + w[None, :] * (lo_l / sc.dsd)
    zb = zsrc[:, None] + w[None, :] * (hi_l / sc.dsd)
    zmin = min(rc.start_pos, rc.end_pos) + margin
    zmax = max(rc.start_pos, rc.end_pos) - margin
    ok = (torch.minimum(za, zb) >= zmin) & (torch.maximum(za, zb) <= zmax)
    return ok[..., None].float()


def refine(plan, raw, x0, steps: int = 10, mode: str = "store", lam: float = None,
           callback=None, operator: str = "ray", mask=None):
    """Gradient descent on 0.5 ||k P x - p||^2 from x0 with step 1 / (k^2 lam).

    operator "ray": P = Pr, the ray-driven line-integral projector, k = 1;
    operator "matched": P = P0 = R^T M^T, k = <P0 x0, p> / <P0 x0, P0 x0>;
    operator "adjoint": P = A^T, k = 1.
    mask (broadcastable to the raw data, e.g. inside_rays(plan)) restricts
    the data term to rays the volume can explain: 0.5 ||m (k P x - p)||^2.
    mode "store": autograd keeps the sinogram-sized residual k P x - p for the
    backward pass; mode "recompute": torch.utils.checkpoint drops it and
    recomputes P x in the backward pass (one more projection per step, one
    sinogram buffer less).  Returns (x, stats).
    """
    from torch.utils.checkpoint import checkpoint

    proj = {"matched": matched_project, "ray": ray_project, "adjoint": project}[operator]
    m = mask if mask is not None else torch.ones((), device=raw.device)
    with torch.no_grad():
        if operator == "matched":
            px0 = matched_project(x0, plan) * m
            k = float(torch.dot(px0.flatten().double(), (raw * m).flatten().double()) /
                      torch.dot(px0.flatten().double(), px0.flatten().double()))
            del px0
        else:
            k = 1.0
    if lam is None:
        lam = power_iteration(plan, x0.shape, operator=operator)
    tau = 1.0 / (k * k * lam)
    x = x0.detach().clone()
    stats = {"operator": operator, "scale": k, "lambda_max": lam, "step": tau, "loss": [],
             "ms_per_step": [], "peak_mb": []}

    def residual_loss(xx):
        r = (k * proj(xx, plan) - raw) * m
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
        stats["loss"].append(float(loss.detach()))
        if callback:
            callback(it, x)
    return x, stats
