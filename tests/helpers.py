"""Shared helpers for the GPU parity tests."""
import math

import numpy as np

from freect_amd import scans as S

HU = S.MU_WATER / 1000.0  # 1 HU in 1/mm


def fov_mask(rc, acq_fov, margin=0.0):
    x = (np.arange(rc.nx) - 0.5 * (rc.nx - 1)) * rc.recon_fov / rc.nx + rc.x_origin
    y = (np.arange(rc.ny) - 0.5 * (rc.ny - 1)) * rc.recon_fov / rc.ny + rc.y_origin
    X, Y = np.meshgrid(x, y)
    return X * X + Y * Y <= (0.5 * acq_fov - margin) ** 2


def hu_errors(a, b, mask=None):
    d = (np.asarray(a, np.float64) - np.asarray(b, np.float64)) / HU
    if mask is not None:
        d = d[..., mask]
    return float(np.sqrt(np.mean(d * d))), float(np.max(np.abs(d)))


def say(*a):
    print(*a, flush=True)


def gpu_ctx():
    from freect_amd import wfbp
    return wfbp.Context(0)


def to_np(t):
    return t.detach().cpu().numpy()
