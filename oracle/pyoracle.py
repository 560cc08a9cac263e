"""ctypes binding of the FP64 CPU restatement -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.  The restatement itself is
oracle/wfbp_oracle.cpp (reference line map in oracle/wfbp_oracle.h).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class OrScan(C.Structure):
    _fields_ = [("n_channels", C.c_int32), ("n_rows", C.c_int32),
                ("n_views", C.c_int32), ("views_per_rotation", C.c_int32),
                ("channel_spacing", C.c_double), ("row_spacing", C.c_double),
                ("dsd", C.c_double), ("r", C.c_double),
                ("central_channel", C.c_double), ("central_row", C.c_double),
                ("angle0", C.c_double), ("z0", C.c_double), ("feed", C.c_double),
                ("ffs_mode", C.c_int32), ("rotation", C.c_int32),
                ("anode_angle", C.c_double)]


class OrRecon(C.Structure):
    _fields_ = [("start_pos", C.c_double), ("end_pos", C.c_double),
                ("recon_fov", C.c_double), ("slice_thickness", C.c_double),
                ("slice_pitch", C.c_double), ("nx", C.c_int32), ("ny", C.c_int32),
                ("x_origin", C.c_double), ("y_origin", C.c_double),
                ("filter_kind", C.c_int32), ("filter_c", C.c_double),
                ("filter_a", C.c_double)]


class OrDerived(C.Structure):
    _fields_ = [("nfft", C.c_int32), ("csw", C.c_double), ("zrot", C.c_double),
                ("theta_cone", C.c_double), ("acq_fov", C.c_double), ("ds", C.c_double),
                ("half_acq_fov_sq", C.c_double), ("z_rot_over_2pi", C.c_double),
                ("tan_half_cone", C.c_double), ("pixel_scale", C.c_double),
                ("two_pi_over_ppr", C.c_double), ("views_per_half_turn", C.c_int32),
                ("n_half_turns", C.c_int32), ("n_slices_native", C.c_int32),
                ("n_slices_requested", C.c_int32)]


class OrEllipsoid(C.Structure):
    _fields_ = [("c", C.c_double * 3), ("a", C.c_double * 3), ("rot", C.c_double * 9),
                ("dmu", C.c_double)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp, fp, i64p, i32p = P(C.c_double), P(C.c_float), P(C.c_int64), P(C.c_int32)
        sig = {
            "or_geometry": (C.c_int, [P(OrScan), P(OrRecon), P(OrDerived)]),
            "or_effective_scan": (None, [P(OrScan), P(OrScan)]),
            "or_weights": (None, [fp, C.c_int, fp, dp]),
            "or_row_weights": (None, [C.c_int, fp, C.c_int, fp, fp]),
            "or_slice_grid": (C.c_int, [C.c_double, C.c_double, C.c_double, dp, C.c_int]),
            "or_filter_kernel": (None, [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, dp]),
            "or_rebin": (None, [P(OrScan), fp, dp, C.c_int, C.c_int]),
            "or_rebin_entries": (C.c_int, [P(OrScan), i32p, dp]),
            "or_adaptive_filter": (None, [P(OrScan), C.c_double, fp, fp]),
            "or_ray_project": (C.c_int, [P(OrScan), P(OrRecon), dp, dp]),
            "or_ray_backproject": (C.c_int, [P(OrScan), P(OrRecon), dp, dp]),
            "or_filter_rows": (None, [dp, C.c_int64, C.c_int, dp, C.c_int, C.c_int]),
            "or_backproject": (C.c_int, [P(OrScan), P(OrRecon), dp, C.c_int, C.c_int, C.c_int,
                                         C.c_int, dp, i64p]),
            "or_backproject_f32": (C.c_int, [P(OrScan), P(OrRecon), fp, C.c_int, C.c_int,
                                             C.c_int, C.c_int, fp, i64p]),
            "or_thicken": (C.c_int, [P(OrScan), P(OrRecon), dp, dp]),
            "or_bp_counts": (C.c_int, [P(OrScan), P(OrRecon), C.c_int, C.c_int, i32p, i64p]),
            "or_bp_coverage": (C.c_int, [P(OrScan), P(OrRecon), C.c_int, C.c_int, dp]),
            "or_bp_selection": (C.c_int, [P(OrScan), P(OrRecon), C.c_int, C.c_int, C.c_int, i32p]),
            "or_project_phantom": (None, [P(OrScan), C.c_double, C.c_double, P(OrEllipsoid),
                                          C.c_int, fp, C.c_int, C.c_int]),
            "or_reconstruct_native": (C.c_int, [P(OrScan), P(OrRecon), fp, C.c_int, C.c_int,
                                                dp, i64p]),
            "or_view_range": (None, [P(OrScan), P(OrRecon), C.c_int, C.c_int, P(C.c_int),
                                     P(C.c_int)]),
            "or_backproject_adjoint": (C.c_int, [P(OrScan), P(OrRecon), dp, C.c_int, C.c_int,
                                                 C.c_int, C.c_int, dp]),
            "or_rebin_adjoint": (None, [P(OrScan), dp, dp, C.c_int, C.c_int]),
            "or_adjoint_native": (C.c_int, [P(OrScan), P(OrRecon), dp, C.c_int, C.c_int, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ctype):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


def scan_struct(scan) -> OrScan:
    return OrScan(scan.n_channels, scan.n_rows, scan.n_views, scan.views_per_rotation,
                  scan.channel_spacing, scan.row_spacing, scan.dsd, scan.r,
                  scan.central_channel, scan.central_row, scan.angle0, scan.z0, scan.feed,
                  getattr(scan, "ffs_mode", 0), getattr(scan, "rotation", 1),
                  getattr(scan, "anode_angle", 0.0))


def _cw(scan) -> bool:
    return getattr(scan, "rotation", 1) < 0


def _mirror(scan, rc):
    """A clockwise scan is the mirror image y -> -y of the counter-clockwise
    scan with tube angles -a_v: reconstruct that one on the grid mirrored in
    y (y_origin -> -y_origin), whose row ny-1-i is row i of the clockwise
    grid.  Returns the counter-clockwise (scan, recon) and whether the rows
    of the result must be reversed."""
    if not _cw(scan):
        return scan, rc, False
    ms = scan.replace(angle0=-scan.angle0, rotation=1)
    mr = rc.replace(y_origin=-rc.y_origin) if rc is not None else None
    return ms, mr, True


def _flip_rows(a, flip):
    return np.ascontiguousarray(a[..., ::-1, :]) if flip else a


def _eff(scan):
    """Parallel-domain functions work on the interleaved detector of an FFS
    scan (ScanGeometry.effective(); checked against or_effective_scan)."""
    return scan.effective() if getattr(scan, "ffs_mode", 0) else scan


def weights(q):
    """The shared FP32 helical weight W(q) (wfbp_math.h) and its FP64
    definition (cuda_kernels.cuh:79-93) at the FP32 points q."""
    q = np.ascontiguousarray(q, dtype=np.float32)
    w32 = np.empty(q.shape, np.float32)
    w64 = np.empty(q.shape, np.float64)
    lib().or_weights(q.ctypes.data_as(C.POINTER(C.c_float)), int(q.size),
                     w32.ctypes.data_as(C.POINTER(C.c_float)),
                     w64.ctypes.data_as(C.POINTER(C.c_double)))
    return w32, w64


def row_weights(nrows, qi):
    """The shared FP32 weight from row positions qi for nrows centred rows:
    (selection-masked fct_w_of_row, unmasked with t clamped to [0, 1])."""
    qi = np.ascontiguousarray(qi, dtype=np.float32)
    ws = np.empty(qi.shape, np.float32)
    wc = np.empty(qi.shape, np.float32)
    fp = C.POINTER(C.c_float)
    lib().or_row_weights(int(nrows), qi.ctypes.data_as(fp), int(qi.size), ws.ctypes.data_as(fp),
                         wc.ctypes.data_as(fp))
    return ws, wc


def effective_scan(scan) -> dict:
    """The interleaved geometry as computed by the C++ restatement."""
    out = OrScan()
    s = scan_struct(scan)
    lib().or_effective_scan(C.byref(s), C.byref(out))
    return {k: getattr(out, k) for k, _ in OrScan._fields_}


def recon_struct(rc) -> OrRecon:
    return OrRecon(rc.start_pos, rc.end_pos, rc.recon_fov, rc.slice_thickness,
                   rc.slice_pitch, rc.nx, rc.ny, rc.x_origin, rc.y_origin,
                   rc.filter_kind, rc.filter_c, rc.filter_a)


def geometry(scan, rc=None) -> OrDerived:
    d = OrDerived()
    s = scan_struct(_eff(scan))
    r = recon_struct(rc) if rc is not None else None
    lib().or_geometry(C.byref(s), C.byref(r) if r is not None else None, C.byref(d))
    return d


def slice_grid(start, end, step) -> np.ndarray:
    n = lib().or_slice_grid(start, end, step, None, 0)
    z = np.zeros(max(n, 1), np.float64)
    lib().or_slice_grid(start, end, step, _ptr(z, C.c_double), n)
    return z[:n]


def filter_kernel(n, ds, kind=0, c=1.0, a=1.0) -> np.ndarray:
    h = np.zeros(n, np.float64)
    lib().or_filter_kernel(n, ds, kind, c, a, _ptr(h, C.c_double))
    return h


def rebin(scan, raw: np.ndarray, v0=0, v1=None) -> np.ndarray:
    """Raw fan data of `scan` -> parallel views [v0, v1) of its interleaved
    detector."""
    e = _eff(scan)
    v1 = e.n_views if v1 is None else v1
    raw = np.ascontiguousarray(raw, np.float32)
    out = np.zeros((v1 - v0, e.n_rows, e.n_channels), np.float64)
    s = scan_struct(scan)
    lib().or_rebin(C.byref(s), _ptr(raw, C.c_float), _ptr(out, C.c_double), v0, v1)
    return out


def adaptive_filter(scan, raw: np.ndarray, strength: float) -> np.ndarray:
    """Adaptive filtration of a raw scan [view][row][ch] (float32 out)."""
    raw = np.ascontiguousarray(raw, np.float32)
    out = np.empty_like(raw)
    s = scan_struct(scan)
    lib().or_adaptive_filter(C.byref(s), float(strength), _ptr(raw, C.c_float),
                             _ptr(out, C.c_float))
    return out


def ray_project(scan, rc, vol: np.ndarray) -> np.ndarray:
    """Joseph line integrals of a native volume along the raw rays (FP64)."""
    vol = np.ascontiguousarray(vol, np.float64)
    raw = np.zeros((scan.n_views, scan.n_rows, scan.n_channels), np.float64)
    s, r = scan_struct(scan), recon_struct(rc)
    if lib().or_ray_project(C.byref(s), C.byref(r), _ptr(vol, C.c_double), _ptr(raw, C.c_double)):
        raise RuntimeError("or_ray_project failed")
    return raw


def ray_backproject(scan, rc, raw: np.ndarray, n_slices: int) -> np.ndarray:
    """Exact transpose of ray_project."""
    raw = np.ascontiguousarray(raw, np.float64)
    vol = np.zeros((n_slices, rc.ny, rc.nx), np.float64)
    s, r = scan_struct(scan), recon_struct(rc)
    if lib().or_ray_backproject(C.byref(s), C.byref(r), _ptr(raw, C.c_double),
                                _ptr(vol, C.c_double)):
        raise RuntimeError("or_ray_backproject failed")
    return vol


def rebin_entries(scan):
    """The rebinning table of the scan (one row parity, two with z FFS):
    int fields (valid, rho, b0, b1, aoff, mmax) and FP64 (fb, fa) per entry."""
    s = scan_struct(scan)
    n = lib().or_rebin_entries(C.byref(s), None, None)
    ient = np.zeros((n, 6), np.int32)
    fent = np.zeros((n, 2), np.float64)
    lib().or_rebin_entries(C.byref(s), _ptr(ient, C.c_int32), _ptr(fent, C.c_double))
    return ient, fent


def filter_rows(rows: np.ndarray, h: np.ndarray, nfft: int, use_fft=True) -> np.ndarray:
    out = np.ascontiguousarray(rows, np.float64).copy()
    ncols = out.shape[-1]
    lib().or_filter_rows(_ptr(out, C.c_double), out.size // ncols, ncols,
                         _ptr(np.ascontiguousarray(h, np.float64), C.c_double), nfft,
                         1 if use_fft else 0)
    return out


def backproject(scan, rc, filtered: np.ndarray, v0: int, s0: int, s1: int):
    scan, rc, flip = _mirror(scan, rc)
    vol, n = _backproject(scan, rc, filtered, v0, s0, s1)
    return _flip_rows(vol, flip), n


def _backproject(scan, rc, filtered: np.ndarray, v0: int, s0: int, s1: int):
    filtered = np.ascontiguousarray(filtered, np.float64)
    v1 = v0 + filtered.shape[0]
    vol = np.zeros((s1 - s0, rc.ny, rc.nx), np.float64)
    n = C.c_int64(0)
    s, r = scan_struct(_eff(scan)), recon_struct(rc)
    rc_ = lib().or_backproject(C.byref(s), C.byref(r), _ptr(filtered, C.c_double), v0, v1,
                               s0, s1, _ptr(vol, C.c_double), C.byref(n))
    if rc_ != 0:
        raise RuntimeError(f"or_backproject failed ({rc_})")
    return vol, n.value


def backproject_f32(scan, rc, filtered: np.ndarray, v0: int, s0: int, s1: int):
    scan, rc, flip = _mirror(scan, rc)
    filtered = np.ascontiguousarray(filtered, np.float32)
    v1 = v0 + filtered.shape[0]
    vol = np.zeros((s1 - s0, rc.ny, rc.nx), np.float32)
    n = C.c_int64(0)
    s, r = scan_struct(_eff(scan)), recon_struct(rc)
    rc_ = lib().or_backproject_f32(C.byref(s), C.byref(r), _ptr(filtered, C.c_float), v0,
                                   v1, s0, s1, _ptr(vol, C.c_float), C.byref(n))
    if rc_ != 0:
        raise RuntimeError(f"or_backproject_f32 failed ({rc_})")
    return _flip_rows(vol, flip), n.value


def thicken(scan, rc, native: np.ndarray) -> np.ndarray:
    d = geometry(scan, rc)
    native = np.ascontiguousarray(native, np.float64)
    out = np.zeros((d.n_slices_requested, rc.ny, rc.nx), np.float64)
    s, r = scan_struct(_eff(scan)), recon_struct(rc)
    lib().or_thicken(C.byref(s), C.byref(r), _ptr(native, C.c_double), _ptr(out, C.c_double))
    return out


def bp_counts(scan, rc, s0, s1):
    scan, rc, flip = _mirror(scan, rc)
    counts = np.zeros((s1 - s0, rc.ny, rc.nx), np.int32)
    n = C.c_int64(0)
    s, r = scan_struct(_eff(scan)), recon_struct(rc)
    lib().or_bp_counts(C.byref(s), C.byref(r), s0, s1, _ptr(counts, C.c_int32), C.byref(n))
    return _flip_rows(counts, flip), n.value


def bp_coverage(scan, rc, s0, s1) -> np.ndarray:
    """Per native voxel: min over theta~ of the weight sum of selected views."""
    scan, rc, flip = _mirror(scan, rc)
    out = np.zeros((s1 - s0, rc.ny, rc.nx), np.float64)
    s, r = scan_struct(_eff(scan)), recon_struct(rc)
    lib().or_bp_coverage(C.byref(s), C.byref(r), s0, s1, _ptr(out, C.c_double))
    return _flip_rows(out, flip)


def covered_mask(scan, rc, threshold=0.5) -> np.ndarray:
    """Requested-grid voxels whose contributing native slices are all fully
    covered (every theta~ has weight sum >= threshold)."""
    d = geometry(scan, rc)
    cov = bp_coverage(scan, rc, 0, d.n_slices_native) >= threshold
    zn = slice_grid(rc.start_pos, rc.end_pos, d.csw)
    pitch = rc.slice_pitch if rc.slice_pitch > 0 else rc.slice_thickness
    zr = slice_grid(rc.start_pos, rc.end_pos, pitch)
    out = np.ones((len(zr), rc.ny, rc.nx), bool)
    for s, z in enumerate(zr):
        used = np.abs(zn - z) < rc.slice_thickness
        out[s] = np.all(cov[used], axis=0) if used.any() else False
    return out


def bp_selection(scan, rc, ix, iy, iz) -> np.ndarray:
    scan, rc, flip = _mirror(scan, rc)
    if flip:
        iy = rc.ny - 1 - iy
    sel = np.zeros(_eff(scan).n_views, np.int32)
    s, r = scan_struct(_eff(scan)), recon_struct(rc)
    lib().or_bp_selection(C.byref(s), C.byref(r), ix, iy, iz, _ptr(sel, C.c_int32))
    return sel


def ellipsoid_array(phantom):
    arr = (OrEllipsoid * max(1, len(phantom.ellipsoids)))()
    for i, e in enumerate(phantom.ellipsoids):
        arr[i].c[:] = list(e["c"])
        arr[i].a[:] = list(e["a"])
        arr[i].rot[:] = list(np.asarray(e["rot"], np.float64).reshape(9))
        arr[i].dmu = e["dmu"]
    return arr


def project_phantom(scan, phantom, v0=0, v1=None) -> np.ndarray:
    v1 = scan.n_views if v1 is None else v1
    out = np.zeros((v1 - v0, scan.n_rows, scan.n_channels), np.float32)
    arr = ellipsoid_array(phantom)
    s = scan_struct(scan)
    lib().or_project_phantom(C.byref(s), phantom.cyl_radius, phantom.cyl_mu, arr,
                             len(phantom.ellipsoids), _ptr(out, C.c_float), v0, v1)
    return out


def view_range(scan, rc, s0, s1):
    a, b = C.c_int(0), C.c_int(0)
    s, r = scan_struct(_eff(scan)), recon_struct(rc)
    lib().or_view_range(C.byref(s), C.byref(r), s0, s1, C.byref(a), C.byref(b))
    return a.value, b.value


def reconstruct_native(scan, rc, raw: np.ndarray, s0: int, s1: int):
    scan, rc, flip = _mirror(scan, rc)
    raw = np.ascontiguousarray(raw, np.float32)
    vol = np.zeros((s1 - s0, rc.ny, rc.nx), np.float64)
    n = C.c_int64(0)
    s, r = scan_struct(scan), recon_struct(rc)
    rc_ = lib().or_reconstruct_native(C.byref(s), C.byref(r), _ptr(raw, C.c_float), s0, s1,
                                      _ptr(vol, C.c_double), C.byref(n))
    if rc_ != 0:
        raise RuntimeError(f"or_reconstruct_native failed ({rc_})")
    return _flip_rows(vol, flip), n.value


def backproject_adjoint(scan, rc, xbar: np.ndarray, v0: int, v1: int, s0: int) -> np.ndarray:
    scan, rc, flip = _mirror(scan, rc)
    xbar = np.ascontiguousarray(_flip_rows(np.asarray(xbar), flip), np.float64)
    s1 = s0 + xbar.shape[0]
    e = _eff(scan)
    g = np.zeros((v1 - v0, e.n_rows, e.n_channels), np.float64)
    s, r = scan_struct(e), recon_struct(rc)
    st = lib().or_backproject_adjoint(C.byref(s), C.byref(r), _ptr(xbar, C.c_double), v0, v1,
                                      s0, s1, _ptr(g, C.c_double))
    if st:
        raise RuntimeError(f"or_backproject_adjoint failed ({st})")
    return g


def rebin_adjoint(scan, gbar: np.ndarray, v0: int = 0) -> np.ndarray:
    gbar = np.ascontiguousarray(gbar, np.float64)
    v1 = v0 + gbar.shape[0]
    p = np.zeros((scan.n_views, scan.n_rows, scan.n_channels), np.float64)
    s = scan_struct(scan)
    lib().or_rebin_adjoint(C.byref(s), _ptr(gbar, C.c_double), _ptr(p, C.c_double), v0, v1)
    return p


def adjoint_native(scan, rc, xbar: np.ndarray, s0: int = 0) -> np.ndarray:
    scan, rc, flip = _mirror(scan, rc)
    xbar = np.ascontiguousarray(_flip_rows(np.asarray(xbar), flip), np.float64)
    s1 = s0 + xbar.shape[0]
    p = np.zeros((scan.n_views, scan.n_rows, scan.n_channels), np.float64)
    s, r = scan_struct(scan), recon_struct(rc)
    st = lib().or_adjoint_native(C.byref(s), C.byref(r), _ptr(xbar, C.c_double), s0, s1,
                                 _ptr(p, C.c_double))
    if st:
        raise RuntimeError(f"or_adjoint_native failed ({st})")
    return p
