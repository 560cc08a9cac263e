"""Python mirror of the fct_wfbp_gpu C ABI (ctypes).

This is the host-side mirror of FreeCT's ReconstructionMachine stages
(ReconstructionMachine.h:16-21) for the test harness and the benchmark:
``Plan.rebin_filter``, ``Plan.backproject``, ``Plan.thicken`` and
``Plan.reconstruct`` call the HIP kernels through the C ABI declared in
include/fct_wfbp_gpu.h.  PyTorch only provides device memory and the stream.
There is no CPU fallback: a missing library raises at import of a plan.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import scans as S

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libfct_wfbp_gpu.so")

PRECISION = {"fp32": 0, "fp16": 1, "bf16": 2}


class FctError(RuntimeError):
    pass


class ScanGeometryC(C.Structure):
    _fields_ = [("n_channels", C.c_int32), ("n_rows", C.c_int32), ("n_views", C.c_int32),
                ("views_per_rotation", C.c_int32), ("channel_spacing", C.c_double),
                ("row_spacing", C.c_double), ("dsd", C.c_double), ("r", C.c_double),
                ("central_channel", C.c_double), ("central_row", C.c_double),
                ("angle0", C.c_double), ("z0", C.c_double), ("feed", C.c_double),
                ("ffs_mode", C.c_int32), ("rotation", C.c_int32),
                ("anode_angle", C.c_double)]


class ReconConfigC(C.Structure):
    _fields_ = [("start_pos", C.c_double), ("end_pos", C.c_double), ("recon_fov", C.c_double),
                ("slice_thickness", C.c_double), ("slice_pitch", C.c_double),
                ("nx", C.c_int32), ("ny", C.c_int32), ("x_origin", C.c_double),
                ("y_origin", C.c_double), ("recon_kernel", C.c_int32),
                ("filter_kind", C.c_int32), ("filter_c", C.c_double), ("filter_a", C.c_double),
                ("tube_angle_offset", C.c_double), ("adaptive_filtration_s", C.c_double)]


class ExecOptionsC(C.Structure):
    _fields_ = [("precision", C.c_int32), ("tile_z", C.c_int32), ("bp_variant", C.c_int32),
                ("adaptive_filtration", C.c_int32)]


class PlanInfoC(C.Structure):
    _fields_ = [("n_slices_native", C.c_int32), ("n_slices_requested", C.c_int32),
                ("nfft", C.c_int32), ("channel_pitch", C.c_int32), ("raw_elems", C.c_int64),
                ("filtered_bytes", C.c_int64), ("scale_elems", C.c_int64),
                ("native_elems", C.c_int64), ("volume_elems", C.c_int64),
                ("collimated_slice_width", C.c_double), ("acquisition_fov", C.c_double),
                ("z_rot", C.c_double), ("tile_z", C.c_int32), ("bp_variant", C.c_int32),
                ("n_parallel_views", C.c_int32), ("n_filtered_rows", C.c_int32),
                ("n_filtered_channels", C.c_int32), ("ffs_states", C.c_int32),
                ("read_views_first", C.c_int32), ("read_views_end", C.c_int32)]


class BatchInfoC(C.Structure):
    _fields_ = [("n_jobs", C.c_int32), ("tile_z", C.c_int32), ("total_tiles", C.c_int32),
                ("grid", C.c_int32), ("device_bytes", C.c_int64)]


class EllipsoidC(C.Structure):
    _fields_ = [("c", C.c_double * 3), ("a", C.c_double * 3), ("rot", C.c_double * 9),
                ("dmu", C.c_double)]


# Every exported entry point and its ctypes signature.
_P, _VP = C.POINTER, C.c_void_p
SIGNATURES = {
    "fct_wfbp_gpu_create": (C.c_int, [C.c_int32, _P(_VP)]),
    "fct_wfbp_gpu_destroy": (C.c_int, [_VP]),
    "fct_wfbp_gpu_last_error": (C.c_char_p, [_VP]),
    "fct_wfbp_gpu_plan_create": (C.c_int, [_VP, _P(ScanGeometryC), _P(ReconConfigC),
                                           _P(ExecOptionsC), _VP, _P(_VP)]),
    "fct_wfbp_gpu_plan_destroy": (C.c_int, [_VP]),
    "fct_wfbp_gpu_plan_info": (C.c_int, [_VP, _P(PlanInfoC)]),
    "fct_wfbp_gpu_set_stream": (C.c_int, [_VP, _VP]),
    "fct_wfbp_gpu_rebin_filter": (C.c_int, [_VP, _VP, _VP, _VP]),
    "fct_wfbp_gpu_backproject": (C.c_int, [_VP, _VP, _VP, _VP]),
    "fct_wfbp_gpu_thicken": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_reconstruct": (C.c_int, [_VP, _VP, _VP, C.c_int32]),
    "fct_wfbp_gpu_reconstruct_stream": (C.c_int, [_VP, _VP, _VP, C.c_int32, C.c_float, _VP]),
    "fct_wfbp_gpu_reconstruct_noalloc": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_set_output_units": (C.c_int, [_VP, C.c_int32, C.c_double]),
    "fct_wfbp_gpu_ray_project": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_ray_backproject": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_backproject_adjoint": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_filter": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_rebin_adjoint": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_thicken_adjoint": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_adjoint": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_count_voxel_views": (C.c_int, [_VP, _VP, _P(C.c_int64)]),
    "fct_wfbp_gpu_voxel_selection": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, _VP]),
    "fct_wfbp_gpu_synchronize": (C.c_int, [_VP]),
    "fct_wfbp_gpu_phantom_project": (C.c_int, [_VP, _P(ScanGeometryC), C.c_double, C.c_double,
                                               _P(EllipsoidC), C.c_int32, C.c_double,
                                               C.c_uint64, C.c_int32, C.c_int32, _VP, _VP]),
    "fct_wfbp_gpu_rebin": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_matched_project": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_matched_backproject": (C.c_int, [_VP, _VP, _VP]),
    "fct_wfbp_gpu_batch_create": (C.c_int, [_VP, C.c_int32, _P(ScanGeometryC),
                                            _P(ReconConfigC), _P(ExecOptionsC), _VP, _P(_VP)]),
    "fct_wfbp_gpu_batch_destroy": (C.c_int, [_VP]),
    "fct_wfbp_gpu_batch_submit": (C.c_int, [_VP, C.c_int32, _VP, _VP]),
    "fct_wfbp_gpu_batch_run": (C.c_int, [_VP, C.c_int32]),
    "fct_wfbp_gpu_batch_synchronize": (C.c_int, [_VP]),
    "fct_wfbp_gpu_batch_backproject": (C.c_int, [_VP]),
    "fct_wfbp_gpu_batch_autotune": (C.c_int, [_VP, _P(C.c_int32), C.c_int32, C.c_int32,
                                              _P(C.c_float)]),
    "fct_wfbp_gpu_batch_info": (C.c_int, [_VP, _P(BatchInfoC)]),
    "fct_wfbp_gpu_batch_job_plan": (C.c_int, [_VP, C.c_int32, _P(_VP)]),
}

_lib = None


def load_library():
    """Load libfct_wfbp_gpu.so; raises FctError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FctError(f"{LIB_PATH} is missing: run __graft_entry__.build() first "
                           "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def scan_struct(scan: S.ScanGeometry) -> ScanGeometryC:
    return ScanGeometryC(scan.n_channels, scan.n_rows, scan.n_views, scan.views_per_rotation,
                         scan.channel_spacing, scan.row_spacing, scan.dsd, scan.r,
                         scan.central_channel, scan.central_row, scan.angle0, scan.z0, scan.feed,
                         scan.ffs_mode, scan.rotation, scan.anode_angle)


def recon_struct(rc: S.ReconConfig) -> ReconConfigC:
    return ReconConfigC(rc.start_pos, rc.end_pos, rc.recon_fov, rc.slice_thickness,
                        rc.slice_pitch, rc.nx, rc.ny, rc.x_origin, rc.y_origin, 0,
                        rc.filter_kind, rc.filter_c, rc.filter_a, 0.0,
                        getattr(rc, "adaptive_filtration_s", 0.0))


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


_DTYPE_NAMES = {"f32": "float32", "i32": "int32"}


def _dev_arg(t, what: str, device: int, *, numel: Optional[int] = None,
             nbytes: Optional[int] = None, dtype: Optional[str] = None,
             optional: bool = False) -> Optional[int]:
    """Checks a device-buffer argument before its address crosses the C ABI
    (whose kernels trust the sizes of the plan): a contiguous CUDA tensor on
    the plan's device, of the given element type ('f32' / 'i32') and element
    count, or of exactly `nbytes` bytes for the opaque filtered buffer.  An
    int is taken as a raw device address and passed through unchecked (the
    C ABI's own contract); None is allowed only where the C ABI allows NULL."""
    if t is None:
        if optional:
            return None
        raise FctError(f"{what}: a device buffer is required")
    if isinstance(t, int):
        return t
    import torch
    if not isinstance(t, torch.Tensor):
        raise FctError(f"{what}: expected a torch tensor or a device address, got {type(t)}")
    if dtype is not None:
        want = {"f32": torch.float32, "i32": torch.int32}[dtype]
        if t.dtype != want:
            raise FctError(f"{what}: dtype must be {_DTYPE_NAMES[dtype]} (got {t.dtype})")
    if not t.is_contiguous():
        raise FctError(f"{what}: must be contiguous")
    if numel is not None and t.numel() != numel:
        raise FctError(f"{what}: {t.numel()} elements, the plan needs {numel}")
    if nbytes is not None and t.numel() * t.element_size() != nbytes:
        raise FctError(f"{what}: {t.numel() * t.element_size()} bytes, the plan needs {nbytes}")
    if not t.is_cuda or t.device.index != device:
        raise FctError(f"{what}: must live on cuda:{device} (got {t.device})")
    return t.data_ptr()


def _stream_handle(stream=None) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class Context:
    def __init__(self, device: int = 0):
        L = load_library()
        h = C.c_void_p()
        st = L.fct_wfbp_gpu_create(device, C.byref(h))
        if st != 0:
            raise FctError(f"fct_wfbp_gpu_create({device}) failed with status {st}")
        self.h = h
        self.device = device

    def check(self, st: int, what: str):
        if st != 0:
            msg = load_library().fct_wfbp_gpu_last_error(self.h)
            raise FctError(f"{what}: status {st}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            load_library().fct_wfbp_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def phantom_project(self, scan: S.ScanGeometry, phantom: S.Phantom, i0: float = 0.0,
                        seed: int = 0, out=None, v0: int = 0, v1: Optional[int] = None,
                        stream=None):
        """Seeded synthetic scan on the GPU: [view][row][ch] float32 tensor."""
        import torch
        v1 = scan.n_views if v1 is None else v1
        if out is None:
            out = torch.empty((v1 - v0, scan.n_rows, scan.n_channels), dtype=torch.float32,
                              device=f"cuda:{self.device}")
        n = len(phantom.ellipsoids)
        arr = (EllipsoidC * max(1, n))()
        for i, e in enumerate(phantom.ellipsoids):
            arr[i].c[:] = list(e["c"])
            arr[i].a[:] = list(e["a"])
            arr[i].rot[:] = list(np.asarray(e["rot"], np.float64).reshape(9))
            arr[i].dmu = e["dmu"]
        sc = scan_struct(scan)
        st = load_library().fct_wfbp_gpu_phantom_project(
            self.h, C.byref(sc), phantom.cyl_radius, phantom.cyl_mu, arr, n, float(i0),
            C.c_uint64(seed & ((1 << 64) - 1)), v0, v1, out.data_ptr(), _stream_handle(stream))
        self.check(st, "phantom_project")
        return out


class Plan:
    """One reconstruction setup (geometry + recon config + execution options)."""

    def __init__(self, ctx: Context, scan: S.ScanGeometry, recon: S.ReconConfig,
                 precision: str = "fp32", tile_z: int = 0, bp_variant: int = 0, stream=None,
                 adaptive_filtration: bool = False):
        self.ctx = ctx
        self.scan = scan
        self.recon = recon
        self.precision = precision
        sc, rc = scan_struct(scan), recon_struct(recon)
        opts = ExecOptionsC(PRECISION[precision], tile_z, bp_variant,
                            1 if adaptive_filtration else 0)
        h = C.c_void_p()
        st = load_library().fct_wfbp_gpu_plan_create(ctx.h, C.byref(sc), C.byref(rc),
                                                     C.byref(opts), _stream_handle(stream),
                                                     C.byref(h))
        ctx.check(st, "plan_create")
        self.h = h
        info = PlanInfoC()
        ctx.check(load_library().fct_wfbp_gpu_plan_info(self.h, C.byref(info)), "plan_info")
        self.info = info

    @classmethod
    def _borrowed(cls, ctx: Context, h, scan, recon, precision):
        """A view of a plan owned by a Batch (never destroyed from here)."""
        self = cls.__new__(cls)
        self.ctx, self.scan, self.recon, self.precision = ctx, scan, recon, precision
        self.h = h
        self._owner = False
        info = PlanInfoC()
        ctx.check(load_library().fct_wfbp_gpu_plan_info(self.h, C.byref(info)), "plan_info")
        self.info = info
        return self

    def close(self):
        if getattr(self, "h", None) and getattr(self, "_owner", True):
            load_library().fct_wfbp_gpu_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- buffers ------------------------------------------------------------
    def _dev(self):
        return f"cuda:{self.ctx.device}"

    def alloc_filtered(self):
        import torch
        i = self.info
        filt = torch.empty(i.filtered_bytes, dtype=torch.uint8, device=self._dev())
        scale = (torch.empty(i.scale_elems, dtype=torch.float32, device=self._dev())
                 if i.scale_elems else None)
        return filt, scale

    def alloc_native(self):
        import torch
        i = self.info
        return torch.empty((i.n_slices_native, self.recon.ny, self.recon.nx),
                           dtype=torch.float32, device=self._dev())

    def alloc_volume(self):
        import torch
        i = self.info
        return torch.empty((i.n_slices_requested, self.recon.ny, self.recon.nx),
                           dtype=torch.float32, device=self._dev())

    def filtered_view(self, filt, scale=None):
        """The filtered projections y as a [view][row][channel_pitch] tensor.

        FP32: a view of the buffer.  FP16 / BF16: the (r, c) slot of every
        quad entry, in the storage type, multiplied by the scan's storage
        scale (scale[0]) when the scale buffer is given."""
        import torch
        dt = {"fp32": torch.float32, "fp16": torch.float16, "bf16": torch.bfloat16}[self.precision]
        i = self.info
        shape = (i.n_parallel_views, i.n_filtered_rows, i.channel_pitch)
        if self.precision == "fp32":
            return filt.view(dt).view(shape)
        y = filt.view(dt).view(shape + (4,))[..., 0]
        return y if scale is None else y.float() * scale[0]

    def filtered_quads(self, filt):
        """FP16 / BF16 only: the quad entries as [view][row][channel_pitch][4]
        (y[r][c], y[r][c+1], y[r+1][c], y[r+1][c+1]) in units of the scale."""
        import torch
        if self.precision == "fp32":
            raise FctError("FP32 projections are stored as plain rows")
        dt = {"fp16": torch.float16, "bf16": torch.bfloat16}[self.precision]
        i = self.info
        return filt.view(dt).view(i.n_parallel_views, i.n_filtered_rows, i.channel_pitch, 4)

    # -- stages -------------------------------------------------------------
    def set_stream(self, stream=None):
        self.ctx.check(load_library().fct_wfbp_gpu_set_stream(self.h, _stream_handle(stream)),
                       "set_stream")

    # Checked arguments: every buffer against the plan's sizes (_dev_arg).
    def _raw(self, t, what="raw"):
        return _dev_arg(t, what, self.ctx.device, numel=self.info.raw_elems, dtype="f32")

    def _filt(self, t, what="filtered"):
        return _dev_arg(t, what, self.ctx.device, nbytes=self.info.filtered_bytes)

    def _scale(self, t, what="scale"):
        return _dev_arg(t, what, self.ctx.device, numel=self.info.scale_elems or None,
                        dtype="f32", optional=True)

    def _native(self, t, what="native"):
        return _dev_arg(t, what, self.ctx.device, numel=self.info.native_elems, dtype="f32")

    def _volume(self, t, what="volume"):
        return _dev_arg(t, what, self.ctx.device, numel=self.info.volume_elems, dtype="f32")

    def _lines(self, t, what="lines"):
        i = self.info
        return _dev_arg(t, what, self.ctx.device,
                        numel=i.n_parallel_views * i.n_filtered_rows * i.channel_pitch,
                        dtype="f32")

    def rebin_filter(self, raw, filt, scale=None):
        self.ctx.check(load_library().fct_wfbp_gpu_rebin_filter(
            self.h, self._raw(raw), self._filt(filt), self._scale(scale)), "rebin_filter")

    def backproject(self, filt, scale, native):
        self.ctx.check(load_library().fct_wfbp_gpu_backproject(
            self.h, self._filt(filt), self._scale(scale), self._native(native)), "backproject")

    def thicken(self, native, volume):
        self.ctx.check(load_library().fct_wfbp_gpu_thicken(
            self.h, self._native(native), self._volume(volume)), "thicken")

    def reconstruct(self, raw, volume, host_buffers: bool = False):
        if host_buffers:
            raise FctError("reconstruct(host_buffers=True) takes numpy arrays: use "
                           "reconstruct_host()")
        self.ctx.check(load_library().fct_wfbp_gpu_reconstruct(
            self.h, self._raw(raw), self._volume(volume), 0), "reconstruct")

    def set_output_units(self, units: str = "mu", mu_water: float = S.MU_WATER):
        """Units of the thickened volume: "mu" (1/mm, default) or "hu"."""
        self.ctx.check(load_library().fct_wfbp_gpu_set_output_units(
            self.h, {"mu": 0, "hu": 1}[units], float(mu_water)), "set_output_units")

    def reconstruct_noalloc(self, raw, volume):
        """Whole chain on device buffers without any device allocation (the
        library's fixed per-GPU scratch; FctError with FCT_ERR_LIMIT when the
        plan does not fit it)."""
        self.ctx.check(load_library().fct_wfbp_gpu_reconstruct_noalloc(
            self.h, self._raw(raw), self._volume(volume)), "reconstruct_noalloc")

    def _host_volume(self, out):
        shape = (self.info.n_slices_requested, self.recon.ny, self.recon.nx)
        if out is None:
            return np.empty(shape, np.float32)
        if (not isinstance(out, np.ndarray) or out.dtype != np.float32 or out.shape != shape
                or not out.flags["C_CONTIGUOUS"]):
            raise FctError(f"volume: expected a C-contiguous float32 array of shape {shape}")
        return out

    def reconstruct_host(self, raw, out: Optional[np.ndarray] = None) -> np.ndarray:
        """Whole chain from a host scan (numpy array, or a CPU torch tensor --
        pinned memory makes the chunked copies asynchronous DMA) to a host
        volume, streamed in chunks of views (fct_wfbp_gpu_reconstruct with
        host_buffers = 1)."""
        if hasattr(raw, "data_ptr"):   # torch CPU tensor
            if raw.is_cuda or raw.dtype.itemsize != 4 or not raw.is_contiguous() or \
                    str(raw.dtype) != "torch.float32":
                raise FctError("raw: expected a contiguous float32 CPU tensor")
            if raw.numel() != self.info.raw_elems:
                raise FctError(f"raw: expected {self.info.raw_elems} elements")
            ptr, keep = raw.data_ptr(), raw
        else:
            keep = np.ascontiguousarray(raw, np.float32)
            if keep.size != self.info.raw_elems:
                raise FctError(f"raw: expected {self.info.raw_elems} elements, got {keep.size}")
            ptr = keep.ctypes.data
        vol = self._host_volume(out)
        self.ctx.check(load_library().fct_wfbp_gpu_reconstruct(
            self.h, ptr, vol.ctypes.data, 1), "reconstruct(host)")
        del keep
        return vol

    READER = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_float))

    def reconstruct_stream(self, read_views, views_per_chunk: int = 0, max_abs_raw: float = 0.0,
                           out: Optional[np.ndarray] = None) -> np.ndarray:
        """Streaming whole chain (fct_wfbp_gpu_reconstruct_stream):
        read_views(v0, n, dst) fills views [v0, v0 + n) into dst, a float32
        numpy view of the plan's pinned staging, shape (n, rows, channels)."""
        rows, chans = self.scan.n_rows, self.scan.n_channels
        errors = []

        def _cb(_user, v0, n, dst):
            try:
                arr = np.ctypeslib.as_array(dst, shape=(n * rows * chans,)).reshape(n, rows, chans)
                read_views(int(v0), int(n), arr)
                return 0
            except Exception as e:   # reported after the call
                errors.append(e)
                return 1

        cb = self.READER(_cb)
        vol = self._host_volume(out)
        st = load_library().fct_wfbp_gpu_reconstruct_stream(
            self.h, C.cast(cb, C.c_void_p), None, int(views_per_chunk), float(max_abs_raw),
            vol.ctypes.data)
        if errors:
            raise errors[0]
        self.ctx.check(st, "reconstruct_stream")
        return vol

    # -- adjoint (transpose of rebin -> filter -> backproject, FP32) ----------
    def alloc_raw(self):
        import torch
        return torch.empty((self.scan.n_views, self.scan.n_rows, self.scan.n_channels),
                           dtype=torch.float32, device=self._dev())

    def alloc_lines(self):
        """FP32 [view][row][channel_pitch] buffer (filtered rows and their cotangents)."""
        import torch
        i = self.info
        return torch.empty((i.n_parallel_views, i.n_filtered_rows, i.channel_pitch),
                           dtype=torch.float32, device=self._dev())

    def backproject_adjoint(self, native_bar, lines_bar):
        self.ctx.check(load_library().fct_wfbp_gpu_backproject_adjoint(
            self.h, self._native(native_bar, "native_bar"), self._lines(lines_bar, "lines_bar")),
            "backproject_adjoint")

    def filter_lines(self, lines_in, lines_out):
        self.ctx.check(load_library().fct_wfbp_gpu_filter(
            self.h, self._lines(lines_in, "lines_in"), self._lines(lines_out, "lines_out")),
            "filter")

    def rebin_lines(self, raw, lines_out):
        """Rebinning without the filter: FP32 [view][row][channel_pitch]."""
        self.ctx.check(load_library().fct_wfbp_gpu_rebin(
            self.h, self._raw(raw), self._lines(lines_out, "lines_out")), "rebin")

    def ray_project(self, native, raw):
        """Line integrals of the native volume along every raw ray (Joseph)."""
        self.ctx.check(load_library().fct_wfbp_gpu_ray_project(
            self.h, self._native(native), self._raw(raw)), "ray_project")

    def ray_backproject(self, raw, native):
        """Exact transpose of ray_project (native overwritten)."""
        self.ctx.check(load_library().fct_wfbp_gpu_ray_backproject(
            self.h, self._raw(raw), self._native(native)), "ray_backproject")

    def matched_project(self, native, lines):
        self.ctx.check(load_library().fct_wfbp_gpu_matched_project(
            self.h, self._native(native), self._lines(lines)), "matched_project")

    def matched_backproject(self, lines, native):
        self.ctx.check(load_library().fct_wfbp_gpu_matched_backproject(
            self.h, self._lines(lines), self._native(native)), "matched_backproject")

    def rebin_adjoint(self, lines_bar, raw_bar):
        self.ctx.check(load_library().fct_wfbp_gpu_rebin_adjoint(
            self.h, self._lines(lines_bar, "lines_bar"), self._raw(raw_bar, "raw_bar")),
            "rebin_adjoint")

    def thicken_adjoint(self, volume_bar, native_bar):
        self.ctx.check(load_library().fct_wfbp_gpu_thicken_adjoint(
            self.h, self._volume(volume_bar, "volume_bar"), self._native(native_bar, "native_bar")),
            "thicken_adjoint")

    def adjoint(self, native_bar, raw_bar):
        self.ctx.check(load_library().fct_wfbp_gpu_adjoint(
            self.h, self._native(native_bar, "native_bar"), self._raw(raw_bar, "raw_bar")),
            "adjoint")

    def forward_native(self, raw, native, filt=None):
        """raw -> native volume (rebin + filter + backproject), FP32 plan."""
        if filt is None:
            filt, _ = self.alloc_filtered()
        self.rebin_filter(raw, filt)
        self.backproject(filt, None, native)
        return filt

    def count_voxel_views(self, per_voxel=None) -> int:
        tot = C.c_int64(0)
        pv = _dev_arg(per_voxel, "per_voxel", self.ctx.device, numel=self.info.native_elems,
                      dtype="i32", optional=True)
        self.ctx.check(load_library().fct_wfbp_gpu_count_voxel_views(self.h, pv,
                                                                    C.byref(tot)),
                       "count_voxel_views")
        return tot.value

    def voxel_selection(self, ix, iy, islice):
        import torch
        sel = torch.empty(self.scan.n_views, dtype=torch.int32, device=self._dev())
        self.ctx.check(load_library().fct_wfbp_gpu_voxel_selection(self.h, ix, iy, islice,
                                                                  sel.data_ptr()),
                       "voxel_selection")
        return sel

    def synchronize(self):
        self.ctx.check(load_library().fct_wfbp_gpu_synchronize(self.h), "synchronize")


class Batch:
    """Uneven reconstructions run together: per-job rebin_filter, one
    persistent backprojection over every job's tiles, per-job thicken,
    recorded into a hipGraph on first run (fct_wfbp_gpu_batch_*)."""

    GRAPH, EAGER = 0, 1

    def __init__(self, ctx: Context, jobs, precision: str = "fp32", tile_z: int = 0,
                 stream=None):
        """jobs: sequence of (ScanGeometry, ReconConfig)."""
        self.ctx = ctx
        self.precision = precision
        self.jobs = list(jobs)
        n = len(self.jobs)
        scans = (ScanGeometryC * n)(*[scan_struct(s) for s, _ in self.jobs])
        recons = (ReconConfigC * n)(*[recon_struct(r) for _, r in self.jobs])
        opts = ExecOptionsC(PRECISION[precision], tile_z, 0, 0)
        h = C.c_void_p()
        ctx.check(load_library().fct_wfbp_gpu_batch_create(ctx.h, n, scans, recons,
                                                           C.byref(opts),
                                                           _stream_handle(stream),
                                                           C.byref(h)), "batch_create")
        self.h = h
        self.plans = []
        for j, (s, r) in enumerate(self.jobs):
            ph = C.c_void_p()
            ctx.check(load_library().fct_wfbp_gpu_batch_job_plan(self.h, j, C.byref(ph)),
                      "batch_job_plan")
            self.plans.append(Plan._borrowed(ctx, ph, s, r, precision))

    def close(self):
        if getattr(self, "h", None):
            for p in self.plans:
                p.h = None
            load_library().fct_wfbp_gpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = BatchInfoC()
        self.ctx.check(load_library().fct_wfbp_gpu_batch_info(self.h, C.byref(i)), "batch_info")
        return {k: getattr(i, k) for k, _ in BatchInfoC._fields_}

    def submit(self, job: int, raw, volume):
        if not 0 <= job < len(self.plans):
            raise FctError(f"batch_submit: job {job} outside [0, {len(self.plans)})")
        p = self.plans[job]
        self.ctx.check(load_library().fct_wfbp_gpu_batch_submit(
            self.h, job, p._raw(raw), p._volume(volume)), "batch_submit")

    def run(self, mode: int = 0):
        self.ctx.check(load_library().fct_wfbp_gpu_batch_run(self.h, mode), "batch_run")

    def synchronize(self):
        self.ctx.check(load_library().fct_wfbp_gpu_batch_synchronize(self.h), "batch_sync")

    def backproject(self):
        """Only the persistent backprojection (eager), on the batch stream."""
        self.ctx.check(load_library().fct_wfbp_gpu_batch_backproject(self.h), "batch_bp")

    def autotune(self, candidates=(4, 8, 16), reps: int = 2) -> dict:
        n = len(candidates)
        tz = (C.c_int32 * n)(*candidates)
        ms = (C.c_float * n)()
        self.ctx.check(load_library().fct_wfbp_gpu_batch_autotune(self.h, tz, n, reps, ms),
                       "batch_autotune")
        return {int(t): float(m) for t, m in zip(candidates, ms)}
