"""Minimal DICOM-CT-PD writer for the reader's tests (no pydicom here): one
projection per file, the standard and private elements FreeCT's
DicomDataSet reads (DicomDataSet.cpp:56-148, 232-302; private tags of
:178-211), in explicit or implicit VR little endian, with the pixel data
stored [row][channel] or transposed, and optionally an undefined-length
sequence the reader has to skip."""
import struct

import numpy as np

EXPLICIT = "1.2.840.10008.1.2.1"
IMPLICIT = "1.2.840.10008.1.2"
BIG_ENDIAN = "1.2.840.10008.1.2.2"


def _pad(b: bytes, fill=b" ") -> bytes:
    return b + fill if len(b) % 2 else b


def _elem(group, elem, vr, value: bytes, explicit=True) -> bytes:
    head = struct.pack("<HH", group, elem)
    if not explicit:
        return head + struct.pack("<I", len(value)) + value
    if vr in ("OB", "OW", "OF", "SQ", "UN", "UT"):
        return head + vr.encode() + b"\0\0" + struct.pack("<I", len(value)) + value
    return head + vr.encode() + struct.pack("<H", len(value)) + value


def _seq_undefined(group, elem, items, explicit=True) -> bytes:
    """A sequence of undefined length, each item of undefined length."""
    head = struct.pack("<HH", group, elem)
    out = head + (b"SQ\0\0" if explicit else b"") + struct.pack("<I", 0xFFFFFFFF)
    for it in items:
        out += struct.pack("<HHI", 0xFFFE, 0xE000, 0xFFFFFFFF) + it
        out += struct.pack("<HHI", 0xFFFE, 0xE00D, 0)
    return out + struct.pack("<HHI", 0xFFFE, 0xE0DD, 0)


def write_projection(path, scan_meta, frame, pixels, slope, intercept,
                     syntax=EXPLICIT, transposed=False, with_sequence=False):
    """scan_meta: manufacturer, rows, channels, transverse/axial spacing,
    shape, r, dsd, ffs_mode, ppr, scan_type, geometry, central (channel,
    row).  frame: angle, z, radial, ffs shifts (angular, axial, radial).
    pixels: uint16 [rows][channels]."""
    ex = syntax != IMPLICIT
    f32 = lambda *v: struct.pack("<%df" % len(v), *v)   # noqa: E731
    u16 = lambda v: struct.pack("<H", v)                 # noqa: E731
    s = lambda t: _pad(t.encode())                       # noqa: E731
    img = pixels.T if transposed else pixels
    els = [
        (0x0008, 0x0070, "LO", s(scan_meta["manufacturer"])),
        (0x0028, 0x0010, "US", u16(img.shape[0])),
        (0x0028, 0x0011, "US", u16(img.shape[1])),
        (0x0028, 0x0100, "US", u16(16)),
        (0x0028, 0x0103, "US", u16(0)),
        (0x0028, 0x1052, "DS", s(repr(float(intercept)))),
        (0x0028, 0x1053, "DS", s(repr(float(slope)))),
        (0x7029, 0x0010, "LO", s("DetectorSystemArrangementModule")),
        (0x7029, 0x1002, "FL", f32(scan_meta["transverse_spacing"])),
        (0x7029, 0x1006, "FL", f32(scan_meta["axial_spacing"])),
        (0x7029, 0x100B, "CS", s(scan_meta["shape"])),
        (0x7029, 0x1010, "US", u16(scan_meta["rows"])),
        (0x7029, 0x1011, "US", u16(scan_meta["channels"])),
        (0x7031, 0x0010, "LO", s("DetectorDynamicsModule")),
        (0x7031, 0x1001, "FL", f32(frame["angle"])),
        (0x7031, 0x1002, "FL", f32(frame["z"])),
        (0x7031, 0x1003, "FL", f32(frame["radial"])),
        (0x7031, 0x1031, "FL", f32(scan_meta["dsd"])),
        (0x7031, 0x1033, "FL", f32(*scan_meta["central"])),
        (0x7033, 0x0010, "LO", s("SourceDynamicsModule")),
        (0x7033, 0x100B, "FL", f32(frame["ffs"][0])),
        (0x7033, 0x100C, "FL", f32(frame["ffs"][1])),
        (0x7033, 0x100D, "FL", f32(frame["ffs"][2])),
        (0x7033, 0x100E, "CS", s(scan_meta["ffs_mode"])),
        (0x7033, 0x1013, "US", u16(scan_meta["ppr"])),
        (0x7037, 0x0010, "LO", s("ProjectionDataDefinitions")),
        (0x7037, 0x1009, "CS", s(scan_meta["scan_type"])),
        (0x7037, 0x100A, "CS", s(scan_meta["geometry"])),
    ]
    body = b""
    if with_sequence:   # Referenced Image Sequence with a nested element
        item = _elem(0x0008, 0x1150, "UI", _pad(b"1.2.3.4", b"\0"), ex)
        body += _seq_undefined(0x0008, 0x1140, [item], ex)
    for g, e, vr, v in els:
        body += _elem(g, e, vr, v, ex)
    body += _elem(0x7FE0, 0x0010, "OW", np.ascontiguousarray(img, "<u2").tobytes(), ex)
    meta = (_elem(0x0002, 0x0001, "OB", b"\0\1") +
            _elem(0x0002, 0x0002, "UI", _pad(b"1.2.840.10008.5.1.4.1.1.2", b"\0")) +
            _elem(0x0002, 0x0003, "UI", _pad(b"1.2.3.4.5", b"\0")) +
            _elem(0x0002, 0x0010, "UI", _pad(syntax.encode(), b"\0")))
    meta = _elem(0x0002, 0x0000, "UL", struct.pack("<I", len(meta))) + meta
    with open(path, "wb") as f:
        f.write(b"\0" * 128 + b"DICM" + meta + body)


def quantize(raw: np.ndarray):
    """uint16 pixels and (slope, intercept) of a float32 scan; the values the
    reader reconstructs are float32(slope) * pixel + float32(intercept)."""
    lo, hi = float(raw.min()), float(raw.max())
    intercept = lo - 0.01
    slope = (hi - intercept) / 65000.0
    px = np.clip(np.rint((raw - intercept) / slope), 0, 65535).astype(np.uint16)
    deq = np.float32(slope) * px.astype(np.float32) + np.float32(intercept)
    return px, slope, intercept, deq.astype(np.float32)


def write_scan(dirpath, sc, raw, angles, zs, ffs_mode="FFSNONE", manufacturer="SIEMENS",
               syntax=EXPLICIT, transposed=False, with_sequence=False, ffs_shifts=None):
    """One file per projection, named so that name order is view order.
    Returns the dequantized float32 scan the reader should produce."""
    import os
    px, slope, intercept, deq = quantize(raw)
    meta = {"manufacturer": manufacturer, "rows": sc.n_rows, "channels": sc.n_channels,
            "transverse_spacing": sc.channel_spacing, "axial_spacing": sc.row_spacing,
            "shape": "CYLINDRICAL", "r": sc.r, "dsd": sc.dsd, "ffs_mode": ffs_mode,
            "ppr": sc.views_per_rotation, "scan_type": "HELICAL", "geometry": "FANBEAM",
            "central": (sc.central_channel, sc.central_row)}
    for v in range(raw.shape[0]):
        shift = ffs_shifts[v] if ffs_shifts is not None else (0.0, 0.0, 0.0)
        frame = {"angle": angles[v], "z": zs[v], "radial": sc.r, "ffs": shift}
        write_projection(os.path.join(dirpath, "proj_%06d.dcm" % v), meta, frame, px[v],
                         slope, intercept, syntax, transposed, with_sequence)
    return deq
