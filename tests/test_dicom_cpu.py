"""The DICOM-CT-PD reader (fct::DicomDataSet, a self-written DICOM parser in
place of DCMTK; DicomDataSet.cpp:56-330) through the command-line driver's
--inspect, no GPU: explicit and implicit VR little endian, pixel data stored
[row][channel] or transposed, an undefined-length sequence to skip, the
rescale to float, the scan metadata of the first file, and the errors for
what it does not read."""
import json
import os
import subprocess

import numpy as np
import pytest

from freect_amd import scans as S
import dicom_writer as DW

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "freect_amd", "lib", "fct_wfbp")


def _inspect(path):
    r = subprocess.run([CLI, "--inspect", str(path)], capture_output=True, text=True, timeout=120)
    return r.returncode, (json.loads(r.stdout) if r.returncode == 0 else r.stdout)


def _scan(oracle):
    w = S.tiny(n_channels=40, n_rows=6, ppr=32, rotations=2.0)
    sc = w.scan
    raw = oracle.project_phantom(sc, w.phantom)
    v = np.arange(sc.n_views)
    return sc, raw, sc.angle0 + v * 2 * np.pi / sc.views_per_rotation, \
        sc.z0 + v * sc.feed / sc.views_per_rotation


@pytest.mark.parametrize("syntax,transposed,seq", [(DW.EXPLICIT, False, False),
                                                   (DW.IMPLICIT, True, True),
                                                   (DW.EXPLICIT, True, True)])
def test_dicom_reader(tmp_path, oracle, syntax, transposed, seq):
    sc, raw, angles, zs = _scan(oracle)
    deq = DW.write_scan(str(tmp_path), sc, raw, angles, zs, syntax=syntax, transposed=transposed,
                        with_sequence=seq)
    rc, g = _inspect(tmp_path)
    assert rc == 0, g
    assert (g["n_views"], g["n_rows"], g["n_channels"], g["ppr"]) == \
        (sc.n_views, sc.n_rows, sc.n_channels, sc.views_per_rotation)
    assert g["manufacturer"] == "SIEMENS" and g["detector_shape"] == "CYLINDRICAL"
    for k, ref in (("dsd", sc.dsd), ("r", sc.r), ("channel_spacing", sc.channel_spacing),
                   ("row_spacing", sc.row_spacing), ("central_channel", sc.central_channel),
                   ("central_row", sc.central_row)):
        assert g[k] == float(np.float32(ref)), k
    assert abs(g["angle0"] - np.float32(sc.angle0)) < 1e-6 and g["rotation"] == 1
    assert abs(g["feed"] - sc.feed) < 1e-4 * abs(sc.feed)
    # The rescaled values, summed in FP64 by the reader, equal the float32
    # slope * pixel + intercept of every sample.
    assert g["sum"] == pytest.approx(deq.astype(np.float64).sum(), rel=1e-12, abs=1e-9)


def test_dicom_reader_errors(tmp_path, oracle):
    sc, raw, angles, zs = _scan(oracle)
    d1 = tmp_path / "be"
    d1.mkdir()
    DW.write_scan(str(d1), sc, raw[:2], angles, zs, syntax=DW.BIG_ENDIAN)
    rc, out = _inspect(d1)
    assert rc != 0 and "transfer syntax" in out
    d2 = tmp_path / "trunc"
    d2.mkdir()
    DW.write_scan(str(d2), sc, raw, angles, zs)
    name = sorted(os.listdir(d2))[5]
    data = open(d2 / name, "rb").read()
    open(d2 / name, "wb").write(data[: len(data) - 100])
    rc, out = _inspect(d2)
    assert rc != 0 and ("past the file end" in out or "pixel data" in out)
    d3 = tmp_path / "empty"
    d3.mkdir()
    rc, out = _inspect(d3)
    assert rc != 0 and "no files" in out
