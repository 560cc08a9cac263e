"""Host C++ mirror of the reference boundary (no GPU): recon YAML parsing
(parse_config.cpp:31-62), the FreeCT two-file format (RawDataSet.cpp:17-107)
with the scan law derived as in ReconstructionMachine.cu:48-64, and the
synthetic phantom recipe shared with freect_amd/scans.py."""
import json
import os
import subprocess

import numpy as np
import pytest

from freect_amd import scans as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "freect_amd", "lib", "fct_wfbp")


def run(*args):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout


def test_phantom_recipe_matches_python():
    for name, seed, zr in [("small", 1000, (-20, 20)), ("full", 1001, (-100, 100))]:
        rc, out = run("--print-phantom", name)
        assert rc == 0
        cpp = json.loads(out)
        assert cpp["seed"] == seed
        py = S.make_phantom(seed, *zr)
        assert len(cpp["ellipsoids"]) == len(py.ellipsoids) == 20
        for a, b in zip(cpp["ellipsoids"], py.ellipsoids):
            assert np.allclose(a["c"], b["c"], rtol=0, atol=1e-12)
            assert np.allclose(a["a"], b["a"], rtol=0, atol=1e-12)
            assert np.allclose(a["rot"], b["rot"], rtol=0, atol=1e-14)
            assert abs(a["dmu"] - b["dmu"]) < 1e-16


def test_parse_config_reference_keys(tmp_path):
    # The keys RawDataSet::writeReconFile emits (RawDataSet.cpp:56-78).
    cfg = tmp_path / "recon.yaml"
    cfg.write_text("raw_data_dir: /data/scan   # comment\noutput_dir: ./\nstart_pos: -12.5\n"
                   "end_pos: 30\nrecon_fov: 500.0\nslice_thickness: 1.0\nnx: 512\nny: 384\n"
                   "recon_kernel: 3\nx_origin: 1.5\ny_origin: -2\ntube_angle_offset: 0.0\n"
                   "adaptive_filtration_s: 0.2\n")
    rc, out = run("--print-config", str(cfg))
    assert rc == 0
    p = json.loads(out)
    assert p["raw_data_dir"] == "/data/scan" and p["output_dir"] == "./"
    assert p["start_pos"] == -12.5 and p["end_pos"] == 30 and p["recon_fov"] == 500
    assert p["nx"] == 512 and p["ny"] == 384 and p["recon_kernel"] == 3
    assert p["x_origin"] == 1.5 and p["y_origin"] == -2
    assert p["slice_pitch"] == 1.0            # 0 -> slice_thickness (parse_config.cpp:59-60)
    assert p["output_file"] == str(cfg) + "_recon.dat"   # parse_config.cpp:43
    assert p["filter"] == "ramp" and p["precision"] == "fp32"


def test_parse_config_missing_keys_default_to_zero(tmp_path):
    cfg = tmp_path / "partial.yaml"
    cfg.write_text("start_pos: 5\nslice_thickness: 2.5\nslice_pitch: 1.25\nfilter: hann\n")
    rc, out = run("--print-config", str(cfg))
    assert rc == 0
    p = json.loads(out)
    assert p["end_pos"] == 0 and p["nx"] == 0 and p["recon_fov"] == 0
    assert p["slice_pitch"] == 1.25 and p["filter"] == "hann"
    rc, out = run("--print-config", str(tmp_path / "missing.yaml"))
    assert rc != 0 and "ERROR" in out


def _write_freect(dirpath, scan, raw, angles, zs, ppr):
    meta = {"manufacturer": "GE", "detector_rows": scan.n_rows,
            "detector_channels": scan.n_channels,
            "detector_transverse_spacing": scan.channel_spacing,
            "detector_axial_spacing": scan.row_spacing, "detector_shape": "CYLINDRICAL",
            "distance_source_to_detector": scan.dsd,
            "distance_source_to_isocenter": scan.r,
            "detector_central_row": scan.central_row,
            "detector_central_channel": scan.central_channel,
            "projection_geometry": "FANBEAM", "scan_type": "HELICAL",
            "projections_per_rotation": ppr, "flying_focal_spot_mode": "FFSNONE",
            "total_num_projections": raw.shape[0]}
    with open(os.path.join(dirpath, "meta.yaml"), "w") as f:
        f.write("# FREECT READER OUTPUT\n")
        for k, v in meta.items():
            f.write(f"{k}: {v!r}\n" if isinstance(v, float) else f"{k}: {v}\n")
    with open(os.path.join(dirpath, "projections.dat"), "wb") as f:
        for v in range(raw.shape[0]):
            np.array([angles[v], zs[v], scan.r], np.float32).tofile(f)
            raw[v].astype(np.float32).tofile(f)


def test_freect_format_reader_and_scan_law(tmp_path, oracle):
    w = S.tiny(n_channels=40, n_rows=4, ppr=32, rotations=2.5)
    sc = w.scan
    raw = oracle.project_phantom(sc, w.phantom)
    v = np.arange(sc.n_views)
    angles = sc.angle0 + v * 2 * np.pi / sc.views_per_rotation
    zs = sc.z0 + v * sc.feed / sc.views_per_rotation
    _write_freect(str(tmp_path), sc, raw, angles, zs, sc.views_per_rotation)
    rc, out = run("--inspect", str(tmp_path))
    assert rc == 0, out
    g = json.loads(out)
    assert g["n_views"] == sc.n_views and g["n_rows"] == 4 and g["n_channels"] == 40
    assert g["ppr"] == 32
    assert abs(g["angle0"] - np.float32(sc.angle0)) < 1e-6
    assert abs(g["z0"] - np.float32(sc.z0)) < 1e-5
    # feed = z(ppr) - z(0) from the float32 frame headers.
    assert abs(g["feed"] - sc.feed) < 1e-4 * abs(sc.feed)
    assert abs(g["sum"] - raw.astype(np.float64).sum()) < 1e-6 * abs(raw).sum()
    assert abs(g["z_last"] - np.float32(zs[-1])) < 1e-5


def test_freect_format_truncated_file_is_an_error(tmp_path, oracle):
    w = S.tiny(n_channels=16, n_rows=2, ppr=16, rotations=1.0)
    sc = w.scan
    raw = np.zeros((sc.n_views, 2, 16), np.float32)
    v = np.arange(sc.n_views)
    _write_freect(str(tmp_path), sc, raw, v * 0.1, v * 0.01, 16)
    data = open(tmp_path / "projections.dat", "rb").read()
    open(tmp_path / "projections.dat", "wb").write(data[: len(data) // 2])
    rc, out = run("--inspect", str(tmp_path))
    assert rc != 0 and "shorter" in out


def test_parse_config_engine_keys(tmp_path):
    cfg = tmp_path / "engine.yaml"
    cfg.write_text("start_pos: 0\nend_pos: 10\nadaptive_filtration_s: 0.25\n"
                   "adaptive_filtration: 1\noutput_units: hu\nmu_water: 0.0195\n")
    rc, out = run("--print-config", str(cfg))
    assert rc == 0
    p = json.loads(out)
    assert p["adaptive_filtration"] == 1 and abs(p["adaptive_filtration_s"] - 0.25) < 1e-7
    assert p["output_units"] == "hu" and abs(p["mu_water"] - 0.0195) < 1e-12
    # A reference recon file (adaptive_filtration_s: 0.2, RawDataSet.cpp:77) alone
    # leaves the filtration off and the output in 1/mm, as the reference does.
    cfg.write_text("start_pos: 0\nend_pos: 10\nadaptive_filtration_s: 0.2\n")
    p = json.loads(run("--print-config", str(cfg))[1])
    assert p["adaptive_filtration"] == 0 and p["output_units"] == "mu"
