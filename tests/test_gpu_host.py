"""The fct_wfbp command-line driver (C++ mirror of FreeCT's main.cpp +
ReconstructionMachine) on the GPU: a recon YAML in, a raw float32
[slice][y][x] volume out, equal to the Python/C-ABI path."""
import os
import subprocess

import numpy as np
import pytest

from freect_amd import scans as S
from helpers import gpu_ctx, say, to_np

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "freect_amd", "lib", "fct_wfbp")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return gpu_ctx()


def _python_volume(ctx, w):
    from freect_amd import wfbp
    raw = ctx.phantom_project(w.scan, w.phantom, w.i0, w.seed)
    p = wfbp.Plan(ctx, w.scan, w.recon)
    vol = p.alloc_volume()
    p.reconstruct(raw, vol)
    p.synchronize()
    return to_np(vol)


def test_cli_synthetic_matches_capi(ctx, tmp_path):
    w = S.small()
    cfg = tmp_path / "recon.yaml"
    rc = w.recon
    cfg.write_text(f"raw_data_dir: synthetic:small\noutput_dir: {tmp_path}\n"
                   f"start_pos: {rc.start_pos}\nend_pos: {rc.end_pos}\nrecon_fov: {rc.recon_fov}\n"
                   f"slice_thickness: {rc.slice_thickness}\nslice_pitch: {rc.slice_pitch}\n"
                   f"nx: {rc.nx}\nny: {rc.ny}\nrecon_kernel: 3\nx_origin: 0\ny_origin: 0\n"
                   f"filter: shepp-logan\n")
    out = tmp_path / "vol.dat"
    say("cli synthetic small")
    r = subprocess.run([CLI, "-v", "-o", str(out), str(cfg)], capture_output=True, text=True,
                       timeout=300)
    say(r.stdout[-800:])
    assert r.returncode == 0
    got = np.fromfile(out, np.float32).reshape(32, rc.ny, rc.nx)
    ref = _python_volume(ctx, w)
    assert np.array_equal(got, ref)


def test_cli_freect_format_round_trip(ctx, tmp_path):
    say("cli write-synthetic small")
    r = subprocess.run([CLI, "--write-synthetic", "small", str(tmp_path)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert (tmp_path / "meta.yaml").exists() and (tmp_path / "projections.dat").exists()
    out = tmp_path / "vol.dat"
    r = subprocess.run([CLI, "-o", str(out), str(tmp_path / "recon.yaml")], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    w = S.small()
    got = np.fromfile(out, np.float32).reshape(32, w.recon.ny, w.recon.nx)
    ref = _python_volume(ctx, w)
    # The two-file format stores per-view angle and z as float32, so the scan
    # law is re-derived from rounded values: agreement to a small fraction of 1 HU.
    d = np.abs(got.astype(np.float64) - ref) / (S.MU_WATER / 1000.0)
    say("freect format round trip: max %.4f HU, mean %.5f HU" % (d.max(), d.mean()))
    assert d.mean() < 0.05 and np.percentile(d, 99.9) < 1.0


def _write_freect(dirpath, sc, raw, angles, zs, ffs="FFSNONE"):
    meta = {"manufacturer": "SYNTHETIC", "detector_rows": sc.n_rows,
            "detector_channels": sc.n_channels,
            "detector_transverse_spacing": sc.channel_spacing,
            "detector_axial_spacing": sc.row_spacing, "detector_shape": "CYLINDRICAL",
            "distance_source_to_detector": sc.dsd, "distance_source_to_isocenter": sc.r,
            "detector_central_row": sc.central_row,
            "detector_central_channel": sc.central_channel, "projection_geometry": "FANBEAM",
            "scan_type": "HELICAL", "projections_per_rotation": sc.views_per_rotation,
            "flying_focal_spot_mode": ffs, "total_num_projections": raw.shape[0]}
    with open(os.path.join(dirpath, "meta.yaml"), "w") as f:
        for k, v in meta.items():
            f.write(f"{k}: {v!r}\n" if isinstance(v, float) else f"{k}: {v}\n")
    with open(os.path.join(dirpath, "projections.dat"), "wb") as f:
        for v in range(raw.shape[0]):
            np.array([angles[v], zs[v], sc.r], np.float32).tofile(f)
            raw[v].astype(np.float32).tofile(f)


@pytest.mark.parametrize("case", ["ccw", "cw", "ffs"])
def test_cli_freect_file_matches_oracle(ctx, oracle, tmp_path, case):
    # A FreeCT two-file scan (RawDataSet.cpp:85-107) written here from the
    # FP64 projector, read by the command-line driver (frame headers first,
    # the projections streamed from disk into the chunked GPU upload) and
    # reconstructed; the FP64 restatement reconstructs the same float32 data
    # on the geometry the driver derives from the headers.  Counter-clockwise,
    # clockwise, and phi+z flying focal spot with per-state deflections in the
    # frame headers.
    import json
    if case == "ffs":
        w = S.tiny(nx=48, n_slices=12, n_channels=96, n_rows=8, ppr=192, rotations=3,
                   ffs_mode=3)
    else:
        w = S.tiny(nx=48, n_slices=12, n_channels=192, n_rows=16, ppr=192, rotations=3)
    sc = w.scan.replace(rotation=-1, angle0=1.1) if case == "cw" else w.scan
    rc = w.recon
    raw = oracle.project_phantom(sc, w.phantom)
    v = np.arange(sc.n_views)
    angles = sc.angle0 + sc.rotation * v * 2 * np.pi / sc.views_per_rotation
    zs = sc.z0 + v * sc.feed / sc.views_per_rotation
    if case == "ffs":   # per focal state deflection offsets, mean zero
        angles = angles + np.array([3e-4, -3e-4, 3e-4, -3e-4])[v % 4]
        zs = zs + np.array([0.05, 0.05, -0.05, -0.05])[v % 4]
    _write_freect(str(tmp_path), sc, raw, angles, zs, "FFSXYZ" if case == "ffs" else "FFSNONE")
    cfg = tmp_path / "recon.yaml"
    cfg.write_text(f"raw_data_dir: {tmp_path}\noutput_dir: {tmp_path}\nstart_pos: {rc.start_pos!r}\n"
                   f"end_pos: {rc.end_pos!r}\nrecon_fov: {rc.recon_fov!r}\n"
                   f"slice_thickness: {rc.slice_thickness!r}\nslice_pitch: {rc.slice_pitch!r}\n"
                   f"nx: {rc.nx}\nny: {rc.ny}\nfilter: {rc.filter}\n"
                   f"anode_angle: {sc.anode_angle!r}\n")
    out = tmp_path / "vol.dat"
    say("cli freect file", case)
    r = subprocess.run([CLI, "-o", str(out), str(cfg)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout
    r = subprocess.run([CLI, "--inspect", str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout
    g = json.loads(r.stdout)
    assert g["rotation"] == sc.rotation and g["ffs_mode"] == sc.ffs_mode
    osc = sc.replace(channel_spacing=g["channel_spacing"], row_spacing=g["row_spacing"],
                     dsd=g["dsd"], r=g["r"], central_channel=g["central_channel"],
                     central_row=g["central_row"], angle0=g["angle0"], z0=g["z0"],
                     feed=g["feed"])
    # The recon keys pass through float32 (recon_structs.h), as in the driver.
    f32 = lambda x: float(np.float32(x))   # noqa: E731
    orc = rc.replace(start_pos=f32(rc.start_pos), end_pos=f32(rc.end_pos),
                     recon_fov=f32(rc.recon_fov), slice_thickness=f32(rc.slice_thickness),
                     slice_pitch=f32(rc.slice_pitch))
    d = oracle.geometry(osc, orc)
    got = np.fromfile(out, np.float32).reshape(d.n_slices_requested, rc.ny, rc.nx)
    nat, _ = oracle.reconstruct_native(osc, orc, raw, 0, d.n_slices_native)
    ref = oracle.thicken(osc, orc, nat)
    from helpers import fov_mask
    cov = oracle.covered_mask(osc, orc) & fov_mask(orc, d.acq_fov)[None]
    dd = (got.astype(np.float64) - ref)[cov] / (S.MU_WATER / 1000.0)
    rmse, mx = float(np.sqrt(np.mean(dd * dd))), float(np.abs(dd).max())
    say("cli freect %s vs fp64: covered %d, rmse %.4f max %.4f HU" % (case, cov.sum(), rmse, mx))
    assert cov.sum() > 0.2 * cov.size
    assert rmse <= 0.5 and mx <= 2.0


def _read_nhdr(path):
    fields, keys = {}, {}
    lines = open(path).read().splitlines()
    assert lines[0] == "NRRD0004"
    for ln in lines[1:]:
        if not ln or ln.startswith("#"):
            continue
        if ":=" in ln:
            k, v = ln.split(":=", 1)
            keys[k] = v
        else:
            k, v = ln.split(": ", 1)
            fields[k] = v
    return fields, keys


def test_cli_hounsfield_output_with_sidecar_header(ctx, tmp_path):
    # output_units: hu and the detached NRRD header written beside the raw
    # volume: the header's sizes, spacings and data file describe it, and the
    # data equal the C-ABI volume in HU.
    w = S.small()
    rc = w.recon
    cfg = tmp_path / "recon.yaml"
    cfg.write_text(f"raw_data_dir: synthetic:small\noutput_dir: {tmp_path}\n"
                   f"start_pos: {rc.start_pos}\nend_pos: {rc.end_pos}\nrecon_fov: {rc.recon_fov}\n"
                   f"slice_thickness: {rc.slice_thickness}\nslice_pitch: {rc.slice_pitch}\n"
                   f"nx: {rc.nx}\nny: {rc.ny}\nfilter: shepp-logan\noutput_units: hu\n"
                   f"mu_water: {S.MU_WATER}\n")
    out = tmp_path / "vol.dat"
    r = subprocess.run([CLI, "-o", str(out), str(cfg)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout
    fields, keys = _read_nhdr(str(out) + ".nhdr")
    nx, ny, nz = (int(v) for v in fields["sizes"].split())
    assert (nx, ny, nz) == (rc.nx, rc.ny, 32)
    sp = [float(v) for v in fields["spacings"].split()]
    assert np.allclose(sp, [rc.recon_fov / rc.nx, rc.recon_fov / rc.ny, rc.slice_pitch])
    assert fields["type"] == "float" and fields["endian"] == "little"
    assert fields["data file"] == "vol.dat" and keys["freect_units"] == "HU"
    got = np.fromfile(tmp_path / fields["data file"], np.float32).reshape(nz, ny, nx)
    from freect_amd import wfbp
    raw = ctx.phantom_project(w.scan, w.phantom, w.i0, w.seed)
    p = wfbp.Plan(ctx, w.scan, w.recon)
    p.set_output_units("hu", S.MU_WATER)
    vol = p.alloc_volume()
    p.reconstruct(raw, vol)
    p.synchronize()
    say("cli HU output: max |diff| %.3g HU" % np.abs(got - to_np(vol)).max())
    assert np.array_equal(got, to_np(vol))


def test_cli_dicom_ct_pd_matches_freect_format(ctx, oracle, tmp_path):
    # The same quantized scan as DICOM-CT-PD files (implicit VR, transposed
    # pixel data) and as a FreeCT two-file directory holding the values the
    # DICOM rescale produces: the driver reconstructs both to the same volume,
    # and that volume matches the FP64 restatement on the covered voxels.
    import dicom_writer as DW
    w = S.tiny(nx=48, n_slices=12, n_channels=192, n_rows=16, ppr=192, rotations=3)
    sc, rc = w.scan, w.recon
    raw = oracle.project_phantom(sc, w.phantom)
    v = np.arange(sc.n_views)
    angles = sc.angle0 + v * 2 * np.pi / sc.views_per_rotation
    zs = sc.z0 + v * sc.feed / sc.views_per_rotation
    ddir, fdir = tmp_path / "dicom", tmp_path / "freect"
    ddir.mkdir()
    fdir.mkdir()
    deq = DW.write_scan(str(ddir), sc, raw, angles, zs, syntax=DW.IMPLICIT, transposed=True)
    _write_freect(str(fdir), sc, deq, angles, zs)
    vols = {}
    for name, d in (("dicom", ddir), ("freect", fdir)):
        cfg = tmp_path / f"{name}.yaml"
        cfg.write_text(f"raw_data_dir: {d}\noutput_dir: {tmp_path}\nstart_pos: {rc.start_pos!r}\n"
                       f"end_pos: {rc.end_pos!r}\nrecon_fov: {rc.recon_fov!r}\n"
                       f"slice_thickness: {rc.slice_thickness!r}\n"
                       f"slice_pitch: {rc.slice_pitch!r}\nnx: {rc.nx}\nny: {rc.ny}\n"
                       f"filter: {rc.filter}\n")
        out = tmp_path / f"{name}.dat"
        say("cli", name)
        r = subprocess.run([CLI, "-o", str(out), str(cfg)], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout
        vols[name] = np.fromfile(out, np.float32)
    assert np.array_equal(vols["dicom"], vols["freect"])
    import json
    g = json.loads(subprocess.run([CLI, "--inspect", str(ddir)], capture_output=True, text=True,
                                  timeout=300).stdout)
    osc = sc.replace(channel_spacing=g["channel_spacing"], row_spacing=g["row_spacing"],
                     dsd=g["dsd"], r=g["r"], central_channel=g["central_channel"],
                     central_row=g["central_row"], angle0=g["angle0"], z0=g["z0"],
                     feed=g["feed"])
    f32 = lambda x: float(np.float32(x))   # noqa: E731
    orc = rc.replace(start_pos=f32(rc.start_pos), end_pos=f32(rc.end_pos),
                     recon_fov=f32(rc.recon_fov), slice_thickness=f32(rc.slice_thickness),
                     slice_pitch=f32(rc.slice_pitch))
    d = oracle.geometry(osc, orc)
    nat, _ = oracle.reconstruct_native(osc, orc, deq, 0, d.n_slices_native)
    ref = oracle.thicken(osc, orc, nat)
    from helpers import fov_mask
    cov = oracle.covered_mask(osc, orc) & fov_mask(orc, d.acq_fov)[None]
    got = vols["dicom"].reshape(ref.shape)
    dd = (got.astype(np.float64) - ref)[cov] / (S.MU_WATER / 1000.0)
    rmse, mx = float(np.sqrt(np.mean(dd * dd))), float(np.abs(dd).max())
    say("cli dicom vs fp64: rmse %.4f max %.4f HU" % (rmse, mx))
    assert rmse <= 0.5 and mx <= 2.0
