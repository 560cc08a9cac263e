"""CPU checks of the scan geometry conventions (no GPU):

* every address the rebinning gather forms (rebin_filter.hip: rebin_sample,
  with the table as the plan uploads it) lies inside the raw scan, for every
  flying-focal-spot mode and view counts that are not multiples of the focal
  states, and the register kernel's windowed 32-bit offsets address the same
  taps;
* a clockwise scan: the FP64 projector against chords computed here from the
  definition (source at angle0 - v dtheta, fan angle measured with the
  rotation), and the host's per-view law check on clockwise and jittered
  FreeCT-format scans;
* whole-chain orientation: an off-centre sphere reconstructs at its analytic
  (x, y, z), counter-clockwise and clockwise.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from freect_amd import scans as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "freect_amd", "lib", "fct_wfbp")


@pytest.mark.parametrize("ffs_mode", [0, 1, 2, 3])
@pytest.mark.parametrize("extra_views", [0, 1, 2, 3])
def test_rebin_gather_addresses_inside_scan(oracle, ffs_mode, extra_views):
    # The gather of rebin_filter.hip reads, for parallel view j, effective
    # row q and table entry e (state = valid ? rho + 1 : 0):
    #   rho' = max(state - 1, 0), m = clamp(j + aoff (+1), max: mmax, min: 0),
    #   raw[((m S + rho') nrows + (q >> zshift)) ncols + b0 / b1]
    # for EVERY entry, selected or not (the value of an entry without a fan
    # sample is discarded afterwards).  All of those must lie inside the scan.
    ppr = 64
    w = S.tiny(n_channels=72, n_rows=6, ppr=ppr, rotations=2.0, ffs_mode=ffs_mode)
    sc = w.scan.replace(n_views=w.scan.n_views + extra_views)
    Sn = sc.ffs_states
    e = sc.effective()
    ient, fent = oracle.rebin_entries(sc)
    zpar = 2 if ffs_mode & 2 else 1
    assert ient.shape[0] == zpar * e.n_channels
    valid, rho, b0, b1, aoff, mmax = (ient[:, i].astype(np.int64) for i in range(6))
    state = np.where(valid != 0, rho + 1, 0)
    rho_r = np.maximum(state - 1, 0)
    assert np.all(mmax >= 0)
    assert np.all((b0 >= 0) & (b0 < sc.n_channels) & (b1 >= 0) & (b1 < sc.n_channels))
    assert np.all((fent[valid != 0] >= 0.0) & (fent[valid != 0] < 1.0))
    j = np.arange(e.n_views, dtype=np.int64)[:, None]
    n_raw = sc.n_views * sc.n_rows * sc.n_channels
    for off in (0, 1):
        m = np.maximum(np.minimum(j + aoff[None, :] + off, mmax[None, :]), 0)
        proj = m * Sn + rho_r[None, :]
        assert proj.min() >= 0 and proj.max() < sc.n_views
        for q in range(e.n_rows):
            row = q >> (1 if zpar == 2 else 0)
            assert 0 <= row < sc.n_rows
            for b in (b0, b1):
                addr = (proj * sc.n_rows + row) * sc.n_channels + b[None, :]
                assert addr.min() >= 0 and addr.max() < n_raw
    # Selected entries read the focal state they name.
    sel = valid != 0
    assert np.all(rho_r[sel] == rho[sel]) and np.all(rho[sel] < Sn)

    # The register kernel's windowed form of the same taps (rebin_filter.hip:
    # raw_window, rebin_taps_w): byte offsets from the first raw view group a
    # line pair (views j0 <= j1 <= j0 + 1) can reach, m_lo = max(0, min(j0 +
    # aoff_min, mmax_min)), with the view clamp taken relative to m_lo.  Every
    # relative view index is >= 0 (24-bit products), every offset fits the
    # 32-bit window and addresses exactly the tap of the plain formula.
    group_bytes = Sn * sc.n_rows * sc.n_channels * 4
    assert group_bytes < 1 << 24
    boff = np.where(valid != 0, (rho * sc.n_rows * sc.n_channels + b0) * 4, 0)
    aoff_min, mmax_min = min(int(aoff.min()), 0), int(mmax.min())
    j0 = np.arange(e.n_views, dtype=np.int64)[:, None]
    m_lo = np.maximum(0, np.minimum(j0 + aoff_min, mmax_min))
    for dj in (0, 1):
        j1 = np.minimum(j0 + dj, e.n_views - 1)
        for off in (0, 1):
            m_abs = np.maximum(np.minimum(j1 + aoff[None, :] + off, mmax[None, :]), 0)
            rel = np.maximum(np.minimum(j1 - m_lo + aoff[None, :] + off, mmax[None, :] - m_lo),
                             -m_lo)
            assert np.array_equal(rel, m_abs - m_lo)
            assert rel.min() >= 0 and rel.max() < 1 << 24
            for q in range(e.n_rows):
                row = q >> (1 if zpar == 2 else 0)
                o0 = rel * group_bytes + boff[None, :] + row * sc.n_channels * 4
                o1 = o0 + (b1 - b0)[None, :] * 4
                assert o0.min() >= 0 and max(o0.max(), o1.max()) + 4 <= 0xFFFFFFF0
                for o, b in ((o0, b0), (o1, b1)):
                    plain = ((m_abs * Sn + rho_r[None, :]) * sc.n_rows + row) * sc.n_channels
                    assert np.array_equal(m_lo * group_bytes + o, (plain + b[None, :]) * 4)


def _sphere_chords_cw(sc, centre, radius, dmu):
    """Chord lengths x dmu of a sphere for a clockwise scan, from the
    definition: source at a_v = angle0 - v dtheta, z_v = z0 + v feed / ppr;
    the fan angle beta = (c - cc) dgamma runs with the rotation, so the ray
    leaves the source along -(cos(a - beta), sin(a - beta)) in the plane and
    reaches row height (r - cr) drow at in-plane distance Dsd."""
    v = np.arange(sc.n_views)[:, None, None]
    row = np.arange(sc.n_rows)[None, :, None]
    c = np.arange(sc.n_channels)[None, None, :]
    a = sc.angle0 - v * 2 * math.pi / sc.views_per_rotation
    beta = (c - sc.central_channel) * sc.channel_spacing / sc.dsd
    src = np.stack(np.broadcast_arrays(sc.r * np.cos(a), sc.r * np.sin(a),
                                       sc.z0 + v * sc.feed / sc.views_per_rotation), -1)
    d = np.stack(np.broadcast_arrays(-sc.dsd * np.cos(a - beta), -sc.dsd * np.sin(a - beta),
                                     (row - sc.central_row) * sc.row_spacing + 0.0 * c), -1)
    u = d / np.linalg.norm(d, axis=-1, keepdims=True)
    p = src - np.asarray(centre)
    t = (p * u).sum(-1)
    d2 = (p * p).sum(-1) - t * t
    return np.where(d2 < radius ** 2, 2 * np.sqrt(np.maximum(radius ** 2 - d2, 0.0)), 0.0) * dmu


def _sphere(centre, radius, hu):
    return {"c": tuple(centre), "a": (radius, radius, radius),
            "rot": np.eye(3).reshape(9).tolist(), "dmu": S.MU_WATER * hu / 1000.0}


def test_clockwise_projector_matches_definition(oracle):
    w = S.tiny(n_channels=96, n_rows=8, ppr=96, rotations=1.5)
    sc = w.scan.replace(rotation=-1, angle0=0.7)
    centre, radius = (40.0, -55.0, 3.0), 30.0
    ph = S.Phantom(cyl_radius=150.0, cyl_mu=0.0, ellipsoids=[_sphere(centre, radius, 1000.0)])
    got = oracle.project_phantom(sc, ph)
    ref = _sphere_chords_cw(sc, centre, radius, ph.ellipsoids[0]["dmu"])
    assert np.abs(ref).max() > 0.5
    assert np.max(np.abs(got - ref)) < 1e-5
    # The same scan read counter-clockwise differs (the convention matters).
    ccw = oracle.project_phantom(sc.replace(rotation=1), ph)
    assert np.max(np.abs(ccw - ref)) > 0.1


@pytest.mark.parametrize("rotation", [1, -1])
def test_offcentre_sphere_reconstructs_at_its_position(oracle, rotation):
    # Whole chain (rebin, filter, backprojection) through the FP64
    # restatement: the centroid of the reconstructed sphere is within a
    # fraction of a voxel of its analytic centre in x, y and z.
    w = S.tiny(n_channels=192, n_rows=16, ppr=192, rotations=3.0, nx=96, n_slices=24)
    sc = w.scan.replace(rotation=rotation, angle0=0.3)
    d0 = oracle.geometry(sc, w.recon)
    rc = w.recon.replace(recon_fov=240.0, x_origin=0.0, y_origin=0.0)
    zc = 0.5 * (rc.start_pos + rc.end_pos) + 0.37 * d0.csw
    centre, radius = (62.0, -41.0, zc), 18.0
    ph = S.Phantom(cyl_radius=150.0, cyl_mu=0.0, ellipsoids=[_sphere(centre, radius, 1000.0)])
    raw = oracle.project_phantom(sc, ph)
    d = oracle.geometry(sc, rc)
    vol, _ = oracle.reconstruct_native(sc, rc, raw, 0, d.n_slices_native)
    xs = (np.arange(rc.nx) - 0.5 * (rc.nx - 1)) * rc.recon_fov / rc.nx + rc.x_origin
    ys = (np.arange(rc.ny) - 0.5 * (rc.ny - 1)) * rc.recon_fov / rc.ny + rc.y_origin
    zs = oracle.slice_grid(rc.start_pos, rc.end_pos, d.csw)
    m = np.where(vol > 0.5 * ph.ellipsoids[0]["dmu"], vol, 0.0)
    tot = m.sum()
    cz = (m.sum(axis=(1, 2)) * zs).sum() / tot
    cy = (m.sum(axis=(0, 2)) * ys).sum() / tot
    cx = (m.sum(axis=(0, 1)) * xs).sum() / tot
    vox = rc.recon_fov / rc.nx
    print("sphere centroid", cx, cy, cz, "truth", centre, "voxel", vox, "csw", d.csw)
    assert abs(cx - centre[0]) < 0.25 * vox and abs(cy - centre[1]) < 0.25 * vox
    assert abs(cz - centre[2]) < 0.25 * d.csw
    # The mirrored and swapped positions are far off.
    assert math.hypot(cx - centre[0], cy + centre[1]) > 20 * vox / 4


def _write_freect(dirpath, sc, raw, angles, zs, ffs="FFSNONE"):
    meta = {"manufacturer": "GE", "detector_rows": sc.n_rows, "detector_channels": sc.n_channels,
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


def _inspect(path):
    import json
    r = subprocess.run([CLI, "--inspect", str(path)], capture_output=True, text=True,
                       timeout=120)
    return r.returncode, (json.loads(r.stdout) if r.returncode == 0 else r.stdout)


def _law(sc):
    v = np.arange(sc.n_views)
    return (sc.angle0 + sc.rotation * v * 2 * np.pi / sc.views_per_rotation,
            sc.z0 + v * sc.feed / sc.views_per_rotation)


def test_clockwise_freect_scan_is_read_as_clockwise(tmp_path, oracle):
    w = S.tiny(n_channels=40, n_rows=4, ppr=32, rotations=2.5)
    sc = w.scan.replace(rotation=-1, angle0=2.0)
    raw = oracle.project_phantom(sc, w.phantom)
    angles, zs = _law(sc)
    # Angles stored wrapped to [0, 2 pi), as DICOM-CT-PD frames hold them.
    _write_freect(str(tmp_path), sc, raw, np.mod(angles, 2 * np.pi), zs)
    rc, g = _inspect(tmp_path)
    assert rc == 0, g
    assert g["rotation"] == -1
    assert abs(g["angle0"] - np.float32(2.0)) < 1e-6
    assert abs(g["feed"] - sc.feed) < 1e-4 * abs(sc.feed)


@pytest.mark.parametrize("what", ["z", "angle", "step"])
def test_off_law_freect_scan_is_rejected(tmp_path, oracle, what):
    w = S.tiny(n_channels=40, n_rows=4, ppr=32, rotations=2.5)
    sc = w.scan
    raw = np.zeros((sc.n_views, sc.n_rows, sc.n_channels), np.float32)
    angles, zs = _law(sc)
    csw = sc.csw
    if what == "z":            # table jitter of 10 % of a slice at one view
        zs = zs.copy()
        zs[37] += 0.1 * csw
    elif what == "angle":      # one view 10 % of a step late
        angles = angles.copy()
        angles[50] += 0.1 * 2 * np.pi / sc.views_per_rotation
    else:                      # views 1.5 steps apart: not 2 pi / ppr
        angles = sc.angle0 + 1.5 * np.arange(sc.n_views) * 2 * np.pi / sc.views_per_rotation
    _write_freect(str(tmp_path), sc, raw, angles, zs)
    rc, out = _inspect(tmp_path)
    assert rc != 0
    assert ("off the uniform helical law" in out) or ("is not +-2pi" in out), out
    # Jitter inside the stated tolerance (1 % of a step / of a slice) passes.
    if what != "step":
        angles, zs = _law(sc)
        if what == "z":
            zs = zs + 0.01 * csw * np.sin(np.arange(sc.n_views))
        else:
            angles = angles + 0.01 * 2 * np.pi / sc.views_per_rotation * np.cos(
                np.arange(sc.n_views))
        _write_freect(str(tmp_path), sc, raw, angles, zs)
        rc, g = _inspect(tmp_path)
        assert rc == 0, g


def test_anode_angle_recon_key(tmp_path):
    import json
    cfg = tmp_path / "r.yaml"
    cfg.write_text("start_pos: 0\nend_pos: 1\nanode_angle: 0.15\n")
    r = subprocess.run([CLI, "--print-config", str(cfg)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0
    assert abs(json.loads(r.stdout)["anode_angle"] - 0.15) < 1e-7


def test_oracle_slices_past_the_scan_are_zero(oracle):
    # Native slices no view reaches: the restatement returns zeros (an empty
    # view range), forward and transposed.
    w = S.tiny()
    rc = w.recon.replace(start_pos=-400.0, end_pos=-380.0)
    raw = oracle.project_phantom(w.scan, w.phantom)
    a, b = oracle.view_range(w.scan, rc, 0, 1)
    assert a == b
    vol, n = oracle.reconstruct_native(w.scan, rc, raw, 0, 2)
    assert n == 0 and not vol.any()
    pbar = oracle.adjoint_native(w.scan, rc, np.ones((2, rc.ny, rc.nx)), 0)
    assert not pbar.any()
