"""Flying focal spot (FFS) in the FP64 restatement (oracle/), CPU only.

The interleave has no reference counterpart (fct_wfbp ignores FFS); what
pins it here:
  * the interleaved geometry: Python ScanGeometry.effective() == the C++
    or_effective_scan (two independent statements of the same rules);
  * the projector: chords of a centred sphere along the deflected rays,
    computed with vector geometry from the focal-spot positions of
    generate_system_matrix_ffs.cpp:110-149, 250-283;
  * the rebinning: an off-centre cylinder along z (an ellipsoid 1e5 mm long)
    has the closed-form parallel projection 2 dmu sqrt(a^2 - (p - p_c)^2),
    p_c = cx sin(theta) - cy cos(theta), at every interleaved ray
    (theta_j, p_k); a half-channel mistake in the phi interleave or a
    one-projection mistake in the state's view offset fails it;
  * the whole chain: a uniform water cylinder reconstructs to mu_water.
"""
import math

import numpy as np
import pytest

from freect_amd import scans as S


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_effective_geometry_python_equals_cpp(oracle, mode):
    sc = S.tiny(n_channels=120, n_rows=6, ppr=96, ffs_mode=mode).scan
    e = sc.effective()
    c = oracle.effective_scan(sc)
    for k in ("n_channels", "n_rows", "n_views", "views_per_rotation", "channel_spacing",
              "row_spacing", "central_channel", "central_row", "dsd", "r", "angle0", "z0",
              "feed", "ffs_mode"):
        assert getattr(e, k) == c[k], k
    if mode == 0:
        assert e == sc


def _ffs_offsets(sc):
    """Focal-spot offsets (e_w, e_u, e_z) per state, from the amplitudes of
    generate_system_matrix_ffs.cpp:66-67 and the order of :250-283."""
    csw = sc.r * sc.row_spacing / sc.dsd
    dr = sc.dsd * csw / (4 * (sc.dsd - sc.r) * math.tan(sc.anode_angle))
    da = sc.dsd * sc.r * math.sin(sc.dgamma) / (4 * (sc.dsd - sc.r))
    dz = dr * math.tan(sc.anode_angle)
    if sc.ffs_mode == 1:
        return [(0, da, 0), (0, -da, 0)]
    if sc.ffs_mode == 2:
        return [(-dr, 0, -dz), (dr, 0, dz)]
    return [(-dr, -da, -dz), (-dr, da, -dz), (dr, -da, dz), (dr, da, dz)]


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_ffs_projector_sphere_chords(oracle, mode):
    sc = S.tiny(n_channels=64, n_rows=4, ppr=64, rotations=0.25, ffs_mode=mode).scan
    rho_s = 60.0
    ph = S.Phantom(cyl_radius=1.0, cyl_mu=0.0,
                   ellipsoids=[{"c": (0.0, 0.0, 0.0), "a": (rho_s,) * 3,
                                "rot": np.eye(3).reshape(9).tolist(), "dmu": 1.0}])
    sc = sc.replace(z0=0.0, feed=1e-9)
    raw = oracle.project_phantom(sc, ph)
    offs = _ffs_offsets(sc)
    ref = np.zeros_like(raw, dtype=np.float64)
    for v in range(sc.n_views):
        a = sc.angle_of_view(v)
        ew = np.array([math.cos(a), math.sin(a), 0.0])
        eu = np.array([-math.sin(a), math.cos(a), 0.0])
        ez = np.array([0.0, 0.0, 1.0])
        base = sc.r * ew + sc.z_of_view(v) * ez
        o = offs[v % len(offs)]
        src = base + o[0] * ew + o[1] * eu + o[2] * ez
        for row in range(sc.n_rows):
            w = (row - sc.central_row) * sc.row_spacing
            for c in range(sc.n_channels):
                g = (c - sc.central_channel) * sc.dgamma
                det = base - sc.dsd * (math.cos(g) * ew + math.sin(g) * eu) + w * ez
                u = (det - src) / np.linalg.norm(det - src)
                d2 = np.dot(src, src) - np.dot(src, u) ** 2
                ref[v, row, c] = 2 * math.sqrt(max(rho_s ** 2 - d2, 0.0))
    assert np.abs(raw - ref).max() < 1e-4


def _cylinder_scan(mode):
    w = S.tiny(n_channels=256, n_rows=4, ppr=512, rotations=1.0, ffs_mode=mode)
    return w.scan.replace(feed=1e-9, z0=0.0)


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_ffs_rebin_offcentre_cylinder(oracle, mode):
    sc = _cylinder_scan(mode)
    e = sc.effective()
    cx, cy, a = 40.0, -25.0, 90.0
    ph = S.Phantom(cyl_radius=1.0, cyl_mu=0.0,
                   ellipsoids=[{"c": (cx, cy, 0.0), "a": (a, a, 1e5),
                                "rot": np.eye(3).reshape(9).tolist(), "dmu": 1.0}])
    raw = oracle.project_phantom(sc, ph)
    par = oracle.rebin(sc, raw)
    assert par.shape == (e.n_views, e.n_rows, e.n_channels)
    ds = e.channel_spacing * e.r / e.dsd
    p = (np.arange(e.n_channels) - e.central_channel) * ds
    # Rows: chord / cos(slant) of the undeflected ray (the deflections change
    # the slant by < 1e-4 relative here).
    wq = (np.arange(e.n_rows) - e.central_row) * e.row_spacing
    secant = np.sqrt(e.dsd ** 2 + wq ** 2) / e.dsd

    def expected(dp=0.0, dth=0.0):
        out = np.zeros_like(par)
        for j in range(e.n_views):
            th = e.angle0 + j * 2 * math.pi / e.views_per_rotation + dth
            pc = cx * math.sin(th) - cy * math.cos(th)
            out[j] = (2 * np.sqrt(np.maximum(a * a - (p + dp - pc) ** 2, 0.0)))[None, :] * \
                secant[:, None]
        return out

    ref = expected()
    # Away from the profile's edge kink, and from the ends of the scan where
    # the rebinning clamps to the first / last projection (the fan half-angle
    # spans `edge` parallel views).
    fan_half = math.asin(min(1.0, 0.5 * (sc.n_channels + 4) * sc.dgamma))
    edge = int(math.ceil(fan_half / (2 * math.pi / sc.views_per_rotation) / sc.ffs_states)) + 2
    mask = np.zeros_like(par, bool)
    for j in range(e.n_views):
        th = e.angle0 + j * 2 * math.pi / e.views_per_rotation
        pc = cx * math.sin(th) - cy * math.cos(th)
        mask[j] = (np.abs(p - pc) < a - 10.0)[None, :]
    mask[:edge] = mask[-edge:] = False
    err = np.abs(par - ref)[mask].max()
    assert err < 2e-3 * ref.max()
    # Planted: an interleave off by one interleaved channel, or a parallel
    # view off by one raw projection angle, is detected.
    if mode & 1:
        assert np.abs(par - expected(dp=ds))[mask].max() > 4 * err
    assert np.abs(par - expected(dth=2 * math.pi / sc.views_per_rotation))[mask].max() > 4 * err


def test_ffs_mode_none_is_plain_path(oracle):
    w = S.tiny(n_channels=96, n_rows=4, ppr=96)
    raw = oracle.project_phantom(w.scan, w.phantom)
    a = oracle.rebin(w.scan, raw)
    b = oracle.rebin(w.scan.replace(ffs_mode=0, anode_angle=0.3), raw)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", [1, 3])
def test_ffs_fbp_uniform_cylinder(oracle, mode):
    w = S.tiny(nx=48, n_slices=4, n_channels=96, n_rows=8, ppr=384, rotations=3,
               ffs_mode=mode)
    sc, rc = w.scan, w.recon
    ph = S.Phantom(cyl_radius=100.0, ellipsoids=[])
    raw = oracle.project_phantom(sc, ph)
    d = oracle.geometry(sc, rc)
    vol, _ = oracle.reconstruct_native(sc, rc, raw, 0, d.n_slices_native)
    x = (np.arange(rc.nx) - 0.5 * (rc.nx - 1)) * rc.recon_fov / rc.nx
    X, Y = np.meshgrid(x, x)
    inside = X ** 2 + Y ** 2 < 80 ** 2
    assert abs(vol[:, inside].mean() - S.MU_WATER) < 3e-3 * S.MU_WATER
