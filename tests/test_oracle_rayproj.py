"""Pins of the FP64 ray-driven projector (Joseph's method; the refinement's
line-integral model, no reference counterpart), by what the mathematics
fixes rather than by its formula:

* the projection of a voxelised sphere (each voxel the sphere's volume
  fraction) approaches the analytic chord length, and a sphere moved by whole
  voxels moves its projection accordingly (planted offsets of a voxel fail);
* projector and transpose satisfy <P x, y> = <x, P^T y> to round-off, for a
  counter-clockwise, a clockwise and a phi+z flying-focal-spot scan.
"""
import math

import numpy as np
import pytest

from freect_amd import scans as S


def _setup(ffs_mode=0, rotation=1, nx=40):
    w = S.tiny(n_channels=96, n_rows=8, ppr=96, rotations=1.0, nx=nx, n_slices=10,
               ffs_mode=ffs_mode)
    sc = w.scan.replace(rotation=rotation, angle0=0.4)
    rc = w.recon.replace(recon_fov=200.0)
    return sc, rc


def _voxelised_sphere(oracle, sc, rc, centre, radius, sub=4):
    d = oracle.geometry(sc, rc)
    zs = oracle.slice_grid(rc.start_pos, rc.end_pos, d.csw)
    dx = rc.recon_fov / rc.nx
    xs = (np.arange(rc.nx) - 0.5 * (rc.nx - 1)) * dx + rc.x_origin
    ys = (np.arange(rc.ny) - 0.5 * (rc.ny - 1)) * dx + rc.y_origin
    off = (np.arange(sub) + 0.5) / sub - 0.5
    vol = np.zeros((len(zs), rc.ny, rc.nx))
    for a in off:
        for b in off:
            for c in off:
                Z, Y, X = np.meshgrid(zs + c * d.csw, ys + b * dx, xs + a * dx, indexing="ij")
                vol += ((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2
                        <= radius ** 2)
    return vol / sub ** 3, d


R_GE, DSD_GE = 541.0, 949.0


def _fine(rotation, nx, drow):
    w = S.tiny(n_channels=192, n_rows=16, ppr=96, rotations=1.0, nx=nx, n_slices=10)
    sc = w.scan.replace(row_spacing=drow * DSD_GE / R_GE, angle0=0.4, rotation=rotation)
    rc = w.recon.replace(recon_fov=200.0, start_pos=-20.0, end_pos=20.0, slice_thickness=drow,
                         slice_pitch=drow)
    return sc, rc


@pytest.mark.parametrize("rotation", [1, -1])
def test_voxelised_sphere_projects_to_its_chords(oracle, rotation):
    # The error against the analytic chords shrinks in proportion to the
    # voxel size (2.5 mm -> 1.25 mm voxels and slices), and a sphere one
    # voxel off in x or y projects several times worse.
    centre, radius = (22.0, -31.0, 0.3), 14.0
    ph = S.Phantom(cyl_radius=150.0, cyl_mu=0.0, ellipsoids=[
        {"c": centre, "a": (radius,) * 3, "rot": np.eye(3).reshape(9).tolist(), "dmu": 1.0}])
    rms = {}
    for nx, drow in ((80, 2.5), (160, 1.25)):
        sc, rc = _fine(rotation, nx, drow)
        vol, _ = _voxelised_sphere(oracle, sc, rc, centre, radius)
        ref = oracle.project_phantom(sc, ph).astype(np.float64)
        hit = ref > 0.25 * radius
        err = oracle.ray_project(sc, rc, vol)[hit] - ref[hit]
        rms[nx] = float(np.sqrt(np.mean(err ** 2)))
        print("voxel %.2f mm: sphere chords rms %.3f mm over %d rays (diameter %.0f mm)"
              % (drow, rms[nx], hit.sum(), 2 * radius))
        assert hit.sum() > 1000
    assert rms[160] < 0.02 * 2 * radius and rms[80] / rms[160] > 1.6
    dx = rc.recon_fov / rc.nx
    for shift in ((dx, 0.0), (0.0, dx)):
        c2 = (centre[0] + shift[0], centre[1] + shift[1], centre[2])
        v2, _ = _voxelised_sphere(oracle, sc, rc, c2, radius)
        e2 = oracle.ray_project(sc, rc, v2)[hit] - ref[hit]
        assert np.sqrt(np.mean(e2 ** 2)) > 2 * rms[160]


@pytest.mark.parametrize("ffs_mode,rotation", [(0, 1), (0, -1), (3, 1)])
def test_ray_projector_dot_product(oracle, ffs_mode, rotation):
    sc, rc = _setup(ffs_mode=ffs_mode, rotation=rotation, nx=24)
    d = oracle.geometry(sc, rc)
    rng = np.random.default_rng(ffs_mode + 10 * (rotation + 1))
    x = rng.standard_normal((d.n_slices_native, rc.ny, rc.nx))
    y = rng.standard_normal((sc.n_views, sc.n_rows, sc.n_channels))
    px = oracle.ray_project(sc, rc, x)
    pty = oracle.ray_backproject(sc, rc, y, d.n_slices_native)
    a, b = float((px * y).sum()), float((x * pty).sum())
    assert abs(a - b) <= 1e-11 * max(abs(a), 1.0) * 100
    assert np.abs(px).max() > 0
