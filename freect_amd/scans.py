"""Scan geometries, recon configurations and the seeded synthetic phantom.

The scan description carries what FreeCT's ``fct::RawDataSet`` exposes
(RawDataSet.h:91-121): detector rows/channels, transverse/axial spacing at the
detector, source-detector and source-isocentre distances, central row/channel,
projections per rotation, total projections, and per-view tube angle / table
position -- here as a linear law (angle0 + v*2pi/ppr, z0 + v*feed/ppr).

The recon configuration carries the YAML keys read by parse_config.cpp:39-57
plus the new keys of this engine (filter, precision).

Workloads (named by what they are):
  small     GE-like 736x16, 1152 views/rot, pitch 1.0, 2 rotations, 256^2 x 32
            slices at 1.25 mm, FOV 250, Shepp-Logan, noiseless.
  full      GE-like 888x64, 984 views/rot, pitch 0.984, 10 rotations (9840
            views), 512^2 x 320 slices at 0.625 mm, FOV 500, ramp, I0 = 5e4.
  ffs       Siemens-like 736x32 with combined phi/z flying focal spot, 512^2 x
            200 slices of 1.0 mm over 200 mm (partly outside the helix).
  batch     24 seeded uneven jobs on the 736x16 detector of `small`.
  adjoint   736x16 detector of `small` at 984 views/rot, pitch 1.0, 3
            rotations (2952 views), 512^2 x 64 slices at 0.625 mm.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

MU_WATER = 0.0193  # 1/mm; 1 HU = MU_WATER/1000


def to_hu(mu):
    return 1000.0 * (np.asarray(mu) - MU_WATER) / MU_WATER


# Filter name -> (kind, c, a).  kind 0 is the generalised-Hamming ramp of
# ReconstructionMachine.cu:493-506 / ramp.m:7-10 (c: cutoff, a: Hamming a);
# kind 1 is the Shepp-Logan kernel.  sharp/medium/smooth are the parameter
# sets of FreeCT's resources/filters/*.txt vectors.
FILTERS = {
    "ramp": (0, 1.0, 1.0),
    "sharp": (0, 1.0, 1.0),
    "medium": (0, 0.65, 0.5),
    "smooth": (0, 0.4, 0.5),
    "hann": (0, 1.0, 0.5),
    "hamming": (0, 1.0, 0.54),
    "shepp-logan": (1, 1.0, 1.0),
}


@dataclass
class ScanGeometry:
    n_channels: int
    n_rows: int
    n_views: int
    views_per_rotation: int
    channel_spacing: float      # mm at the detector (arc length)
    row_spacing: float          # mm at the detector
    dsd: float                  # source-detector distance (mm)
    r: float                    # source-isocentre distance (mm)
    central_channel: float
    central_row: float
    angle0: float = 0.0         # tube angle of view 0 (rad)
    z0: float = 0.0             # source z of view 0 (mm)
    feed: float = 0.0           # table feed per rotation (mm, signed)
    ffs_mode: int = 0           # 0 none, 1 phi, 2 z, 3 phi+z (focal state = view mod S)
    anode_angle: float = 0.0    # rad (z-FFS deflection)
    # +1 counter-clockwise (tube angle grows with the view), -1 clockwise: the
    # mirror image y -> -y of the counter-clockwise scan with angles -a_v
    # (channel fan angles and phi-FFS deflections run with the rotation).
    rotation: int = 1

    @property
    def ffs_states(self):
        return {0: 1, 1: 2, 2: 2, 3: 4}[self.ffs_mode]

    def effective(self) -> "ScanGeometry":
        """The interleaved detector the reconstruction works on: phi FFS
        doubles the channels (spacing / 2, central channel 2 cc + 1/2), z FFS
        the rows (spacing / 2, central row 2 cr + 1/2), and S focal states
        make n_views // S parallel views of S projection angles each."""
        S = self.ffs_states
        phi, z = bool(self.ffs_mode & 1), bool(self.ffs_mode & 2)
        return dataclasses.replace(
            self,
            n_channels=self.n_channels * (2 if phi else 1),
            channel_spacing=self.channel_spacing / (2 if phi else 1),
            central_channel=2 * self.central_channel + 0.5 if phi else self.central_channel,
            n_rows=self.n_rows * (2 if z else 1),
            row_spacing=self.row_spacing / (2 if z else 1),
            central_row=2 * self.central_row + 0.5 if z else self.central_row,
            n_views=self.n_views // S, views_per_rotation=self.views_per_rotation // S,
            ffs_mode=0)

    @property
    def dgamma(self):
        return self.channel_spacing / self.dsd

    @property
    def csw(self):
        return self.r * self.row_spacing / self.dsd

    def z_of_view(self, v):
        return self.z0 + v * self.feed / self.views_per_rotation

    def angle_of_view(self, v):
        sgn = -1.0 if self.rotation < 0 else 1.0
        return self.angle0 + sgn * v * 2.0 * math.pi / self.views_per_rotation

    def replace(self, **kw):
        return dataclasses.replace(self, **kw)


@dataclass
class ReconConfig:
    start_pos: float
    end_pos: float
    recon_fov: float
    slice_thickness: float
    slice_pitch: float = 0.0    # 0 -> slice_thickness (parse_config.cpp:59-60)
    nx: int = 512
    ny: int = 512
    x_origin: float = 0.0
    y_origin: float = 0.0
    filter: str = "ramp"
    precision: str = "fp32"     # fp32 | fp16 (filtered projections storage)
    adaptive_filtration_s: float = 0.0   # strength, used when a plan enables it

    @property
    def filter_kind(self):
        return FILTERS[self.filter][0]

    @property
    def filter_c(self):
        return FILTERS[self.filter][1]

    @property
    def filter_a(self):
        return FILTERS[self.filter][2]

    def replace(self, **kw):
        return dataclasses.replace(self, **kw)


@dataclass
class Phantom:
    cyl_radius: float = 150.0
    cyl_mu: float = MU_WATER
    ellipsoids: List[dict] = field(default_factory=list)

    def mu_at(self, x, y, z):
        """Attenuation at points (broadcast arrays), for error-vs-truth checks."""
        x, y, z = np.broadcast_arrays(np.asarray(x, float), np.asarray(y, float),
                                      np.asarray(z, float))
        mu = np.where(x * x + y * y <= self.cyl_radius ** 2, self.cyl_mu, 0.0)
        for e in self.ellipsoids:
            R = np.asarray(e["rot"], float).reshape(3, 3)
            d = np.stack([x - e["c"][0], y - e["c"][1], z - e["c"][2]], -1)
            loc = d @ R  # R^T d per point
            r2 = ((loc / np.asarray(e["a"], float)) ** 2).sum(-1)
            mu = mu + np.where(r2 <= 1.0, e["dmu"], 0.0)
        return mu


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): seed -> stream of 64-bit words."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = seed & self.MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def choice(self, seq):
        return seq[self.next() % len(seq)]

    def randint(self, lo, hi):  # inclusive
        return lo + self.next() % (hi - lo + 1)


def euler_zyz(phi, theta, psi):
    """R = Rz(phi) Ry(theta) Rz(psi) (intrinsic z-y-z Euler angles)."""
    def rz(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])

    def ry(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, 0, s], [0, 1.0, 0], [-s, 0, c]])

    def mul(A, B):  # same summation order as the C++ host (fct_host.cpp matmul3)
        C = np.empty((3, 3))
        for i in range(3):
            for j in range(3):
                C[i, j] = A[i, 0] * B[0, j] + A[i, 1] * B[1, j] + A[i, 2] * B[2, j]
        return C

    return mul(mul(rz(phi), ry(theta)), rz(psi))


def make_phantom(seed: int, z_lo: float, z_hi: float, n: int = 20,
                 cyl_radius: float = 150.0) -> Phantom:
    """Water cylinder (radius 150 mm, 0 HU) + n ellipsoids.

    Centres uniform inside the cylinder cross-section (area-uniform) with z
    uniform in [z_lo, z_hi]; semi-axes uniform in [5, 60] mm; z-y-z Euler
    angles uniform in [0,2pi) x [0,pi) x [0,2pi); values uniform in
    [-900, 1500] HU entered as additive dmu = MU_WATER * HU / 1000.
    """
    rng = SplitMix64(seed)
    ells = []
    for _ in range(n):
        rad = cyl_radius * math.sqrt(rng.uniform())
        ang = rng.uniform(0.0, 2.0 * math.pi)
        c = (rad * math.cos(ang), rad * math.sin(ang), rng.uniform(z_lo, z_hi))
        a = (rng.uniform(5, 60), rng.uniform(5, 60), rng.uniform(5, 60))
        R = euler_zyz(rng.uniform(0, 2 * math.pi), rng.uniform(0, math.pi),
                      rng.uniform(0, 2 * math.pi))
        hu = rng.uniform(-900.0, 1500.0)
        ells.append({"c": c, "a": a, "rot": R.reshape(9).tolist(),
                     "dmu": MU_WATER * hu / 1000.0, "hu": hu})
    return Phantom(cyl_radius=cyl_radius, cyl_mu=MU_WATER, ellipsoids=ells)


# ---------------------------------------------------------------------------
# Detector presets.
# ---------------------------------------------------------------------------
def _ge(n_channels, n_rows, dgamma, ppr):
    R, Dsd = 541.0, 949.0
    return dict(n_channels=n_channels, n_rows=n_rows, views_per_rotation=ppr,
                channel_spacing=dgamma * Dsd, row_spacing=0.625 * Dsd / R, dsd=Dsd, r=R,
                central_channel=0.5 * (n_channels - 1) + 0.25,
                central_row=0.5 * (n_rows - 1))


def helical(det: dict, pitch: float, rotations: float, angle0: float = 0.0,
            n_views: int = None) -> ScanGeometry:
    ppr = det["views_per_rotation"]
    nv = n_views if n_views is not None else int(round(rotations * ppr))
    csw = det["r"] * det["row_spacing"] / det["dsd"]
    feed = pitch * det["n_rows"] * csw
    z0 = -0.5 * (nv - 1) * feed / ppr  # scan centred on z = 0
    return ScanGeometry(n_views=nv, angle0=angle0, z0=z0, feed=feed, **det)


def ge_small_detector():
    # Delta gamma chosen for a 500 mm acquisition FOV: 2 asin(250/R) / 735.
    return _ge(736, 16, 2.0 * math.asin(250.0 / 541.0) / 735.0, 1152)


def ge_full_detector():
    return _ge(888, 64, 0.0010831, 984)


def siemens_detector():
    # resources/scanners/definitionas.scanner:5-14 (R 595, Dsd 1085.6,
    # fan increment 0.0011844 rad, central channel 366.25), 0.6 mm rows at iso.
    R, Dsd = 595.0, 1085.6
    return dict(n_channels=736, n_rows=32, views_per_rotation=1152,
                channel_spacing=0.0011844 * Dsd, row_spacing=0.6 * Dsd / R, dsd=Dsd, r=R,
                central_channel=366.25, central_row=15.5)


@dataclass
class Workload:
    name: str
    scan: ScanGeometry
    recon: ReconConfig
    phantom: Phantom
    i0: float = 0.0             # 0: noiseless
    seed: int = 0


def small(seed: int = 1000) -> Workload:
    scan = helical(ge_small_detector(), pitch=1.0, rotations=2)
    rc = ReconConfig(start_pos=-20.0, end_pos=20.0, recon_fov=250.0, slice_thickness=1.25,
                     slice_pitch=1.25, nx=256, ny=256, filter="shepp-logan")
    return Workload("small", scan, rc, make_phantom(seed, -20.0, 20.0), 0.0, seed)


def full(seed: int = 1001) -> Workload:
    scan = helical(ge_full_detector(), pitch=0.984, rotations=10)
    rc = ReconConfig(start_pos=-100.0, end_pos=100.0, recon_fov=500.0, slice_thickness=0.625,
                     slice_pitch=0.625, nx=512, ny=512, filter="ramp")
    return Workload("full", scan, rc, make_phantom(seed, -100.0, 100.0), 5e4, seed)


def adjoint(seed: int = 1004) -> Workload:
    """The small scan's 736 x 16 detector at 984 views per rotation (3
    rotations = 2952 views), pitch 1.0, noiseless; 512^2 x 64 native slices
    of 0.625 mm (40 mm, of which the 30 mm of table travel covers the
    central part fully)."""
    det = ge_small_detector()
    det["views_per_rotation"] = 984
    scan = helical(det, pitch=1.0, rotations=3)
    rc = ReconConfig(start_pos=-20.0, end_pos=20.0, recon_fov=500.0, slice_thickness=0.625,
                     slice_pitch=0.625, nx=512, ny=512, filter="ramp")
    return Workload("adjoint", scan, rc, make_phantom(seed, -20.0, 20.0), 0.0, seed)


def ffs(seed: int = 1002) -> Workload:
    """Siemens-like scan with combined phi + z flying focal spot: 736 x 32
    rows (0.6 mm at iso), 2304 projections per rotation in a 4-state focal
    cycle (interleaved to 1472 channels x 64 rows, 576 parallel views per
    rotation), pitch 0.8, 8 rotations, Poisson noise at I0 = 2.5e4; 512^2 x 200
    slices of 1.0 mm at 1.0 mm spacing over z in [-100, 100) mm, of which the
    123 mm of table travel covers the central part (HU metrics on the covered
    mask), sharp kernel (c = 1, a = 1)."""
    det = siemens_detector()
    det["views_per_rotation"] = 2304
    scan = helical(det, pitch=0.8, rotations=8).replace(ffs_mode=3, anode_angle=0.1222)
    rc = ReconConfig(start_pos=-100.0, end_pos=100.0, recon_fov=500.0, slice_thickness=1.0,
                     slice_pitch=1.0, nx=512, ny=512, filter="sharp")
    return Workload("ffs", scan, rc, make_phantom(seed, -100.0, 100.0), 2.5e4, seed)


def batch_jobs(seed: int = 1003, n_jobs: int = 24) -> List[Workload]:
    """Uneven jobs on the 736x16 detector of `small` (1152 views/rot): scan
    lengths 2-20 rotations, pitch in {0.516, 0.984, 1.375}, matrix in
    {256^2, 512^2}, 16-240 slices of {0.625, 1.25, 2.5} mm, filters {ramp,
    shepp-logan, hann}, I0 in [1e4, 1e5].  The slice count drawn is cut to the
    z-range the helix covers fully (at least 16 slices, which may reach past
    it on the shortest scans)."""
    rng = SplitMix64(seed)
    det = ge_small_detector()
    jobs = []
    for j in range(n_jobs):
        rot = rng.randint(2, 20)
        pitch = rng.choice([0.516, 0.984, 1.375])
        n = rng.choice([256, 512])
        thick = rng.choice([0.625, 1.25, 2.5])
        filt = rng.choice(["ramp", "shepp-logan", "hann"])
        i0 = rng.uniform(1e4, 1e5)
        scan = helical(det, pitch=pitch, rotations=rot)
        # Slices inside the fully covered z-range of the scan.
        half_cov = 0.5 * (scan.n_views - 1) * scan.feed / scan.views_per_rotation
        margin = 0.5 * det["n_rows"] * scan.csw * (det["r"] + 250.0) / det["r"] + scan.feed
        zmax = max(half_cov - margin, 5.0)
        n_sl = rng.randint(16, 240)
        span = min(n_sl * thick, 2 * zmax)
        n_sl = max(16, int(span / thick))
        rc = ReconConfig(start_pos=-0.5 * n_sl * thick, end_pos=0.5 * n_sl * thick,
                         recon_fov=500.0, slice_thickness=thick, slice_pitch=thick, nx=n,
                         ny=n, filter=filt)
        ph = make_phantom(seed * 100 + j, rc.start_pos, rc.end_pos)
        jobs.append(Workload(f"batch{j}", scan, rc, ph, i0, seed * 100 + j))
    return jobs


def tiny(seed: int = 7, n_channels=96, n_rows=8, ppr=96, rotations=2.5, nx=32,
         n_slices=8, filter="ramp", ffs_mode=0, anode_angle=0.1222) -> Workload:
    """A scaled-down GE-like scan for fast oracle tests."""
    R, Dsd = 541.0, 949.0
    dgamma = 2.0 * math.asin(250.0 / R) / (n_channels - 1)
    det = dict(n_channels=n_channels, n_rows=n_rows, views_per_rotation=ppr,
               channel_spacing=dgamma * Dsd, row_spacing=5.0 * Dsd / R, dsd=Dsd, r=R,
               central_channel=0.5 * (n_channels - 1) + 0.25,
               central_row=0.5 * (n_rows - 1))
    scan = helical(det, pitch=1.0, rotations=rotations, angle0=0.1)
    if ffs_mode:
        scan = scan.replace(ffs_mode=ffs_mode, anode_angle=anode_angle)
    csw = scan.effective().csw
    rc = ReconConfig(start_pos=-0.5 * n_slices * csw, end_pos=0.5 * n_slices * csw,
                     recon_fov=400.0, slice_thickness=csw, slice_pitch=csw, nx=nx, ny=nx,
                     filter=filter)
    return Workload("tiny", scan, rc, make_phantom(seed, rc.start_pos, rc.end_pos, n=6),
                    0.0, seed)
