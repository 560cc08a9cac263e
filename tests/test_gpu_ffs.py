"""Flying focal spot on the GPU against the FP64 restatement (same interleave).

Tolerances as for the plain path: projector to a few FP32 ulps, rebin +
filter within the FP32 FFT round-off bound, rebinning transpose to FP32
round-off, volumes within 0.5 HU RMSE / 2 HU max on the fully covered voxels.
"""
import math

import numpy as np
import pytest

from freect_amd import scans as S
from helpers import fov_mask, gpu_ctx, say, to_np

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return gpu_ctx()


def _w(mode, **kw):
    args = dict(n_channels=96, n_rows=8, ppr=192, rotations=3.0, nx=40, n_slices=10)
    args.update(kw)
    return S.tiny(ffs_mode=mode, **args)


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_ffs_phantom_matches_oracle(ctx, oracle, mode):
    w = _w(mode)
    say("ffs phantom mode", mode)
    raw = to_np(ctx.phantom_project(w.scan, w.phantom, 0.0, w.seed))
    ref = oracle.project_phantom(w.scan, w.phantom)
    assert np.max(np.abs(raw - ref)) <= 4 * np.spacing(np.float32(np.abs(ref).max()))


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_ffs_rebin_filter_matches_oracle(ctx, oracle, mode):
    from freect_amd import wfbp
    w = _w(mode, filter="shepp-logan")
    sc, rc = w.scan, w.recon
    e = sc.effective()
    raw_t = ctx.phantom_project(sc, w.phantom, 0.0, w.seed)
    p = wfbp.Plan(ctx, sc, rc)
    i = p.info
    assert (i.n_parallel_views, i.n_filtered_rows, i.n_filtered_channels, i.ffs_states) == \
        (e.n_views, e.n_rows, e.n_channels, sc.ffs_states)
    filt_t, _ = p.alloc_filtered()
    p.rebin_filter(raw_t, filt_t)
    p.synchronize()
    got = to_np(p.filtered_view(filt_t))[..., : e.n_channels]
    d = oracle.geometry(sc, rc)
    x = oracle.rebin(sc, to_np(raw_t))
    h = oracle.filter_kernel(d.nfft, d.ds, rc.filter_kind, rc.filter_c, rc.filter_a)
    ref = oracle.filter_rows(x, h, d.nfft, use_fft=False)
    Hk = np.real(np.fft.fft(np.fft.ifftshift(h)))
    tol = 1e-7 * math.log2(d.nfft) * np.abs(Hk).max() * np.abs(x).max()
    err = np.max(np.abs(got - ref))
    say("ffs rebin_filter mode %d nfft %d err %.3g tol %.3g" % (mode, d.nfft, err, tol))
    assert err < tol


def test_ffs_rebin_adjoint_matches_oracle(ctx, oracle):
    from freect_amd import wfbp
    w = _w(3)
    sc, rc = w.scan, w.recon
    e = sc.effective()
    p = wfbp.Plan(ctx, sc, rc)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((e.n_views, e.n_rows, p.info.channel_pitch)).astype(np.float32)
    g[..., e.n_channels:] = 0.0
    g_t = torch.from_numpy(g).cuda()
    pbar = p.alloc_raw()
    p.rebin_adjoint(g_t, pbar)
    p.synchronize()
    ref = oracle.rebin_adjoint(sc, g[..., : e.n_channels].astype(np.float64))
    got = to_np(pbar)
    assert got.shape == ref.shape == (sc.n_views, sc.n_rows, sc.n_channels)
    err = np.abs(got - ref).max()
    say("ffs rebin adjoint err", err, "max", np.abs(ref).max())
    assert err <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("mode", [1, 3])
def test_ffs_reconstruct_matches_oracle(ctx, oracle, mode):
    from freect_amd import wfbp
    w = _w(mode)
    sc, rc = w.scan, w.recon
    d = oracle.geometry(sc, rc)
    raw_t = ctx.phantom_project(sc, w.phantom, 0.0, w.seed)
    p = wfbp.Plan(ctx, sc, rc)
    vol = p.alloc_volume()
    p.reconstruct(raw_t, vol)
    p.synchronize()
    ref_nat, n_sel = oracle.reconstruct_native(sc, rc, to_np(raw_t), 0, d.n_slices_native)
    ref = oracle.thicken(sc, rc, ref_nat)
    cov = oracle.covered_mask(sc, rc) & fov_mask(rc, d.acq_fov)[None]
    d_hu = (to_np(vol).astype(np.float64) - ref) / (S.MU_WATER / 1000.0)
    rmse, mx = float(np.sqrt(np.mean(d_hu[cov] ** 2))), float(np.abs(d_hu[cov]).max())
    say("ffs recon mode %d covered %d: rmse %.4f max %.4f HU" % (mode, cov.sum(), rmse, mx))
    assert cov.sum() > 0.2 * cov.size
    assert rmse <= 0.5 and mx <= 2.0
    assert p.count_voxel_views() == n_sel


def test_ffs_plan_errors(ctx):
    from freect_amd import wfbp
    w = _w(3)
    with pytest.raises(wfbp.FctError):
        wfbp.Plan(ctx, w.scan.replace(views_per_rotation=196), w.recon)   # not a multiple of 8
    with pytest.raises(wfbp.FctError):
        wfbp.Plan(ctx, w.scan.replace(anode_angle=0.0), w.recon)
