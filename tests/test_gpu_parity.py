"""GPU parity of the HIP path against the FP64 CPU restatement (oracle/).

Everything runs through the C ABI (include/fct_wfbp_gpu.h) via ctypes.
Tolerances: index selection bit-exact; FP32 GPU volumes within 0.5 HU RMSE
and 2 HU max of the FP64 oracle; filtered projections within the FP32 FFT
round-off bound eps*log2(N)*max|H|*max|x| of the FP64 direct sums.
"""
import math

import numpy as np
import pytest

from freect_amd import scans as S
from helpers import fov_mask, gpu_ctx, hu_errors, say, to_np

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return gpu_ctx()


def _plan(ctx, w, **kw):
    from freect_amd import wfbp
    return wfbp.Plan(ctx, w.scan, w.recon, **kw)


def _raw_gpu(ctx, w, i0=None):
    return ctx.phantom_project(w.scan, w.phantom, w.i0 if i0 is None else i0, w.seed)


def test_phantom_matches_oracle(ctx, oracle):
    w = S.tiny()
    say("phantom tiny", w.scan.n_views, w.scan.n_rows, w.scan.n_channels)
    raw = to_np(_raw_gpu(ctx, w, 0.0))
    ref = oracle.project_phantom(w.scan, w.phantom)
    assert np.max(np.abs(raw - ref)) <= 4 * np.spacing(np.float32(np.abs(ref).max()))


def test_phantom_noise_statistics(ctx):
    # Uniform line integral p: the noisy values average to ~p with variance
    # ~ exp(p)/I0 (delta method on ln of a Poisson count).
    w = S.tiny(n_channels=128, n_rows=8, ppr=128, rotations=4)
    ph = S.Phantom(cyl_radius=1000.0, cyl_mu=0.0, ellipsoids=[])
    clean = to_np(ctx.phantom_project(w.scan, ph, 0.0, 1))
    assert np.abs(clean).max() == 0.0
    i0 = 1e4
    noisy = to_np(ctx.phantom_project(w.scan, ph, i0, 123))
    assert abs(noisy.mean()) < 3 * math.sqrt(1.0 / i0 / noisy.size) + 1e-4
    assert abs(noisy.var() * i0 - 1.0) < 0.05
    again = to_np(ctx.phantom_project(w.scan, ph, i0, 123))
    assert np.array_equal(noisy, again)  # seeded: deterministic


@pytest.mark.parametrize("filt", ["ramp", "shepp-logan", "hann", "medium"])
def test_rebin_filter_matches_oracle(ctx, oracle, filt):
    w = S.tiny(filter=filt)
    sc, rc = w.scan, w.recon
    say("rebin_filter", filt)
    raw_t = _raw_gpu(ctx, w, 0.0)
    p = _plan(ctx, w)
    filt_t, _ = p.alloc_filtered()
    p.rebin_filter(raw_t, filt_t)
    p.synchronize()
    got = to_np(p.filtered_view(filt_t))[..., : sc.n_channels]
    raw = to_np(raw_t)
    d = oracle.geometry(sc, rc)
    ref = oracle.rebin(sc, raw)
    h = oracle.filter_kernel(d.nfft, d.ds, rc.filter_kind, rc.filter_c, rc.filter_a)
    ref = oracle.filter_rows(ref, h, d.nfft, use_fft=False)
    # FP32 FFT round-off scales with the input line and the filter gain:
    # |err| <~ eps * log2(N) * max|H| * max|x| (a one-channel shift or a wrong
    # tap is ~1e3 times larger).
    Hk = np.real(np.fft.fft(np.fft.ifftshift(h)))
    x = oracle.rebin(sc, raw)
    tol = 1e-7 * math.log2(d.nfft) * np.abs(Hk).max() * np.abs(x).max()
    err = np.max(np.abs(got - ref))
    say("rebin_filter err %.3g tol %.3g" % (err, tol))
    assert err < tol


@pytest.mark.parametrize("n_channels", [96, 200, 400, 888, 1500])
def test_rebin_filter_fft_sizes(ctx, oracle, n_channels):
    # nfft 256 (generic LDS kernel) and 512 / 1024 / 2048 / 4096 (the
    # register-resident kernel, one instantiation each), FP32 against the
    # FP64 oracle, then FP16 / BF16 storage against the FP32 output: the
    # stored value times the scan's power-of-two scale s is within half an ulp
    # of the format (2^-11 / 2^-8 relative, FP16's subnormal spacing 2^-24 s
    # below 2^-14 s) plus the FFT round-off; s is the documented one
    # (max|y| < 2^15 s, 2^14 s <= max|raw| sum|h| (1 + 2^-7)); and every slot
    # of the 2 x 2 quads holds the stored (r, c) value of its neighbour, 0
    # past the last channel of the pitch and past the last row.
    w = S.tiny(n_channels=n_channels, n_rows=4, ppr=96, rotations=2.0)
    sc, rc = w.scan, w.recon
    raw_t = _raw_gpu(ctx, w, 0.0)
    raw = to_np(raw_t)
    d = oracle.geometry(sc, rc)
    say("rebin_filter sizes", n_channels, "nfft", d.nfft)
    x = oracle.rebin(sc, raw)
    h = oracle.filter_kernel(d.nfft, d.ds, rc.filter_kind, rc.filter_c, rc.filter_a)
    ref = oracle.filter_rows(x, h, d.nfft, use_fft=n_channels > 400)
    Hk = np.real(np.fft.fft(np.fft.ifftshift(h)))
    tol = 1e-7 * math.log2(d.nfft) * np.abs(Hk).max() * np.abs(x).max()
    f32 = None
    for prec, rel in (("fp32", 0.0), ("fp16", 2.0 ** -11), ("bf16", 2.0 ** -8)):
        p = _plan(ctx, w, precision=prec)
        filt_t, scale_t = p.alloc_filtered()
        p.rebin_filter(raw_t, filt_t, scale_t)
        p.synchronize()
        full_rows = to_np(p.filtered_view(filt_t).float())
        assert np.all(full_rows[..., sc.n_channels:] == 0.0)   # pitch padding is zero
        got = full_rows[..., : sc.n_channels]
        if prec == "fp32":
            err = np.max(np.abs(got - ref))
            say("  fp32 err %.3g tol %.3g" % (err, tol))
            assert err < tol
            f32 = got
            continue
        s = float(to_np(scale_t)[0])
        bound = np.abs(raw).max() * np.abs(h).sum() * (1 + 2.0 ** -7)
        assert s == 2.0 ** round(math.log2(s))
        assert np.abs(f32).max() < s * 2.0 ** 15 and s * 2.0 ** 14 <= bound * (1 + 1e-6)
        deq = got.astype(np.float64) * s
        err = np.abs(deq - f32) - (rel * np.abs(f32) + s * 2.0 ** -24 + tol)
        say("  %s s %g worst margin %.3g" % (prec, s, err.max()))
        assert np.all(err <= 0.0)
        q = to_np(p.filtered_quads(filt_t).float())
        y = q[..., 0]
        assert np.array_equal(q[:, :, :-1, 1], y[:, :, 1:]) and np.all(q[:, :, -1, 1] == 0)
        assert np.array_equal(q[:, :-1, :, 2], y[:, 1:, :]) and np.all(q[:, -1, :, 2] == 0)
        assert np.array_equal(q[:, :-1, :-1, 3], y[:, 1:, 1:])
        assert np.all(q[:, -1, :, 3] == 0) and np.all(q[:, :, -1, 3] == 0)


@pytest.mark.parametrize("n_channels,n_rows", [(150, 5), (96, 5), (150, 6)])
def test_fp16_quads_every_slot(ctx, oracle, n_channels, n_rows):
    # The FP16 quad writer on both kernels (nfft 512: register kernel, which
    # stores row pairs whole when a line pair is rows 2i, 2i+1 of one view and
    # falls back to 16-bit slot stores when an odd row count makes pairs
    # straddle views; nfft 256: the LDS kernel), checked slot by slot against
    # the stored (r, c) values, and the stored values against the FP32 rows.
    w = S.tiny(n_channels=n_channels, n_rows=n_rows, ppr=96, rotations=2.0)
    sc = w.scan
    raw_t = _raw_gpu(ctx, w, 0.0)
    say("fp16 quads", n_channels, "channels", n_rows, "rows")
    p32 = _plan(ctx, w)
    f32_t, _ = p32.alloc_filtered()
    p32.rebin_filter(raw_t, f32_t)
    p = _plan(ctx, w, precision="fp16")
    filt_t, scale_t = p.alloc_filtered()
    p.rebin_filter(raw_t, filt_t, scale_t)
    p.synchronize()
    f32 = to_np(p32.filtered_view(f32_t)).astype(np.float64)
    q = to_np(p.filtered_quads(filt_t).float())
    s = float(to_np(scale_t)[0])
    y = q[..., 0]
    assert np.array_equal(q[:, :, :-1, 1], y[:, :, 1:]) and np.all(q[:, :, -1, 1] == 0)
    assert np.array_equal(q[:, :-1, :, 2], y[:, 1:, :]) and np.all(q[:, -1, :, 2] == 0)
    assert np.array_equal(q[:, :-1, :-1, 3], y[:, 1:, 1:])
    assert np.all(q[:, -1, :, 3] == 0) and np.all(q[:, :, -1, 3] == 0)
    err = np.abs(y.astype(np.float64) * s - f32) - (2.0 ** -11 * np.abs(f32) + s * 2.0 ** -24)
    assert np.all(err <= 1e-6 * np.abs(f32).max())


@pytest.mark.parametrize("factor", [0.0, 1e-6, 1e6])
def test_fp16_storage_scale_extremes(ctx, oracle, factor):
    # The per-scan storage scale follows the data: a scan multiplied by 1e6
    # (FP16 would overflow without it) or by 1e-6 (subnormal without it)
    # reconstructs to the FP32 volume times the same factor within the
    # mixed-precision error, and an all-zero scan gives scale 1 and zeros.
    w = S.tiny(nx=32, n_slices=8, n_channels=96, n_rows=8, ppr=96, rotations=2.0)
    raw_t = _raw_gpu(ctx, w, 0.0)
    say("fp16 scale factor", factor)
    vols = {}
    for prec in ("fp32", "fp16"):
        p = _plan(ctx, w, precision=prec)
        filt_t, sc_t = p.alloc_filtered()
        p.rebin_filter(raw_t * factor, filt_t, sc_t)
        nat = p.alloc_native()
        p.backproject(filt_t, sc_t, nat)
        p.synchronize()
        vols[prec] = to_np(nat).astype(np.float64)
        if prec == "fp16":
            s = float(to_np(sc_t)[0])
            assert np.isfinite(vols[prec]).all()
            if factor == 0.0:
                assert s == 1.0 and not vols[prec].any()
    if factor:
        ref = vols["fp32"]
        err = np.abs(vols["fp16"] - ref).max() / np.abs(ref).max()
        say("  rel max err %.2e (scale %g)" % (err, s))
        assert err < 2e-3


def test_fp16_direct_variant_reads_quads(ctx, oracle):
    # The direct kernel reads the same FP16 quads (one scale per scan) as the
    # lean kernel: the two volumes agree to FP32 round-off, and both are within
    # the mixed-precision bound of the FP64 restatement.
    w = S.tiny(nx=48, n_slices=12, n_channels=192, n_rows=16, ppr=192, rotations=3)
    sc, rc = w.scan, w.recon
    d = oracle.geometry(sc, rc)
    raw_t = _raw_gpu(ctx, w, 5e4)
    ref, _ = oracle.reconstruct_native(sc, rc, to_np(raw_t), 0, d.n_slices_native)
    vols = {}
    for variant in (1, 2):
        p = _plan(ctx, w, precision="fp16", bp_variant=variant)
        filt_t, sc_t = p.alloc_filtered()
        p.rebin_filter(raw_t, filt_t, sc_t)
        nat = p.alloc_native()
        p.backproject(filt_t, sc_t, nat)
        p.synchronize()
        vols[variant] = to_np(nat).astype(np.float64)
    mask = fov_mask(rc, d.acq_fov)
    diff = hu_errors(vols[1], vols[2], mask)
    say("fp16 direct vs lean: rmse %.2e max %.2e HU" % diff)
    assert diff[1] < 0.01
    for variant in (1, 2):
        rmse, mx = hu_errors(vols[variant], ref, mask)
        say("fp16 variant %d vs fp64: rmse %.4f max %.4f HU" % (variant, rmse, mx))
        assert rmse <= 0.5


@pytest.mark.parametrize("cr_shift", [0.0, 0.37])
def test_selection_bit_exact_counts(ctx, oracle, cr_shift):
    # cr_shift != 0 moves the central row off (nrows-1)/2 and exercises the
    # general selection test of every variant.
    w = S.tiny(nx=48, n_slices=12)
    sc = w.scan.replace(central_row=w.scan.central_row + cr_shift)
    rc = w.recon
    d = oracle.geometry(sc, rc)
    ref, ref_total = oracle.bp_counts(sc, rc, 0, d.n_slices_native)
    for variant in (1, 2):
        for tz in (8, 16, 32):
            p = wfbp_plan(ctx, sc, rc, tile_z=tz, bp_variant=variant)
            counts = torch.zeros((d.n_slices_native, rc.ny, rc.nx), dtype=torch.int32,
                                 device="cuda")
            total = p.count_voxel_views(counts)
            say("counts variant", variant, "tz", tz, total, ref_total)
            assert total == ref_total
            assert np.array_equal(to_np(counts), ref)


def wfbp_plan(ctx, scan, recon, **kw):
    from freect_amd import wfbp
    return wfbp.Plan(ctx, scan, recon, **kw)


def test_selection_bit_exact_per_view(ctx, oracle):
    w = S.small()
    sc, rc = w.scan, w.recon
    p = _plan(ctx, w)
    for (ix, iy, iz) in [(128, 128, 31), (20, 200, 5), (250, 3, 60), (77, 180, 40)]:
        got = to_np(p.voxel_selection(ix, iy, iz))
        ref = oracle.bp_selection(sc, rc, ix, iy, iz)
        say("selection", ix, iy, iz, int((ref >= 0).sum()))
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("variant,tz,cr_shift", [(1, 16, 0.0), (2, 8, 0.0), (2, 16, 0.0),
                                                 (2, 32, 0.0), (2, 16, 0.37)])
def test_backproject_matches_oracle(ctx, oracle, variant, tz, cr_shift):
    # Same filtered input to both (the GPU's FP32 filtered rows): BP only.
    w = S.tiny(nx=64, n_slices=16, n_channels=192, n_rows=16, ppr=192, rotations=3)
    sc = w.scan.replace(central_row=w.scan.central_row + cr_shift)
    rc = w.recon
    d = oracle.geometry(sc, rc)
    raw_t = ctx.phantom_project(sc, w.phantom, 0.0, w.seed)
    p = wfbp_plan(ctx, sc, rc, tile_z=tz, bp_variant=variant)
    filt_t, _ = p.alloc_filtered()
    p.rebin_filter(raw_t, filt_t)
    nat = p.alloc_native()
    p.backproject(filt_t, None, nat)
    p.synchronize()
    filt = to_np(p.filtered_view(filt_t))[..., : sc.n_channels].astype(np.float64)
    ref, _ = oracle.backproject(sc, rc, filt, 0, 0, d.n_slices_native)
    cov = oracle.bp_coverage(sc, rc, 0, d.n_slices_native) >= 0.5
    mask = cov & fov_mask(rc, d.acq_fov)[None]
    g = to_np(nat).astype(np.float64)
    dd = (g - ref)[mask] / (S.MU_WATER / 1000.0)
    rmse, mx = float(np.sqrt(np.mean(dd * dd))), float(np.abs(dd).max())
    rmse_all, mx_all = hu_errors(g, ref, fov_mask(rc, d.acq_fov))
    say("bp variant %d tz %d: covered rmse %.4f max %.4f HU; whole FOV rmse %.4f max %.4f"
        % (variant, tz, rmse, mx, rmse_all, mx_all))
    assert mask.sum() > 0.2 * mask.size
    assert rmse <= 0.5 and mx <= 2.0
    assert rmse_all <= 0.5 and mx_all <= 20.0


def test_thicken_matches_oracle(ctx, oracle):
    w = S.tiny(nx=32, n_slices=12)
    d0 = oracle.geometry(w.scan, w.recon)
    rc = w.recon.replace(slice_thickness=2.5 * d0.csw, slice_pitch=1.5 * d0.csw)
    w = S.Workload("thk", w.scan, rc, w.phantom)
    p = _plan(ctx, w)
    rng = np.random.default_rng(0)
    native = rng.standard_normal((p.info.n_slices_native, rc.ny, rc.nx)).astype(np.float32)
    vol = p.alloc_volume()
    p.thicken(torch.from_numpy(native).cuda(), vol)
    p.synchronize()
    ref = oracle.thicken(w.scan, rc, native.astype(np.float64))
    assert np.max(np.abs(to_np(vol) - ref)) < 1e-5


def test_small_workload_end_to_end(ctx, oracle):
    """The small GE-like scan: 736x16, 2304 views, 256^2 x 32 slices at 1.25 mm,
    Shepp-Logan."""
    w = S.small()
    sc, rc = w.scan, w.recon
    d = oracle.geometry(sc, rc)
    say("small e2e: native", d.n_slices_native, "requested", d.n_slices_requested)
    raw_t = _raw_gpu(ctx, w)
    p = _plan(ctx, w)
    vol = p.alloc_volume()
    p.reconstruct(raw_t, vol)
    p.synchronize()
    raw = to_np(raw_t)
    ref_nat, n_sel = oracle.reconstruct_native(sc, rc, raw, 0, d.n_slices_native)
    ref = oracle.thicken(sc, rc, ref_nat)
    # HU metrics on the fully covered voxels (every theta~ has weight sum >=
    # 0.5); the 40 mm slice range of this workload exceeds the 20 mm helical
    # travel, and where only |q| ~ 1 views reach a voxel the FP32 and FP64
    # weights W(q) ~ 0 differ in relative terms.  The whole-FOV error is
    # reported and bounded loosely.
    cov = oracle.covered_mask(sc, rc) & fov_mask(rc, d.acq_fov)[None]
    g = to_np(vol)
    d_hu = (g.astype(np.float64) - ref) / (S.MU_WATER / 1000.0)
    rmse, mx = float(np.sqrt(np.mean(d_hu[cov] ** 2))), float(np.abs(d_hu[cov]).max())
    rmse_all, mx_all = hu_errors(g, ref, fov_mask(rc, d.acq_fov))
    say("small e2e covered (%d voxels): rmse %.4f max %.4f HU; whole FOV rmse %.4f max %.4f;"
        " voxel-views %d" % (cov.sum(), rmse, mx, rmse_all, mx_all, n_sel))
    assert cov.sum() > 0.2 * cov.size
    assert rmse <= 0.5 and mx <= 2.0
    assert rmse_all <= 0.5 and mx_all <= 20.0
    assert p.count_voxel_views() == n_sel


def test_plan_moved_to_another_stream(ctx):
    # Plan creation queues the backprojection's z-interval table on the
    # plan's stream; set_stream waits for it, so a plan moved at once to
    # another stream reconstructs bitwise what a plan kept on one stream does.
    w = S.tiny(nx=48, n_slices=12, n_channels=192, n_rows=16, ppr=192, rotations=3)
    raw_t = _raw_gpu(ctx, w)
    torch.cuda.synchronize()
    ref_plan = _plan(ctx, w, precision="fp16")
    ref = ref_plan.alloc_volume()
    ref_plan.reconstruct(raw_t, ref)
    ref_plan.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    p = _plan(ctx, w, precision="fp16", stream=s1)
    p.set_stream(s2)
    vol = p.alloc_volume()
    p.reconstruct(raw_t, vol)
    p.synchronize()
    assert torch.equal(vol, ref)


@pytest.mark.parametrize("precision,rmse_bound", [("fp16", 0.5), ("bf16", 8.0)])
def test_mixed_precision_error(ctx, oracle, precision, rmse_bound):
    w = S.tiny(nx=64, n_slices=16, n_channels=192, n_rows=16, ppr=192, rotations=3)
    sc, rc = w.scan, w.recon
    d = oracle.geometry(sc, rc)
    raw_t = _raw_gpu(ctx, w, 5e4)
    ref, _ = oracle.reconstruct_native(sc, rc, to_np(raw_t), 0, d.n_slices_native)
    p = _plan(ctx, w, precision=precision)
    filt_t, sc_t = p.alloc_filtered()
    p.rebin_filter(raw_t, filt_t, sc_t)
    nat = p.alloc_native()
    p.backproject(filt_t, sc_t, nat)
    p.synchronize()
    rmse, mx = hu_errors(to_np(nat), ref, fov_mask(rc, d.acq_fov))
    say(f"{precision}: rmse {rmse:.4f} max {mx:.4f} HU")
    assert rmse <= rmse_bound


def test_error_paths(ctx):
    from freect_amd import wfbp
    w = S.tiny()
    with pytest.raises(wfbp.FctError, match="views_per_rotation"):
        wfbp.Plan(ctx, w.scan.replace(views_per_rotation=95), w.recon)
    with pytest.raises(wfbp.FctError, match="nx, ny"):
        wfbp.Plan(ctx, w.scan, w.recon.replace(nx=0))
    with pytest.raises(wfbp.FctError, match="tile_z"):
        wfbp.Plan(ctx, w.scan, w.recon, tile_z=12)
    with pytest.raises(wfbp.FctError, match="bp_variant"):
        wfbp.Plan(ctx, w.scan, w.recon, bp_variant=9)
    p = wfbp.Plan(ctx, w.scan, w.recon, precision="fp16")
    with pytest.raises(wfbp.FctError, match="scale buffer"):
        p.backproject(1, None, 1)


def test_rebinning_limits_are_errors(ctx):
    # Scans outside the rebinning's addressing are refused at plan creation
    # (FCT_ERR_LIMIT, before any device buffer is allocated): fan view offsets
    # beyond the packed table's -2048..2047 (a 64-channel fan of +-0.6 rad
    # sampled 30000 times per rotation: offsets near +-2900), and a raw view
    # group of 16 MiB (2048 rows x 2048 channels of floats).
    from freect_amd import wfbp
    w = S.tiny()
    sc = w.scan
    wide = sc.replace(n_channels=64, central_channel=31.5, channel_spacing=0.6 / 32 * sc.dsd,
                      views_per_rotation=30000, n_views=30000)
    with pytest.raises(wfbp.FctError, match="packed form"):
        wfbp.Plan(ctx, wide, w.recon)
    big = sc.replace(n_rows=2048, central_row=1023.5, row_spacing=0.01, n_channels=2048,
                     central_channel=1023.5, channel_spacing=0.3, views_per_rotation=16,
                     n_views=20)
    with pytest.raises(wfbp.FctError, match="raw view group"):
        wfbp.Plan(ctx, big, w.recon)
    # A plan within the limits still works after the refusals.
    wfbp.Plan(ctx, sc, w.recon)


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_non_finite_raw_reads_as_zero(ctx, precision):
    # A dead (NaN) or saturated (+/-Inf) detector reading enters the
    # rebinning as 0: the volume is bitwise that of the scan with those
    # samples set to 0 (finite everywhere; the FP16 storage scale ignores
    # them too).
    w = S.tiny(nx=48, n_slices=12, n_channels=192, n_rows=16, ppr=192, rotations=3)
    raw_t = _raw_gpu(ctx, w, 5e4)
    bad = raw_t.clone()
    zero = raw_t.clone()
    spots = [(10, 3, 50), (300, 0, 191), (301, 15, 0), (500, 8, 96)]
    for (v, r, c), val in zip(spots, [float("inf"), float("nan"), -float("inf"), float("nan")]):
        bad[v, r, c] = val
        zero[v, r, c] = 0.0
    vols = []
    for src in (bad, zero):
        p = _plan(ctx, w, precision=precision)
        vol = p.alloc_volume()
        p.reconstruct(src, vol)
        p.synchronize()
        vols.append(vol)
    say("non-finite raw", precision, "finite:", bool(torch.isfinite(vols[0]).all()))
    assert torch.isfinite(vols[0]).all()
    assert torch.equal(vols[0], vols[1])


@pytest.mark.parametrize("ffs_mode", [0, 3])
def test_clockwise_scan_matches_oracle(ctx, oracle, ffs_mode):
    # A clockwise scan (tube angle angle0 - v dtheta, fan angles and phi-FFS
    # deflections running with the rotation): the GPU projector against the
    # FP64 one (which traces the clockwise rays directly), then the whole
    # chain against the FP64 restatement (which reconstructs the mirrored
    # counter-clockwise scan and reverses the rows), with an off-centre grid.
    w = S.tiny(nx=48, n_slices=12, n_channels=192, n_rows=16, ppr=192, rotations=3,
               ffs_mode=ffs_mode)
    sc = w.scan.replace(rotation=-1, angle0=0.9)
    rc = w.recon.replace(x_origin=-3.0, y_origin=7.0)
    d = oracle.geometry(sc, rc)
    raw_t = ctx.phantom_project(sc, w.phantom, 0.0, w.seed)
    raw = to_np(raw_t)
    ref_raw = oracle.project_phantom(sc, w.phantom)
    assert np.max(np.abs(raw - ref_raw)) <= 4 * np.spacing(np.float32(np.abs(ref_raw).max()))
    p = wfbp_plan(ctx, sc, rc)
    filt_t, _ = p.alloc_filtered()
    p.rebin_filter(raw_t, filt_t)
    nat = p.alloc_native()
    p.backproject(filt_t, None, nat)
    counts = torch.zeros((d.n_slices_native, rc.ny, rc.nx), dtype=torch.int32, device="cuda")
    total = p.count_voxel_views(counts)
    p.synchronize()
    ref, n_sel = oracle.reconstruct_native(sc, rc, raw, 0, d.n_slices_native)
    cnt, _ = oracle.bp_counts(sc, rc, 0, d.n_slices_native)
    assert total == n_sel and np.array_equal(to_np(counts), cnt)
    cov = (oracle.bp_coverage(sc, rc, 0, d.n_slices_native) >= 0.5) & fov_mask(rc, d.acq_fov)[None]
    g = to_np(nat).astype(np.float64)
    dd = (g - ref)[cov] / (S.MU_WATER / 1000.0)
    rmse, mx = float(np.sqrt(np.mean(dd * dd))), float(np.abs(dd).max())
    say("clockwise ffs %d: covered rmse %.4f max %.4f HU" % (ffs_mode, rmse, mx))
    assert cov.sum() > 0.2 * cov.size
    assert rmse <= 0.5 and mx <= 2.0
