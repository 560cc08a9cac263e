"""Streaming whole chain (fct_wfbp_gpu_reconstruct_stream, and the host path
of fct_wfbp_gpu_reconstruct): views arrive in chunks, are copied on the
plan's copy stream, and rebin+filter / backprojection of what has arrived
run while later chunks are copied.

Pins: the streamed volume is bitwise the device-resident one (every line
pair and tile is computed by the same code) for FP32, for FP16 with a given
bound on max|raw|, for a flying-focal-spot scan, an odd row count and a
table moving backwards and a scan much longer than the volume (whose
leading and trailing views, read by no slice, are not requested); FP16 with
the running bound (the stored projections
rescaled when a later chunk raises the scale) agrees with the one-shot FP16
volume to the rounding of values below the format's normal range; the
small workload streamed matches the FP64 restatement (<= 0.5 HU RMSE,
<= 2 HU max on covered voxels).
"""
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


def _device_volume(ctx, w, raw_t, precision):
    from freect_amd import wfbp
    p = wfbp.Plan(ctx, w.scan, w.recon, precision=precision)
    vol = p.alloc_volume()
    p.reconstruct(raw_t, vol)
    p.synchronize()
    return to_np(vol)


def _check_calls(calls, n_views):
    # One contiguous, increasing run of views inside the scan.
    assert calls and calls[0][0] >= 0 and calls[-1][0] + calls[-1][1] <= n_views
    for (a, n), (b, _) in zip(calls, calls[1:]):
        assert n > 0 and b == a + n


def _reader(raw):
    calls = []

    def read(v0, n, dst):
        calls.append((v0, n))
        dst[...] = raw[v0:v0 + n]
    return read, calls


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_small_workload_streamed_matches_device_and_oracle(ctx, oracle, precision):
    from freect_amd import wfbp
    w = S.small()
    sc, rc = w.scan, w.recon
    raw_t = ctx.phantom_project(sc, w.phantom, 0.0, w.seed)
    torch.cuda.synchronize()
    raw = to_np(raw_t)
    ref = _device_volume(ctx, w, raw_t, precision)
    p = wfbp.Plan(ctx, sc, rc, precision=precision)
    bound = float(np.abs(raw).max()) if precision == "fp16" else 0.0
    for vpc in (0, 100, 257):
        read, calls = _reader(raw)
        got = p.reconstruct_stream(read, views_per_chunk=vpc, max_abs_raw=bound)
        n_chunks = len(calls)
        _check_calls(calls, sc.n_views)
        say(f"stream {precision} vpc {vpc}: {n_chunks} chunks, max|diff| "
            f"{np.abs(got - ref).max():.3g}")
        assert np.array_equal(got, ref)
    host = p.reconstruct_host(raw)              # chunked copies from the caller's array
    assert np.array_equal(host, ref)
    pinned = torch.from_numpy(raw).pin_memory()
    host = p.reconstruct_host(pinned)           # pinned: asynchronous DMA
    assert np.array_equal(host, ref)
    # Pinned source and destination: every upload queued at once, the volume
    # copied back in slice ranges as the slab groups finish.
    out = torch.full(ref.shape, float("nan")).pin_memory()
    host = p.reconstruct_host(pinned, out=out.numpy())
    assert np.array_equal(host, ref)
    if precision == "fp32":
        d = oracle.geometry(sc, rc)
        ref_nat, _ = oracle.reconstruct_native(sc, rc, raw, 0, d.n_slices_native)
        ref64 = oracle.thicken(sc, rc, ref_nat)
        cov = oracle.covered_mask(sc, rc) & fov_mask(rc, d.acq_fov)[None]
        dd = (host.astype(np.float64) - ref64)[cov] / (S.MU_WATER / 1000.0)
        rmse, mx = float(np.sqrt(np.mean(dd * dd))), float(np.abs(dd).max())
        say("streamed small vs fp64: rmse %.4f max %.4f HU" % (rmse, mx))
        assert rmse <= 0.5 and mx <= 2.0


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_running_bound_rescales_stored_projections(ctx, precision):
    # Later views 4x brighter: the running abs-max raises the storage scale
    # twice while the scan streams in, and the stored quads are rescaled.
    w = S.tiny(nx=48, n_slices=12, n_channels=192, n_rows=16, ppr=192, rotations=3)
    raw_t = ctx.phantom_project(w.scan, w.phantom, 5e4, w.seed)
    ramp = torch.linspace(1.0, 4.0, w.scan.n_views, device="cuda")[:, None, None]
    raw_t = (raw_t * ramp).contiguous()
    torch.cuda.synchronize()
    raw = to_np(raw_t)
    ref = _device_volume(ctx, w, raw_t, precision)
    from freect_amd import wfbp
    p = wfbp.Plan(ctx, w.scan, w.recon, precision=precision)
    read, calls = _reader(raw)
    got = p.reconstruct_stream(read, views_per_chunk=48)
    diff = np.abs(got.astype(np.float64) - ref)
    say(f"{precision} running bound: {len(calls)} chunks, max|diff| {diff.max():.3g} "
        f"({int((diff > 0).sum())} voxels differ) of max|vol| {np.abs(ref).max():.3g}")
    assert diff.max() <= 1e-5 * np.abs(ref).max()


def test_running_bound_growth_with_groups_in_flight(ctx):
    # 64 slices in slabs of 4 over a 10-rotation helix whose views brighten
    # 8x along the scan: slab groups start while the scan still streams in,
    # and the storage scale grows three times after the first group, so each
    # growth waits for the groups in flight before the stored quads are
    # rescaled.  Pinned source and destination (every upload queued at once,
    # slice ranges copied back as groups finish) and a view reader.
    from freect_amd import wfbp
    w = S.tiny(nx=32, n_slices=64, n_channels=96, n_rows=8, ppr=96, rotations=10)
    raw_t = ctx.phantom_project(w.scan, w.phantom, 5e4, w.seed)
    ramp = torch.logspace(0.0, 3.0, w.scan.n_views, base=2.0, device="cuda")[:, None, None]
    raw_t = (raw_t * ramp).contiguous()
    torch.cuda.synchronize()
    raw = to_np(raw_t)
    p_dev = wfbp.Plan(ctx, w.scan, w.recon, precision="fp16", tile_z=4)
    vol = p_dev.alloc_volume()
    p_dev.reconstruct(raw_t, vol)
    p_dev.synchronize()
    ref = to_np(vol)
    p = wfbp.Plan(ctx, w.scan, w.recon, precision="fp16", tile_z=4)
    assert p.info.n_slices_native >= 64
    pinned = torch.from_numpy(raw).pin_memory()
    out = torch.empty(ref.shape).pin_memory().numpy()
    got = p.reconstruct_host(pinned, out=out)
    read, calls = _reader(raw)
    got2 = p.reconstruct_stream(read, views_per_chunk=32)
    for name, g in (("pinned", got), ("reader", got2)):
        diff = np.abs(g.astype(np.float64) - ref)
        say(f"growth in flight ({name}): max|diff| {diff.max():.3g} of max|vol| "
            f"{np.abs(ref).max():.3g}")
        assert diff.max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("case", ["ffs", "odd_rows", "backwards", "long_scan"])
def test_streamed_equals_device_special_scans(ctx, case):
    from freect_amd import wfbp
    if case == "long_scan":   # 6 rotations (+-120 mm) around a 40 mm volume
        w = S.tiny(nx=40, n_slices=8, n_channels=128, n_rows=8, ppr=128, rotations=6)
    elif case == "ffs":
        w = S.tiny(nx=48, n_slices=12, n_channels=96, n_rows=8, ppr=192, rotations=3,
                   ffs_mode=3)
    elif case == "odd_rows":
        w = S.tiny(nx=40, n_slices=10, n_channels=150, n_rows=5, ppr=96, rotations=3)
    else:
        w = S.tiny(nx=40, n_slices=12, n_channels=128, n_rows=8, ppr=128, rotations=3)
        w.scan = w.scan.replace(feed=-w.scan.feed, z0=-w.scan.z0)
    raw_t = ctx.phantom_project(w.scan, w.phantom, 0.0, w.seed)
    torch.cuda.synchronize()
    raw = to_np(raw_t)
    for precision in ("fp32", "fp16"):
        ref = _device_volume(ctx, w, raw_t, precision)
        p = wfbp.Plan(ctx, w.scan, w.recon, precision=precision)
        bound = float(np.abs(raw).max()) if precision == "fp16" else 0.0
        read, calls = _reader(raw)
        got = p.reconstruct_stream(read, views_per_chunk=37, max_abs_raw=bound)
        say(f"{case} {precision}: {len(calls)} chunks, views [{calls[0][0]}, "
            f"{calls[-1][0] + calls[-1][1]}) of {w.scan.n_views}, "
            f"max|diff| {np.abs(got - ref).max():.3g}")
        _check_calls(calls, w.scan.n_views)
        # Exactly the views some slice reads are requested.
        assert (calls[0][0], calls[-1][0] + calls[-1][1]) == \
            (p.info.read_views_first, p.info.read_views_end)
        if case == "long_scan":   # the views no slice reads are not requested
            assert calls[0][0] > 0 and calls[-1][0] + calls[-1][1] < w.scan.n_views
        assert np.array_equal(got, ref)
    if case == "long_scan":   # FP16 with the running bound of the requested views only
        ref = _device_volume(ctx, w, raw_t, "fp16")
        p = wfbp.Plan(ctx, w.scan, w.recon, precision="fp16")
        for pin in (False, True):
            src = torch.from_numpy(raw).pin_memory() if pin else raw
            out = torch.empty(ref.shape).pin_memory().numpy() if pin else None
            got = p.reconstruct_host(src, out=out)
            diff = np.abs(got.astype(np.float64) - ref)
            say(f"long_scan fp16 running bound (pinned {pin}): max|diff| {diff.max():.3g}")
            assert diff.max() <= 1e-5 * np.abs(ref).max()


def test_reader_failure_is_an_error_and_plan_stays_usable(ctx):
    from freect_amd import wfbp
    w = S.tiny()
    raw_t = ctx.phantom_project(w.scan, w.phantom, 0.0, w.seed)
    torch.cuda.synchronize()
    raw = to_np(raw_t)
    p = wfbp.Plan(ctx, w.scan, w.recon)

    def bad(v0, n, dst):
        if v0 >= 64:
            raise IOError("frame unreadable")
        dst[...] = raw[v0:v0 + n]
    with pytest.raises(IOError):
        p.reconstruct_stream(bad, views_per_chunk=32)
    ref = _device_volume(ctx, w, raw_t, "fp32")
    read, _ = _reader(raw)
    assert np.array_equal(p.reconstruct_stream(read, views_per_chunk=32), ref)
    with pytest.raises(wfbp.FctError):
        p.reconstruct_stream(read, views_per_chunk=32, max_abs_raw=float("nan"))
