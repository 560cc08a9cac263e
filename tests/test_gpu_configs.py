"""GPU parity at the full sizes of the benchmark workloads.

* full: GE-like 888 x 64, 9840 views (984/rot, pitch 0.984), 512^2 x 320
  native slices at 0.625 mm, ramp, Poisson I0 = 5e4.  FP32 and FP16 storage
  against the FP64 restatement on native slices at both ends and in the
  middle of the volume; per-voxel selection counts bit-exact there.
* ffs: Siemens-like 736 x 32, 18432 projections with phi+z flying focal spot
  (4608 parallel views of 1472 x 64), 512^2 x 200 slices of 1.0 mm over
  [-100, 100) mm; the central native slice, one near the start of the helix
  (partly covered) and the first slice (no view reaches it).
* batch: 24 seeded uneven jobs on the 736 x 16 detector of the small scan;
  the batched volumes bitwise equal to one-by-one plans, and two jobs against
  the FP64 restatement.

Tolerances (FP64 restatement as reference, HU = 1000 mu / mu_water):
FP32 <= 0.5 HU RMSE and <= 2 HU max on fully covered voxels (every theta~
has weight sum >= 0.5), FP16 storage <= 1 HU RMSE; the CPU side only
reconstructs the sampled slices, so each case runs in seconds.
"""
import numpy as np
import pytest

from freect_amd import scans as S
from helpers import fov_mask, gpu_ctx, say, to_np

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HU = S.MU_WATER / 1000.0


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return gpu_ctx()


def _native(ctx, w, raw_t, precision, **kw):
    from freect_amd import wfbp
    p = wfbp.Plan(ctx, w.scan, w.recon, precision=precision, **kw)
    filt, scale = p.alloc_filtered()
    nat = p.alloc_native()
    p.rebin_filter(raw_t, filt, scale)
    p.backproject(filt, scale, nat)
    p.synchronize()
    del filt, scale
    return p, nat


def _compare(tag, got, ref, mask):
    d = (np.asarray(got, np.float64) - ref)[mask] / HU
    rmse, mx = float(np.sqrt(np.mean(d * d))), float(np.abs(d).max())
    say("%s: %d voxels, rmse %.4f max %.4f HU" % (tag, mask.sum(), rmse, mx))
    return rmse, mx


def test_full_workload_matches_oracle(ctx, oracle):
    w = S.full()
    sc, rc = w.scan, w.recon
    d = oracle.geometry(sc, rc)
    n = d.n_slices_native
    say("full workload: views", sc.n_views, "native slices", n)
    raw_t = ctx.phantom_project(sc, w.phantom, w.i0, w.seed)
    torch.cuda.synchronize()
    raw = to_np(raw_t)
    ranges = [(0, 4), (n // 2 - 4, n // 2 + 4), (n - 4, n)]
    fov = fov_mask(rc, d.acq_fov)
    got = {}
    for prec in ("fp32", "fp16"):
        p, nat = _native(ctx, w, raw_t, prec)
        got[prec] = {r: to_np(nat[r[0]:r[1]]) for r in ranges}
        if prec == "fp32":
            counts = torch.zeros((n, rc.ny, rc.nx), dtype=torch.int32, device="cuda")
            total = p.count_voxel_views(counts)
            got["counts"] = {r: to_np(counts[r[0]:r[1]]) for r in ranges}
            say("full workload: voxel-views", total)
            del counts
        del nat, p
        torch.cuda.empty_cache()
    for (s0, s1) in ranges:
        ref, n_sel = oracle.reconstruct_native(sc, rc, raw, s0, s1)
        cnt, n_ref = oracle.bp_counts(sc, rc, s0, s1)
        assert np.array_equal(got["counts"][(s0, s1)], cnt)
        assert int(got["counts"][(s0, s1)].sum(dtype=np.int64)) == n_ref == n_sel
        cov = (oracle.bp_coverage(sc, rc, s0, s1) >= 0.5) & fov[None]
        assert cov.sum() > 0.9 * fov.sum() * (s1 - s0)   # inside the helix
        r32, m32 = _compare(f"full fp32 slices {s0}..{s1 - 1}", got["fp32"][(s0, s1)], ref, cov)
        r16, m16 = _compare(f"full fp16 slices {s0}..{s1 - 1}", got["fp16"][(s0, s1)], ref, cov)
        assert r32 <= 0.5 and m32 <= 2.0
        assert r16 <= 1.0


def test_ffs_workload_matches_oracle(ctx, oracle):
    w = S.ffs()
    sc, rc = w.scan, w.recon
    d = oracle.geometry(sc, rc)
    n = d.n_slices_native
    zs = oracle.slice_grid(rc.start_pos, rc.end_pos, d.csw)
    z_first = min(sc.z_of_view(0), sc.z_of_view(sc.n_views - 1))
    s_edge = int(np.searchsorted(zs, z_first + 0.4 * d.zrot))
    picks = [n // 2, s_edge, 0]
    say("ffs workload: projections", sc.n_views, "native slices", n, "picks", picks)
    raw_t = ctx.phantom_project(sc, w.phantom, w.i0, w.seed)
    torch.cuda.synchronize()
    raw = to_np(raw_t)
    fov = fov_mask(rc, d.acq_fov)
    got = {}
    for prec in ("fp32", "fp16"):
        p, nat = _native(ctx, w, raw_t, prec)
        got[prec] = {s: to_np(nat[s]) for s in picks}
        del nat, p
        torch.cuda.empty_cache()
    for s in picks:
        ref, _ = oracle.reconstruct_native(sc, rc, raw, s, s + 1)
        cov = (oracle.bp_coverage(sc, rc, s, s + 1)[0] >= 0.5) & fov
        if s == 0:
            # z = -100 mm lies past the end of the helix: no view reaches it.
            assert not ref.any() and not got["fp32"][s].any() and not got["fp16"][s].any()
            continue
        assert cov.sum() > 0.1 * fov.sum()
        r32, m32 = _compare(f"ffs fp32 slice {s}", got["fp32"][s], ref[0], cov)
        r16, m16 = _compare(f"ffs fp16 slice {s}", got["fp16"][s], ref[0], cov)
        assert r32 <= 0.5 and m32 <= 2.0
        assert r16 <= 1.0


def test_batch_workload_24_jobs(ctx, oracle):
    from freect_amd import wfbp
    ws = S.batch_jobs()
    assert len(ws) == 24 and all(w.scan.n_channels == 736 and w.scan.n_rows == 16 for w in ws)
    raws = [ctx.phantom_project(w.scan, w.phantom, w.i0, w.seed) for w in ws]
    torch.cuda.synchronize()
    b = wfbp.Batch(ctx, [(w.scan, w.recon) for w in ws], precision="fp16", tile_z=8)
    vols = [p.alloc_volume() for p in b.plans]
    for j, (r, v) in enumerate(zip(raws, vols)):
        b.submit(j, r, v)
    say("batch workload:", b.info())
    for v in vols:
        v.fill_(float("nan"))
    b.run(wfbp.Batch.GRAPH)
    b.synchronize()
    for j, (w, r, v) in enumerate(zip(ws, raws, vols)):
        p = wfbp.Plan(ctx, w.scan, w.recon, precision="fp16", tile_z=8, bp_variant=2)
        ref = p.alloc_volume()
        p.reconstruct(r, ref)
        p.synchronize()
        assert torch.equal(v, ref), f"job {j}"
        del p, ref
    say("batch workload: 24 batched volumes bitwise equal to single plans")
    b.close()
    del vols
    # Two jobs (the shortest and the longest scan) against the FP64
    # restatement, FP32 storage, on their central native slices.
    order = sorted(range(len(ws)), key=lambda j: ws[j].scan.n_views)
    for j in (order[0], order[-1]):
        w = ws[j]
        sc, rc = w.scan, w.recon
        d = oracle.geometry(sc, rc)
        s0 = d.n_slices_native // 2 - 1
        p, nat = _native(ctx, w, raws[j], "fp32")
        ref, _ = oracle.reconstruct_native(sc, rc, to_np(raws[j]), s0, s0 + 2)
        cov = (oracle.bp_coverage(sc, rc, s0, s0 + 2) >= 0.5) & fov_mask(rc, d.acq_fov)[None]
        assert cov.sum() > 0
        rmse, mx = _compare(f"batch job {j} ({sc.n_views} views, {rc.nx}^2, {rc.filter}) fp32",
                            to_np(nat[s0:s0 + 2]), ref, cov)
        assert rmse <= 0.5 and mx <= 2.0
        del p, nat
