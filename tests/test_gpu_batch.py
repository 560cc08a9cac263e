"""Batched uneven reconstructions (fct_wfbp_gpu_batch_*): one persistent
backprojection launch over every job's tiles (the grid kernel for a batch of
one job), recorded into a hipGraph.

Pin: each job's volume is bitwise identical to a single-plan reconstruction
with the same precision and tile depth (a tile is computed by the same code
whichever wave takes it), in eager mode, on the first graph replay and on
later replays with other GPU work in between, and after the autotuner has
changed the tile depth.
"""
import pytest

from freect_amd import scans as S
from helpers import gpu_ctx, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return gpu_ctx()


def _jobs():
    """Uneven jobs: scan length, pitch-like feed, detector width, matrix,
    slice count and filter all differ."""
    ws = [S.tiny(seed=11, n_channels=96, rotations=2.5, nx=32, n_slices=8, filter="ramp"),
          S.tiny(seed=12, n_channels=128, rotations=4.0, nx=48, n_slices=20,
                 filter="shepp-logan"),
          S.tiny(seed=13, n_channels=200, n_rows=12, ppr=128, rotations=3.0, nx=40,
                 n_slices=12, filter="hann")]
    # An off-centre central row exercises the general selection form.
    ws[2].scan = ws[2].scan.replace(central_row=ws[2].scan.central_row + 0.3)
    return ws


def _single(ctx, w, raw, precision, tile_z):
    from freect_amd import wfbp
    p = wfbp.Plan(ctx, w.scan, w.recon, precision=precision, tile_z=tile_z, bp_variant=2)
    vol = p.alloc_volume()
    p.reconstruct(raw, vol)
    p.synchronize()
    return vol


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_batch_equals_single_plans(ctx, precision):
    from freect_amd import wfbp
    ws = _jobs()
    raws = [ctx.phantom_project(w.scan, w.phantom, 0.0, w.seed) for w in ws]
    torch.cuda.synchronize()
    b = wfbp.Batch(ctx, [(w.scan, w.recon) for w in ws], precision=precision, tile_z=8)
    vols = [p.alloc_volume() for p in b.plans]
    for j, (r, v) in enumerate(zip(raws, vols)):
        b.submit(j, r, v)
    info = b.info()
    say("batch", precision, info)
    assert info["n_jobs"] == 3 and info["tile_z"] == 8 and info["grid"] % 8 == 0

    refs = [_single(ctx, w, r, precision, 8) for w, r in zip(ws, raws)]
    for mode in (wfbp.Batch.EAGER, wfbp.Batch.GRAPH, wfbp.Batch.GRAPH, wfbp.Batch.GRAPH):
        for v in vols:
            v.fill_(float("nan"))
        torch.cuda.synchronize()
        b.run(mode)
        b.synchronize()
        for j, (v, ref) in enumerate(zip(vols, refs)):
            say("  mode", mode, "job", j, "max|diff|", (v - ref).abs().max().item())
            assert torch.equal(v, ref)
        # Unrelated GPU work between replays.
        torch.randn(1 << 20, device="cuda").sum().item()

    tuned = b.autotune((4, 8, 16), reps=1)
    info = b.info()
    say("autotune", tuned, "->", info["tile_z"])
    assert set(tuned) == {4, 8, 16} and all(ms > 0 for ms in tuned.values())
    assert info["tile_z"] == min(tuned, key=tuned.get)
    refs = [_single(ctx, w, r, precision, info["tile_z"]) for w, r in zip(ws, raws)]
    for v in vols:
        v.zero_()
    b.run(wfbp.Batch.GRAPH)
    b.synchronize()
    for v, ref in zip(vols, refs):
        assert torch.equal(v, ref)
    b.close()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_single_job_batch_equals_single_plan(ctx, precision):
    # A batch of one job backprojects with the single-plan grid kernel: the
    # same bitwise volume, eagerly and over graph replays, before and after
    # the autotuner.
    from freect_amd import wfbp
    w = _jobs()[2]
    raw = ctx.phantom_project(w.scan, w.phantom, 0.0, w.seed)
    torch.cuda.synchronize()
    b = wfbp.Batch(ctx, [(w.scan, w.recon)], precision=precision, tile_z=8)
    vol = b.plans[0].alloc_volume()
    b.submit(0, raw, vol)
    ref = _single(ctx, w, raw, precision, 8)
    for mode in (wfbp.Batch.EAGER, wfbp.Batch.GRAPH, wfbp.Batch.GRAPH):
        vol.fill_(float("nan"))
        torch.cuda.synchronize()
        b.run(mode)
        b.synchronize()
        assert torch.equal(vol, ref)
        torch.randn(1 << 20, device="cuda").sum().item()
    tuned = b.autotune((4, 8), reps=1)
    tz = b.info()["tile_z"]
    say("single-job batch", precision, tuned, "->", tz)
    ref = _single(ctx, w, raw, precision, tz)
    vol.zero_()
    b.run(wfbp.Batch.GRAPH)
    b.synchronize()
    assert torch.equal(vol, ref)
    b.close()


def test_batch_errors(ctx):
    from freect_amd import wfbp
    w = S.tiny()
    b = wfbp.Batch(ctx, [(w.scan, w.recon)])
    with pytest.raises(wfbp.FctError):
        b.run()                      # nothing bound
    with pytest.raises(wfbp.FctError):
        b.autotune((5,))             # not a tile depth
    b.close()
