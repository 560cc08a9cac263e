"""The step before and after the path: adaptive filtration of the raw scan
(the recon key adaptive_filtration_s, parsed but never applied by the
reference, parse_config.cpp:57, README.md:14) and Hounsfield output (the
reference writes 1/mm only, ReconstructionMachine.cu:436-455).

Pins: the whole chain with adaptive filtration against the FP64 restatement
fed the restatement's own adaptively filtered scan (<= 0.5 HU RMSE, <= 2 HU
max on covered voxels), the streamed chain bitwise equal to the device one,
and HU output equal to 1000 (mu - mu_water) / mu_water of the 1/mm volume.
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


def test_adaptive_filtration_matches_oracle(ctx, oracle):
    from freect_amd import wfbp
    w = S.tiny(nx=48, n_slices=12, n_channels=192, n_rows=16, ppr=192, rotations=3)
    rc = w.recon.replace(adaptive_filtration_s=0.3)
    sc = w.scan
    raw_t = ctx.phantom_project(sc, w.phantom, 2e4, w.seed)
    torch.cuda.synchronize()
    raw = to_np(raw_t)
    p = wfbp.Plan(ctx, sc, rc, adaptive_filtration=True)
    vol = p.alloc_volume()
    p.reconstruct(raw_t, vol)
    p.synchronize()
    got = to_np(vol)
    plain = wfbp.Plan(ctx, sc, rc)
    vol0 = plain.alloc_volume()
    plain.reconstruct(raw_t, vol0)
    plain.synchronize()
    d = oracle.geometry(sc, rc)
    raw_af = oracle.adaptive_filter(sc, raw, 0.3)
    nat, _ = oracle.reconstruct_native(sc, rc, raw_af, 0, d.n_slices_native)
    ref = oracle.thicken(sc, rc, nat)
    cov = oracle.covered_mask(sc, rc) & fov_mask(rc, d.acq_fov)[None]
    dd = (got.astype(np.float64) - ref)[cov] / (S.MU_WATER / 1000.0)
    rmse, mx = float(np.sqrt(np.mean(dd * dd))), float(np.abs(dd).max())
    change = float(np.abs(got - to_np(vol0)).max() / (S.MU_WATER / 1000.0))
    say("adaptive filtration vs fp64: rmse %.4f max %.4f HU (change vs unfiltered %.1f HU)"
        % (rmse, mx, change))
    assert change > 1.0            # it does something on this noisy scan
    assert rmse <= 0.5 and mx <= 2.0
    # Streamed (filtered chunk by chunk as views arrive): the same volume.
    def read(v0, n, dst):
        dst[...] = raw[v0:v0 + n]
    assert np.array_equal(p.reconstruct_stream(read, views_per_chunk=50), got)
    with pytest.raises(wfbp.FctError, match="adaptive_filtration_s"):
        wfbp.Plan(ctx, sc, w.recon, adaptive_filtration=True)   # s = 0


def test_hounsfield_output(ctx):
    from freect_amd import wfbp
    w = S.tiny(nx=48, n_slices=12, n_channels=192, n_rows=16, ppr=192, rotations=3)
    raw_t = ctx.phantom_project(w.scan, w.phantom, 0.0, w.seed)
    p = wfbp.Plan(ctx, w.scan, w.recon)
    mu = p.alloc_volume()
    p.reconstruct(raw_t, mu)
    p.set_output_units("hu", S.MU_WATER)
    hu = p.alloc_volume()
    p.reconstruct(raw_t, hu)
    p.synchronize()
    mu, hu = to_np(mu).astype(np.float64), to_np(hu).astype(np.float64)
    ref = 1000.0 * (mu - S.MU_WATER) / S.MU_WATER
    err = np.abs(hu - ref).max()
    say("HU output: max |HU - 1000 (mu - mu_w) / mu_w| = %.2e" % err)
    assert err < 2e-3
    p.set_output_units("mu")
    again = p.alloc_volume()
    p.reconstruct(raw_t, again)
    p.synchronize()
    assert np.array_equal(to_np(again).astype(np.float64), mu)
    with pytest.raises(wfbp.FctError):
        p.set_output_units("hu", 0.0)
