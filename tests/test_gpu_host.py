"""The fct_wfbp command-line driver (C++ mirror of FreeCT's main.cpp +
ReconstructionMachine) on the GPU: a recon YAML in, a raw float32
[slice][y][x] volume out, equal to the Python/C-ABI path."""
import os
import subprocess

import numpy as np
import pytest

from freect_amd import scans as S
from helpers import gpu_ctx, say, to_np

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "freect_amd", "lib", "fct_wfbp")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return gpu_ctx()


def _python_volume(ctx, w):
    from freect_amd import wfbp
    raw = ctx.phantom_project(w.scan, w.phantom, w.i0, w.seed)
    p = wfbp.Plan(ctx, w.scan, w.recon)
    vol = p.alloc_volume()
    p.reconstruct(raw, vol)
    p.synchronize()
    return to_np(vol)


def test_cli_synthetic_matches_capi(ctx, tmp_path):
    w = S.small()
    cfg = tmp_path / "recon.yaml"
    rc = w.recon
    cfg.write_text(f"raw_data_dir: synthetic:small\noutput_dir: {tmp_path}\n"
                   f"start_pos: {rc.start_pos}\nend_pos: {rc.end_pos}\nrecon_fov: {rc.recon_fov}\n"
                   f"slice_thickness: {rc.slice_thickness}\nslice_pitch: {rc.slice_pitch}\n"
                   f"nx: {rc.nx}\nny: {rc.ny}\nrecon_kernel: 3\nx_origin: 0\ny_origin: 0\n"
                   f"filter: shepp-logan\n")
    out = tmp_path / "vol.dat"
    say("cli synthetic small")
    r = subprocess.run([CLI, "-v", "-o", str(out), str(cfg)], capture_output=True, text=True,
                       timeout=300)
    say(r.stdout[-800:])
    assert r.returncode == 0
    got = np.fromfile(out, np.float32).reshape(32, rc.ny, rc.nx)
    ref = _python_volume(ctx, w)
    assert np.array_equal(got, ref)


def test_cli_freect_format_round_trip(ctx, tmp_path):
    say("cli write-synthetic small")
    r = subprocess.run([CLI, "--write-synthetic", "small", str(tmp_path)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert (tmp_path / "meta.yaml").exists() and (tmp_path / "projections.dat").exists()
    out = tmp_path / "vol.dat"
    r = subprocess.run([CLI, "-o", str(out), str(tmp_path / "recon.yaml")], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    w = S.small()
    got = np.fromfile(out, np.float32).reshape(32, w.recon.ny, w.recon.nx)
    ref = _python_volume(ctx, w)
    # The two-file format stores per-view angle and z as float32, so the scan
    # law is re-derived from rounded values: agreement to a small fraction of 1 HU.
    d = np.abs(got.astype(np.float64) - ref) / (S.MU_WATER / 1000.0)
    say("freect format round trip: max %.4f HU, mean %.5f HU" % (d.max(), d.mean()))
    assert d.mean() < 0.05 and np.percentile(d, 99.9) < 1.0
