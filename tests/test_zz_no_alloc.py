"""fct_wfbp_gpu_reconstruct_noalloc: the whole chain without any device
allocation (the library's fixed per-GPU scratch array), called through the
C ABI.  Its volume is bitwise that of fct_wfbp_gpu_reconstruct (the same
kernels), on two calls in a row with the second on a smaller scan, in FP32
and FP16; a plan that does not fit the scratch is rejected with
FCT_ERR_LIMIT before anything runs."""
import ctypes as C

import pytest

from freect_amd import scans as S
from helpers import gpu_ctx, say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FCT_ERR_LIMIT = 5


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return gpu_ctx()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_noalloc_equals_reconstruct_twice(ctx, precision):
    from freect_amd import wfbp
    lib = wfbp.load_library()
    # The small scan (FP16: 226 MB of quads + 17 MB native), then a smaller one.
    for w in (S.small(), S.tiny(nx=48, n_slices=12, n_channels=192, n_rows=16, ppr=192,
                                rotations=3)):
        raw = ctx.phantom_project(w.scan, w.phantom, w.i0, w.seed)
        p = wfbp.Plan(ctx, w.scan, w.recon, precision=precision)
        ref = p.alloc_volume()
        p.reconstruct(raw, ref)
        got = p.alloc_volume()
        got.fill_(float("nan"))
        p.synchronize()
        st = lib.fct_wfbp_gpu_reconstruct_noalloc(p.h, raw.data_ptr(), got.data_ptr())
        assert st == 0, lib.fct_wfbp_gpu_last_error(ctx.h)
        p.synchronize()
        say(f"noalloc {precision} {w.name}: max|diff| {(got - ref).abs().max().item():.3g}")
        assert torch.equal(got, ref)


def test_noalloc_rejects_what_does_not_fit(ctx):
    from freect_amd import wfbp
    w = S.full()
    p = wfbp.Plan(ctx, w.scan, w.recon, precision="fp32")
    lib = wfbp.load_library()
    st = lib.fct_wfbp_gpu_reconstruct_noalloc(p.h, C.c_void_p(16), C.c_void_p(16))
    assert st == FCT_ERR_LIMIT
    assert b"scratch" in lib.fct_wfbp_gpu_last_error(ctx.h)
    # Adaptive filtration needs a raw-sized buffer: not on this path.
    t = S.tiny()
    p = wfbp.Plan(ctx, t.scan, t.recon.replace(adaptive_filtration_s=0.2),
                  adaptive_filtration=True)
    assert lib.fct_wfbp_gpu_reconstruct_noalloc(p.h, C.c_void_p(16), C.c_void_p(16)) == 2
