"""The hand-written adjoint (matched helical forward projector) on the GPU.

Checks, all through the C ABI:
  * dot-product identity <A p, x> = <p, A^T x> per stage and for the whole
    chain over 16 seeded Gaussian pairs (FP64 inner products), <= 1e-5 rel;
  * the GPU adjoint against the FP64 oracle adjoint;
  * FP64 central finite differences of f(p) = <A_64 p, x> (oracle forward)
    along 64 random directions against <d, A^T x> from the GPU;
  * the torch.autograd wrappers' gradients equal the adjoint.
"""
import numpy as np
import pytest

from freect_amd import scans as S
from helpers import gpu_ctx, say, to_np

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return gpu_ctx()


def _tiny():
    return S.tiny(nx=32, n_slices=8, n_channels=96, n_rows=8, ppr=96, rotations=2.5)


def _plan(ctx, w, **kw):
    from freect_amd import wfbp
    return wfbp.Plan(ctx, w.scan, w.recon, **kw)


def _dot(a, b):
    return float(torch.dot(a.flatten().double(), b.flatten().double()))


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_stage_dot_products(ctx):
    w = _tiny()
    p = _plan(ctx, w)
    gen = torch.Generator(device="cuda").manual_seed(5)
    sc = w.scan
    nch, cp = sc.n_channels, p.info.channel_pitch
    for trial in range(4):
        raw = torch.randn((sc.n_views, sc.n_rows, nch), device="cuda", generator=gen)
        lines = torch.zeros((sc.n_views, sc.n_rows, cp), device="cuda")
        lines[..., :nch] = torch.randn((sc.n_views, sc.n_rows, nch), device="cuda", generator=gen)
        nat = torch.randn((p.info.n_slices_native, w.recon.ny, w.recon.nx), device="cuda",
                          generator=gen)
        # filter (self-adjoint)
        f1, f2 = p.alloc_lines(), p.alloc_lines()
        lines2 = torch.zeros_like(lines)
        lines2[..., :nch] = torch.randn((sc.n_views, sc.n_rows, nch), device="cuda", generator=gen)
        p.filter_lines(lines, f1)
        p.filter_lines(lines2, f2)
        torch.cuda.synchronize()
        a, b = _dot(f1[..., :nch], lines2[..., :nch]), _dot(lines[..., :nch], f2[..., :nch])
        assert _rel(a, b) < 1e-5, (a, b)
        # backprojection
        bp = p.alloc_native()
        p.backproject(lines, None, bp)
        gt = p.alloc_lines()
        p.backproject_adjoint(nat, gt)
        torch.cuda.synchronize()
        a, b = _dot(bp, nat), _dot(lines[..., :nch], gt[..., :nch])
        say("bp dot", a, b, _rel(a, b))
        assert _rel(a, b) < 1e-5
        # rebin: forward via rebin_filter on the identity filter is not
        # exposed; use the whole-chain identity below and the oracle checks.
        rb = p.alloc_raw()
        p.rebin_adjoint(lines, rb)
        torch.cuda.synchronize()
        assert torch.isfinite(rb).all()


def _chain_dot_test(ctx, w, seed, tag):
    """<A p, x> against <p, A^T x> over 16 seeded Gaussian pairs, FP64 inner
    products of the FP32 results.  The pairs are Gaussian with mean 1 and
    unit variance: with zero-mean pairs <A p, x> is itself a zero-mean sum,
    the smallest of 16 falls near zero and |a - b| / |a| measures FP32
    round-off divided by that accident rather than how exactly A^T transposes
    A.  The zero-mean pairs' error against the Cauchy-Schwarz scale
    ||A p|| ||x|| is reported beside."""
    p = _plan(ctx, w)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    worst, worst_cs = 0.0, 0.0
    shp_raw = (w.scan.n_views, w.scan.n_rows, w.scan.n_channels)
    shp_x = (p.info.n_slices_native, w.recon.ny, w.recon.nx)
    for trial in range(16):
        for mean in (1.0, 0.0):
            raw = torch.randn(shp_raw, device="cuda", generator=gen) + mean
            x = torch.randn(shp_x, device="cuda", generator=gen) + mean
            ax = p.alloc_native()
            p.forward_native(raw, ax)
            atx = p.alloc_raw()
            p.adjoint(x, atx)
            torch.cuda.synchronize()
            a, b = _dot(ax, x), _dot(raw, atx)
            if mean:
                worst = max(worst, _rel(a, b))
            else:
                worst_cs = max(worst_cs, abs(a - b) / (float(ax.double().norm()) *
                                                       float(x.double().norm())))
    say("%s dot test: worst rel %.3g over 16 pairs (zero-mean pairs: %.3g of ||Ap|| ||x||)"
        % (tag, worst, worst_cs))
    assert worst < 1e-5
    assert worst_cs < 1e-7


def test_whole_chain_dot_products_16_pairs(ctx):
    _chain_dot_test(ctx, _tiny(), 2024, "whole-chain")


def test_adjoint_workload_dot_products(ctx):
    """The adjoint workload: 736 x 16 detector at 984 views/rot, 3 rotations
    (2952 views), 512^2 x 64 native slices."""
    _chain_dot_test(ctx, S.adjoint(), 77, "adjoint workload")


def test_adjoint_matches_oracle(ctx, oracle):
    w = _tiny()
    sc, rc = w.scan, w.recon
    p = _plan(ctx, w)
    d = oracle.geometry(sc, rc)
    rng = np.random.default_rng(9)
    x = rng.standard_normal((d.n_slices_native, rc.ny, rc.nx)).astype(np.float32)
    atx = p.alloc_raw()
    p.adjoint(torch.from_numpy(x).cuda(), atx)
    torch.cuda.synchronize()
    ref = oracle.adjoint_native(sc, rc, x.astype(np.float64), 0)
    got = to_np(atx).astype(np.float64)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say("adjoint vs FP64 oracle adjoint: max rel %.3g" % err)
    assert err < 1e-4
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-5


def test_adjoint_fp64_finite_differences(ctx, oracle):
    # f(p) = <A_64 p, x>; its directional derivative along d is <d, A^T x>.
    w = S.tiny(nx=16, n_slices=4, n_channels=48, n_rows=4, ppr=48, rotations=2.5)
    sc, rc = w.scan, w.recon
    p = _plan(ctx, w)
    d = oracle.geometry(sc, rc)
    rng = np.random.default_rng(31)
    x = rng.standard_normal((d.n_slices_native, rc.ny, rc.nx))
    atx = p.alloc_raw()
    p.adjoint(torch.from_numpy(x.astype(np.float32)).cuda(), atx)
    torch.cuda.synchronize()
    g = to_np(atx).astype(np.float64)
    x32 = x.astype(np.float32).astype(np.float64)
    p0 = rng.standard_normal((sc.n_views, sc.n_rows, sc.n_channels)).astype(np.float32)
    eps = 0.5   # A is linear: central differences are exact up to FP64 rounding
    worst = 0.0
    for k in range(64):
        dirn = rng.standard_normal(p0.shape).astype(np.float32)
        fp, _ = oracle.reconstruct_native(sc, rc, p0 + eps * dirn, 0, d.n_slices_native)
        fm, _ = oracle.reconstruct_native(sc, rc, p0 - eps * dirn, 0, d.n_slices_native)
        fd = (np.vdot(fp, x32) - np.vdot(fm, x32)) / (2 * eps)
        an = np.vdot(dirn.astype(np.float64), g)
        worst = max(worst, abs(fd - an) / max(abs(fd), 1e-12))
    say("FP64 central differences vs GPU adjoint: worst rel %.3g over 64 directions" % worst)
    # The difference is <(A_32 - A_64) d, x>: the FP32 GPU operator against the
    # FP64 one, about 1e-5 relative on images and larger on the random-sign
    # inner products here.  A wrong transpose is off by O(1).
    assert worst < 3e-4


def test_autograd_wrappers(ctx):
    from freect_amd import autograd as AG
    w = _tiny()
    p = _plan(ctx, w)
    gen = torch.Generator(device="cuda").manual_seed(3)
    raw = torch.randn((w.scan.n_views, w.scan.n_rows, w.scan.n_channels), device="cuda",
                      generator=gen, requires_grad=True)
    wgt = torch.randn((p.info.n_slices_native, w.recon.ny, w.recon.nx), device="cuda",
                      generator=gen)
    (AG.wfbp(raw, p) * wgt).sum().backward()
    atx = p.alloc_raw()
    p.adjoint(wgt, atx)
    torch.cuda.synchronize()
    assert torch.allclose(raw.grad, atx, rtol=1e-5, atol=1e-6 * atx.abs().max().item())
    x = wgt.clone().requires_grad_(True)
    r = torch.randn_like(raw)
    (AG.project(x, p) * r).sum().backward()
    ar = p.alloc_native()
    p.forward_native(r.detach(), ar)
    torch.cuda.synchronize()
    assert torch.allclose(x.grad, ar, rtol=1e-5, atol=1e-6 * ar.abs().max().item())


def test_matched_projector_dot_products_and_refinement(ctx):
    """P0 = R^T M^T (fct_wfbp_gpu_matched_project, rebin_adjoint) and
    P0^T = M R (fct_wfbp_gpu_rebin, matched_backproject): 16 seeded Gaussian
    pairs agree to 1e-5; P0 of a reconstructed water cylinder follows its
    line integrals on the rays the volume explains (the least-squares scale
    leaves a misfit from the voxel-driven footprint, reported); and the
    refinement's loss decreases monotonically with the 1/(k^2 lambda) step."""
    from freect_amd import autograd as AG
    w = _tiny()
    p = _plan(ctx, w)
    gen = torch.Generator(device="cuda").manual_seed(11)
    worst = 0.0
    for _ in range(16):
        raw = torch.randn((w.scan.n_views, w.scan.n_rows, w.scan.n_channels), device="cuda",
                          generator=gen)
        x = torch.randn((p.info.n_slices_native, w.recon.ny, w.recon.nx), device="cuda",
                        generator=gen)
        a = _dot(AG._p0(p, x), raw)
        b = _dot(x, AG._p0t(p, raw))
        worst = max(worst, _rel(a, b))
    say("matched projector dot test: worst rel %.3g" % worst)
    assert worst < 1e-5
    # Calibration: a centred water cylinder reconstructed and re-projected,
    # on the rays whose path through the FOV stays inside the volume's z-range.
    ph = S.Phantom(cyl_radius=100.0, ellipsoids=[])
    raw = ctx.phantom_project(w.scan, ph, 0.0, 0)
    x0 = AG.wfbp(raw, p).detach()
    m = AG.inside_rays(p)
    assert m.sum() > 0.05 * m.numel()
    px = AG._p0(p, x0) * m
    ref = raw * m
    k = _dot(px, ref) / _dot(px, px)
    rel = float((k * px - ref).norm() / ref.norm())
    say("matched projector: scale %.5g, relative misfit %.4f on %d rays" % (k, rel, int(m.sum())))
    # Coarse here (12.5 mm voxels against 5 mm rows and ~5 mm channels): the
    # voxel-driven footprint leaves a misfit of ~0.25; the transpose of the
    # W-normalised backprojection (no R / l, view window) measured 0.49.
    assert rel < 0.35
    x, st = AG.refine(p, raw, x0, steps=4, operator="matched", mask=m)
    say("refine losses", st["loss"])
    assert all(b <= a * (1 + 1e-6) for a, b in zip(st["loss"], st["loss"][1:]))


@pytest.mark.parametrize("ffs_mode,rotation", [(0, 1), (0, -1), (3, 1)])
def test_ray_projector_matches_oracle(ctx, oracle, ffs_mode, rotation):
    # The ray-driven projector (fct_wfbp_gpu_ray_project) and its transpose
    # against the FP64 restatement (which traces the physical rays, also the
    # clockwise and FFS ones, directly), and the GPU dot-product identity.
    w = S.tiny(nx=40, n_slices=10, n_channels=96, n_rows=8, ppr=96, rotations=1.5,
               ffs_mode=ffs_mode)
    sc = w.scan.replace(rotation=rotation, angle0=0.4)
    rc = w.recon.replace(recon_fov=220.0, y_origin=4.0)
    from freect_amd import wfbp
    p = wfbp.Plan(ctx, sc, rc)
    nz = p.info.n_slices_native
    gen = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand((nz, rc.ny, rc.nx), device="cuda", generator=gen)
    y = torch.rand((sc.n_views, sc.n_rows, sc.n_channels), device="cuda", generator=gen)
    px, pty = p.alloc_raw(), p.alloc_native()
    p.ray_project(x, px)
    p.ray_backproject(y, pty)
    p.synchronize()
    ref_px = oracle.ray_project(sc, rc, to_np(x).astype(np.float64))
    ref_pty = oracle.ray_backproject(sc, rc, to_np(y).astype(np.float64), nz)
    e1 = np.abs(to_np(px) - ref_px).max() / np.abs(ref_px).max()
    e2 = np.abs(to_np(pty) - ref_pty).max() / np.abs(ref_pty).max()
    rel = _rel(_dot(px, y), _dot(x, pty))
    say("ray projector ffs %d rot %d: P rel %.2e, P^T rel %.2e, dot %.2e" % (ffs_mode, rotation,
                                                                          e1, e2, rel))
    assert e1 < 2e-5 and e2 < 2e-5 and rel < 1e-5


def test_ray_refinement_lowers_hu_error(ctx):
    # The data-consistency refinement on a tiny analogue of the adjoint
    # workload (noiseless, analytic phantom): ten gradient steps on
    # 0.5 ||m (Pr x - p)||^2 from the WFBP volume, Pr the ray-driven
    # line-integral projector, step 1 / lambda_max from power iteration.  The
    # loss falls monotonically and the HU RMSE against the phantom (sampled
    # at the voxel centres, and averaged over each voxel) falls too.
    from freect_amd import autograd as AG
    from freect_amd import wfbp
    w = S.tiny(n_channels=192, n_rows=16, ppr=192, rotations=3, nx=64, n_slices=16, seed=5)
    sc, rc = w.scan, w.recon.replace(recon_fov=300.0)
    p = wfbp.Plan(ctx, sc, rc)
    raw = ctx.phantom_project(sc, w.phantom, 0.0, w.seed)
    x0 = AG.wfbp(raw, p).detach()
    info = p.info
    nz = info.n_slices_native
    zs = rc.start_pos + info.collimated_slice_width * np.arange(nz)
    dx = rc.recon_fov / rc.nx
    xs = (np.arange(rc.nx) - 0.5 * (rc.nx - 1)) * dx
    Z, Y, X = np.meshgrid(zs, xs, xs, indexing="ij")
    pts = w.phantom.mu_at(X, Y, Z)
    off = (np.arange(3) + 0.5) / 3 - 0.5
    vox = sum(w.phantom.mu_at(X + a * dx, Y + b * dx, Z + c * info.collimated_slice_width)
              for a in off for b in off for c in off) / 27.0
    m3 = ((X * X + Y * Y) <= (0.45 * info.acquisition_fov) ** 2)
    m3 &= np.isin(np.arange(nz), np.arange(nz // 4, 3 * nz // 4))[:, None, None]
    HU = S.MU_WATER / 1000.0

    def rmse(x, t):
        e = (to_np(x).astype(np.float64) - t)[m3] / HU
        return float(np.sqrt(np.mean(e * e)))

    mask = AG.inside_rays(p)
    x, st = AG.refine(p, raw, x0, steps=10, operator="ray", mask=mask)
    r0 = (rmse(x0, pts), rmse(x0, vox))
    r1 = (rmse(x, pts), rmse(x, vox))
    say("ray refinement: HU RMSE vs phantom %.2f -> %.2f (voxel-averaged %.2f -> %.2f); "
        "losses %s" % (r0[0], r1[0], r0[1], r1[1], ["%.4g" % v for v in st["loss"]]))
    assert all(b <= a * (1 + 1e-6) for a, b in zip(st["loss"], st["loss"][1:]))
    assert r1[0] < r0[0] and r1[1] < r0[1] - 1.0
