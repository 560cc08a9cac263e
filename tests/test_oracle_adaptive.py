"""Pins of the FP64 restatement of the adaptive filtration (no reference
implementation exists: FreeCT parses adaptive_filtration_s,
parse_config.cpp:57, and lists the filter as missing, README.md:14), by the
properties that define it rather than by its formula:

* samples below (1 - s) vmax of their view are unchanged, bit for bit;
* a view that is constant, or linear in (row, channel), is a fixed point
  away from the detector edges (the 3 x 3 mean of a plane is its centre);
* the most attenuated sample of a view becomes the mean of its 3 x 3
  neighbourhood (lambda = 1);
* on a noisy high-attenuation plateau the variance drops to about 1/9;
* a larger strength never smooths less, and non-finite samples count as 0.
"""
import numpy as np
import pytest

from freect_amd import scans as S


def _scan(nv=6, nr=8, nc=40):
    w = S.tiny(n_channels=nc, n_rows=nr, ppr=32, rotations=2.0)
    return w.scan.replace(n_views=nv)


def test_below_threshold_unchanged_and_max_is_local_mean(oracle):
    sc = _scan()
    rng = np.random.default_rng(1)
    raw = rng.uniform(0.0, 5.0, (sc.n_views, sc.n_rows, sc.n_channels)).astype(np.float32)
    s = 0.2
    out = oracle.adaptive_filter(sc, raw, s)
    for v in range(sc.n_views):
        vm = np.abs(raw[v]).max()
        low = raw[v] < (1 - s) * vm
        assert low.any() and np.array_equal(out[v][low], raw[v][low])
        r, c = np.unravel_index(np.argmax(raw[v]), raw[v].shape)
        nb = raw[v][max(r - 1, 0):r + 2, max(c - 1, 0):c + 2].astype(np.float64)
        assert abs(out[v, r, c] - nb.mean()) <= 1e-6 * vm


@pytest.mark.parametrize("kind", ["constant", "plane"])
def test_constant_and_plane_are_fixed_points(oracle, kind):
    sc = _scan()
    r = np.arange(sc.n_rows)[:, None]
    c = np.arange(sc.n_channels)[None, :]
    view = np.full((sc.n_rows, sc.n_channels), 7.0) if kind == "constant" else \
        3.0 + 0.25 * r + 0.1 * c
    raw = np.broadcast_to(view, (sc.n_views,) + view.shape).astype(np.float32).copy()
    out = oracle.adaptive_filter(sc, raw, 1.0)
    inner = (slice(None), slice(1, -1), slice(1, -1))
    assert np.max(np.abs(out[inner] - raw[inner])) <= 2e-6 * raw.max()
    if kind == "plane":   # at the edges the neighbourhood is one-sided: the plane moves
        assert np.max(np.abs(out - raw)) > 1e-3


def test_noise_on_a_plateau_is_reduced(oracle):
    sc = _scan(nv=4, nr=32, nc=128)
    rng = np.random.default_rng(2)
    raw = (8.0 + 0.05 * rng.standard_normal((sc.n_views, sc.n_rows, sc.n_channels))).astype(
        np.float32)
    out = oracle.adaptive_filter(sc, raw, 1.0)   # every sample is above 0 = (1 - s) vmax
    inner = (slice(None), slice(2, -2), slice(2, -2))
    ratio = out[inner].var() / raw[inner].var()
    assert 0.08 < ratio < 0.2    # about 1/9 (lambda ~ 1 on the plateau)


def test_strength_is_monotone_and_non_finite_reads_zero(oracle):
    sc = _scan()
    rng = np.random.default_rng(3)
    raw = rng.uniform(0.0, 6.0, (sc.n_views, sc.n_rows, sc.n_channels)).astype(np.float32)
    dev = [np.abs(oracle.adaptive_filter(sc, raw, s) - raw).sum() for s in (0.05, 0.2, 0.6, 1.0)]
    assert all(a <= b for a, b in zip(dev, dev[1:])) and dev[0] > 0
    bad = raw.copy()
    bad[1, 3, 7] = np.nan
    bad[2, 0, 0] = np.inf
    zero = raw.copy()
    zero[1, 3, 7] = 0.0
    zero[2, 0, 0] = 0.0
    assert np.array_equal(oracle.adaptive_filter(sc, bad, 0.3),
                          oracle.adaptive_filter(sc, zero, 0.3))
