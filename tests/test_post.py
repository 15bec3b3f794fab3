"""Gradient post-processing: the oracle (oracle/post_oracle.c) against the
reference's own taper_grad_shot / taper_grad / smooth compiled from source
(oracle/_ref), and against brute force (sorted edge-padded windows with the
reference's off-by-one).  Bit-exact for the tapers; the smoothing is a
selection, so equal values are required (a zero's sign may differ: the
reference's sort does not order -0 and +0)."""
import numpy as np
import pytest

import oracle_lib as ol

needs_ref = pytest.mark.skipif(not ol.ref_available(), reason="reference tree not built into oracle/_ref")


def _grad(ny, nx, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((ny, nx)).astype(np.float32)
    g[rng.random((ny, nx)) < 0.05] = 0.0
    return g


@needs_ref
@pytest.mark.parametrize("shape", [1, 2])
@pytest.mark.parametrize("radius,filtsize,DH", [(10.0, 1, 2.0), (25.0, 3, 2.0), (3.0, 0, 1.0), (40.0, 2, 5.0)])
@pytest.mark.parametrize("pos", [(30, 12), (4, 3), (2, 20), (60, 39)])
def test_taper_shot_matches_reference(shape, radius, filtsize, DH, pos, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    ny, nx = 40, 64
    g = _grad(ny, nx, 11)
    isrc, jsrc = [17, pos[0]], [9, pos[1]]
    ref = ol.taper_sources(g, DH, radius, shape, filtsize, isrc, jsrc, which="ref", ishot=2)
    orc = ol.taper_sources(g, DH, radius, shape, filtsize, isrc, jsrc, ishot=2)
    np.testing.assert_array_equal(orc.view(np.uint32), ref.view(np.uint32))
    # the taper and the box did something
    assert not np.array_equal(orc, g)


@needs_ref
def test_taper_shot_box_wraps_left_of_the_grid(tmp_path, monkeypatch):
    """A box reaching two or more columns left of the grid zeroes cells at
    the right end of the row above, as the reference's scratch matrix does."""
    monkeypatch.chdir(tmp_path)
    ny, nx = 24, 32
    g = np.ones((ny, nx), np.float32)
    ref = ol.taper_sources(g, 1.0, 2.0, 1, 4, [1], [10], which="ref", ishot=1)
    orc = ol.taper_sources(g, 1.0, 2.0, 1, 4, [1], [10], ishot=1)
    np.testing.assert_array_equal(orc, ref)
    # box rows 6..14 (1-based), columns -3..5; columns -3, -2 land in the
    # previous row at NX+2+i = 31, 32
    assert orc[4, 30] == 0.0 and orc[4, 31] == 0.0 and orc[12, 31] == 0.0
    assert orc[14, 31] != 0.0


@needs_ref
@pytest.mark.parametrize("shape", [1, 2])
def test_taper_all_sources_matches_reference(shape, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    ny, nx = 36, 80
    g = _grad(ny, nx, 5)
    isrc, jsrc = [10, 30, 31, 70], [5, 5, 18, 30]
    ref = ol.taper_sources(g, 2.0, 12.0, shape, 2, isrc, jsrc, which="ref", ishot=0)
    orc = ol.taper_sources(g, 2.0, 12.0, shape, 2, isrc, jsrc, ishot=0)
    np.testing.assert_array_equal(orc.view(np.uint32), ref.view(np.uint32))


@needs_ref
@pytest.mark.parametrize("gradt", [(5, 15, 50, 70), (1, 2, 78, 79), (20, 21, 30, 60)])
def test_taper_vertical_matches_reference(gradt, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    ny, nx = 30, 80
    g = _grad(ny, nx, 3)
    rec = np.array([[x, 4] for x in range(3, 78, 7)], np.int32)
    ref = ol.taper_vertical(g, gradt, [12, 40], [3, 6], rec, which="ref")
    orc = ol.taper_vertical(g, gradt, [12, 40], [3, 6], rec)
    np.testing.assert_array_equal(orc.view(np.uint32), ref.view(np.uint32))
    assert orc[3, 37] == 0.0 and orc[5, 39] == 0.0  # receiver (38, 4), source (40, 6)


@needs_ref
@pytest.mark.parametrize("gradt", [(3, 10, 0, 0), (1, 1, 0, 0), (12, 25, 0, 0)])
def test_taper_horizontal_matches_reference(gradt, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    ny, nx = 30, 20
    g = _grad(ny, nx, 4)
    ref = ol.taper_horizontal(g, gradt, which="ref")
    orc = ol.taper_horizontal(g, gradt)
    np.testing.assert_array_equal(orc.view(np.uint32), ref.view(np.uint32))


def test_taper_horizontal_closed_form():
    """Rows above GRADT1 vanish, rows below GRADT2 pass, the flank is
    exp(-(3 k / (len/2))^2 / 2) with k = row - GRADT2."""
    ny, nx, t1, t2 = 20, 3, 4, 9
    out = ol.taper_horizontal(np.ones((ny, nx), np.float32), (t1, t2))
    L = t2 - t1
    for j in range(1, ny + 1):
        if j < t1:
            want = 0.0
        elif j <= t2:
            k = j - t1 + 1
            want = np.exp(-0.5 * (3.0 * (k - L) / (L / 2.0)) ** 2) if k <= L else 1.0
        else:
            want = 1.0
        np.testing.assert_allclose(out[j - 1], want, rtol=1e-6, atol=0)


def test_taper_shot_closed_form_centre_and_far_field():
    """Far from the source the taper is 1; the box around it is 0; the
    erf shape rises monotonically with distance along a row."""
    ny, nx = 60, 60
    out = ol.taper_sources(np.ones((ny, nx), np.float32), 1.0, 8.0, 1, 1, [30], [30])
    assert out[0, 0] == 1.0 and out[59, 59] == 1.0
    assert np.all(out[28:31, 28:31] == 0.0)
    row = out[29, 31:40]
    assert np.all(np.diff(row) >= 0) and row[-1] == 1.0


def _median_numpy(g, fs):
    """smooth()'s value by brute force: rank k//2+1 of the window with its
    bottom-right corner replaced by a zero (median2d sorts t[0..k-1] of a
    vector whose data sit in t[1..k], t[0] = 0)."""
    if fs == 0:
        return g.copy()
    if fs % 2 == 0:
        fs += 1 if fs > 0 else -1
    h = abs(fs) // 2
    if h == 0:
        return g.copy()
    p = np.pad(g, h, mode="edge")
    ny, nx = g.shape
    out = np.empty_like(g)
    for j in range(ny):
        for i in range(nx):
            w = p[j:j + 2 * h + 1, i:i + 2 * h + 1].copy()
            w[-1, -1] = 0.0
            out[j, i] = np.sort(w, axis=None)[w.size // 2 + 1]
    return out


@pytest.mark.parametrize("fs", [0, 1, 3, 4, 5, -3, 7])
def test_median_oracle_closed_form(fs):
    g = _grad(17, 23, 8)
    g[3:6, 4:9] = 1.5  # duplicates
    np.testing.assert_array_equal(ol.median_smooth(g, fs), _median_numpy(g, fs))


@needs_ref
@pytest.mark.parametrize("fs", [0, 1, 3, 4, 5, -3, 9])
def test_median_matches_reference(fs):
    g = _grad(21, 26, 9)
    g[3:6, 4:9] = 1.5
    ref = ol.median_smooth(g, fs, which="ref")
    orc = ol.median_smooth(g, fs)
    np.testing.assert_array_equal(orc, ref)


# --- the product's host taper builder (ifos_taper_coeff) against the oracle


def _apply(g, coef, zero):
    out = g * coef
    out[zero.astype(bool)] = 0.0
    return out


@pytest.mark.parametrize("shape,radius,fs,DH", [(1, 14.0, 4, 2.0), (2, 9.0, 1, 2.0), (1, 3.0, 0, 1.0),
                                                (2, 40.0, 7, 5.0)])
def test_host_source_taper_matches_oracle(shape, radius, fs, DH):
    from ifos_amd import opt as P
    NY, NX = 45, 70
    g = _grad(NY, NX, 21)
    src = np.array([[12, 4], [40, 6], [2, 20], [69, 44]], np.int32)
    coef, zero = P.taper_coeff(P.TAPER_SOURCES, NX, NY, DH=DH, SRTRADIUS=radius, SRTSHAPE=shape, FILTSIZE=fs,
                               srcpos=src)
    want = ol.taper_sources(g, DH, radius, shape, fs, src[:, 0], src[:, 1], ishot=0)
    np.testing.assert_array_equal(_apply(g, coef, zero).view(np.uint32), want.view(np.uint32))
    # one shot: the table helpers give the same coefficients
    for k in range(len(src)):
        c1, z1 = P.taper_coeff(P.TAPER_SOURCES, NX, NY, DH=DH, SRTRADIUS=radius, SRTSHAPE=shape, FILTSIZE=fs,
                               srcpos=src[k:k + 1])
        want1 = ol.taper_sources(g, DH, radius, shape, fs, src[k:k + 1, 0], src[k:k + 1, 1], ishot=1)
        np.testing.assert_array_equal(_apply(g, c1, z1).view(np.uint32), want1.view(np.uint32))


@pytest.mark.parametrize("gradt", [(6, 14, 50, 62), (1, 2, 68, 69)])
def test_host_vertical_taper_matches_oracle(gradt):
    from ifos_amd import opt as P
    NY, NX = 30, 70
    g = _grad(NY, NX, 22)
    src = np.array([[12, 4], [40, 6]], np.int32)
    rec = np.array([[x, 3] for x in range(5, 66, 6)], np.int32)
    coef, zero = P.taper_coeff(P.TAPER_VERTICAL, NX, NY, gradt=gradt, srcpos=src, recpos=rec)
    want = ol.taper_vertical(g, gradt, src[:, 0], src[:, 1], rec)
    np.testing.assert_array_equal(_apply(g, coef, zero).view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("gradt", [(4, 11, 0, 0), (1, 1, 0, 0)])
def test_host_horizontal_taper_matches_oracle(gradt):
    from ifos_amd import opt as P
    NY, NX = 30, 20
    g = _grad(NY, NX, 23)
    coef, zero = P.taper_coeff(P.TAPER_HORIZONTAL, NX, NY, gradt=gradt)
    assert not zero.any()
    want = ol.taper_horizontal(g, gradt)
    np.testing.assert_array_equal(_apply(g, coef, zero).view(np.uint32), want.view(np.uint32))


@needs_ref
@pytest.mark.parametrize("size,first", [(3, 1), (5, 4), (1, 10), (0, 1)])
def test_spat_filt_matches_reference(size, first):
    g = _grad(25, 40, 31)
    ref = ol.spat_filt(g, size, first, which="ref")
    orc = ol.spat_filt(g, size, first)
    np.testing.assert_array_equal(orc.view(np.uint32), ref.view(np.uint32))


def test_spat_filt_is_a_boxcar_mean():
    """The integer (-1/2) in the reference's Gaussian makes every weight 1:
    inside the filtered window each value is the mean of 2S+1 neighbours."""
    g = _grad(12, 30, 32)
    out = ol.spat_filt(g, 2, 3)
    want = g.copy()
    for j in range(3, 12 - 2 + 1):
        for i in range(3, 30 - 2 + 1):
            want[j - 1, i - 1] = g[j - 1, i - 3:i + 2].astype(np.float64).mean()
    np.testing.assert_allclose(out, want, rtol=2e-6, atol=1e-7)
    np.testing.assert_array_equal(out[:2], g[:2])
