"""Pins of the oracle's per-nucleus features (ObjFeatures.cpp / Operators.cpp restated in
oracle/oracle.cpp) against what the mathematics fixes: closed forms for rectangles and single
pixels, numpy / scipy statistics of the same pixels, numpy's gradient.  CPU only."""
import ctypes

import numpy as np
import pytest
from scipy import ndimage as ndi
from scipy import stats

from tests._lib import boxes_numpy, oracle, ptr

HS = 0.70710678118  # the reference's table value for a half-diagonal cell


def feats(lab, gray=None):
    """Oracle features of canonical labels 1..n (bg 0): dict of arrays."""
    h, w = lab.shape
    lab = np.ascontiguousarray(lab.astype(np.int32))
    box, _ = boxes_numpy(lab, bg=0)
    box = np.ascontiguousarray(box)
    n = box.shape[1]
    o = oracle()
    area = np.zeros(n, np.int32)
    o.orc_feat_area(ptr(box), n, ptr(lab), w, h, ptr(area))
    ma, mi, ec = (np.zeros(n) for _ in range(3))
    o.orc_feat_ellipse(ptr(box), ptr(area), n, ptr(lab), w, h, ptr(ma), ptr(mi), ptr(ec))
    per = np.zeros(n)
    o.orc_feat_perimeter(ptr(box), n, ptr(lab), w, h, ptr(per))
    ext, circ = np.zeros(n), np.zeros(n)
    o.orc_feat_extent_circ(ptr(box), ptr(area), ptr(per), n, ptr(ext), ptr(circ))
    out = dict(box=box, area=area, major=ma, minor=mi, ecc=ec, perimeter=per, extent=ext, circ=circ)
    if gray is not None:
        g = np.ascontiguousarray(gray.astype(np.uint8))
        f8 = np.zeros(8 * n, np.float32)
        o.orc_feat_intensity(ptr(box), n, ptr(lab), ptr(g), w, h, ptr(f8))
        f6 = np.zeros(6 * n, np.float32)
        o.orc_feat_gradient(ptr(box), n, ptr(lab), ptr(g), w, h, ptr(f6))
        out["intensity"] = f8.reshape(n, 8)
        out["gradient"] = f6.reshape(n, 6)
    return out


def test_rectangles_and_single_pixels_closed_forms():
    lab = np.zeros((60, 80), np.int32)
    rects = [(3, 5, 7, 4), (20, 10, 1, 1), (30, 40, 12, 3), (50, 2, 2, 9)]  # y, x, height, width
    for k, (y, x, hh, ww) in enumerate(rects, 1):
        lab[y:y + hh, x:x + ww] = k
    f = feats(lab)
    for k, (y, x, hh, ww) in enumerate(rects):
        a, b = max(hh, ww), min(hh, ww)
        assert f["area"][k] == hh * ww
        # normalised second moments of an a x b block (with the 1/12 pixel term): a^2/12, b^2/12
        assert np.isclose(f["major"][k], 2 / np.sqrt(3) * a, rtol=1e-12)
        assert np.isclose(f["minor"][k], 2 / np.sqrt(3) * b, rtol=1e-12)
        assert np.isclose(f["ecc"][k], np.sqrt(1 - (b / a) ** 2), rtol=1e-9, atol=1e-12)
        # marching squares through the pixel centres: straight runs plus four cut corners
        assert np.isclose(f["perimeter"][k], 2 * (hh - 1) + 2 * (ww - 1) + 4 * HS, rtol=1e-12)
        assert f["extent"][k] == 1.0
        assert np.isclose(f["circ"][k], 4 * 3.14159265359 * hh * ww / f["perimeter"][k] ** 2, rtol=1e-12)


def _blobs(seed, h=120, w=140):
    rng = np.random.default_rng(seed)
    m = ndi.gaussian_filter(rng.random((h, w)), 2.5) > 0.52
    lab, n = ndi.label(m, structure=np.ones((3, 3)))
    return lab.astype(np.int32), n


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_moments_and_areas_against_numpy(seed):
    lab, n = _blobs(seed)
    assert n > 5
    f = feats(lab)
    assert np.array_equal(f["area"], np.bincount(lab.ravel())[1:])
    for k in range(n):
        ys, xs = np.nonzero(lab == k + 1)
        b = f["box"][:, k]
        cx = xs - (b[1] + b[2]) / 2.0
        cy = ys - (b[3] + b[4]) / 2.0
        mxx = np.mean((cx - cx.mean()) ** 2) + 1 / 12
        myy = np.mean((cy - cy.mean()) ** 2) + 1 / 12
        mxy = np.mean((cx - cx.mean()) * (cy - cy.mean()))
        ev = np.linalg.eigvalsh(np.array([[mxx, mxy], [mxy, myy]]))  # principal second moments
        assert np.isclose(f["major"][k], np.sqrt(8) * np.sqrt(2 * ev[1]), rtol=1e-9)
        assert np.isclose(f["minor"][k], np.sqrt(8) * np.sqrt(max(2 * ev[0], 0)), rtol=1e-7, atol=1e-7)
        box_px = (b[2] - b[1] + 1) * (b[4] - b[3] + 1)
        assert np.isclose(f["extent"][k], len(xs) / box_px, rtol=1e-15)


@pytest.mark.parametrize("seed", [4, 5])
def test_perimeter_against_padded_marching_squares(seed):
    """The reference's box-edge half cells equal marching squares over the box padded by one empty
    pixel: every 2x2 cell of the padded box, looked up in the same table."""
    lab, n = _blobs(seed)
    f = feats(lab)
    lut = np.zeros(16)
    lut[[1, 2, 4, 8, 7, 11, 13, 14]] = HS
    lut[[3, 6, 9, 12]] = 1.0
    lut[[5, 10]] = 1.41421356237
    for k in range(n):
        b = f["box"][:, k]
        m = np.pad((lab[b[3]:b[4] + 1, b[1]:b[2] + 1] == k + 1).astype(int), 1)
        code = (m[:-1, :-1] << 3) | (m[:-1, 1:] << 2) | (m[1:, 1:] << 1) | m[1:, :-1]
        assert np.isclose(f["perimeter"][k], lut[code].sum(), rtol=1e-12)


@pytest.mark.parametrize("seed", [6, 7])
def test_intensity_statistics_against_scipy(seed):
    lab, n = _blobs(seed)
    gray = np.random.default_rng(seed).integers(0, 256, lab.shape).astype(np.uint8)
    f = feats(lab, gray)["intensity"]
    for k in range(n):
        v = gray[lab == k + 1].astype(np.float64)
        p = np.bincount(v.astype(int), minlength=256) / v.size
        exp = [v.mean(), v.max(), v.min(), v.std(), stats.entropy(p, base=2), np.sum(p ** 2),
               stats.skew(v), stats.kurtosis(v, fisher=False)]
        exp = np.nan_to_num(exp)  # a constant region: the reference's guards give 0 skewness / kurtosis
        assert np.allclose(f[k], exp, rtol=2e-5, atol=2e-5), (k, f[k], exp)


def test_gradient_image_against_numpy_gradient():
    from tests._lib import oracle as orc
    rng = np.random.default_rng(8)
    g = rng.integers(0, 256, (37, 53)).astype(np.uint8)
    out = np.zeros_like(g)
    orc().orc_gradient_u8(ptr(g), ptr(out), 53, 37)
    gy, gx = np.gradient(g.astype(np.float64))  # central inside, one-sided on the border
    mag = np.sqrt(gx ** 2 + gy ** 2)
    exp = np.clip(np.floor(mag + 0.5), 0, 255).astype(np.uint8)
    assert np.array_equal(out, exp)


def test_gradient_features_against_scipy():
    lab, n = _blobs(9)
    gray = ndi.gaussian_filter(np.random.default_rng(9).integers(0, 256, lab.shape).astype(float), 1.0)
    gray = gray.astype(np.uint8)
    f = feats(lab, gray)["gradient"]
    gy, gx = np.gradient(gray.astype(np.float64))
    gm = np.clip(np.floor(np.sqrt(gx ** 2 + gy ** 2) + 0.5), 0, 255)
    for k in range(n):
        v = gm[lab == k + 1]
        p = np.bincount(v.astype(int), minlength=256) / v.size
        exp = np.nan_to_num([v.mean(), v.std(), stats.entropy(p, base=2), np.sum(p ** 2), stats.skew(v),
                             stats.kurtosis(v, fisher=False)])
        assert np.allclose(f[k], exp, rtol=2e-5, atol=2e-5), (k, f[k], exp)
