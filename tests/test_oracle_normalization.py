"""Pins of the oracle's colour normalization (Normalization.cpp restated in oracle/oracle.cpp)
against what the mathematics fixes: the discriminant's exact rational form over all 2^24
colours, Reinhard's transforms evaluated in float64 (the published method, as the reference's
MATLAB code writes it: src/normalization/matlab/ColorNormalization/rgb2lab.m, lab2rgb.m), the
round trip through the two 4-digit matrices, and the defining property of the transfer (the
foreground statistics move to the target).  CPU only."""
import numpy as np
import pytest

from tests._lib import (all_colours, o_norm_bgr, o_norm_fg, o_norm_lab, o_norm_target, o_normalize)

RGB2LMS = np.array([[0.3811, 0.5783, 0.0402], [0.1967, 0.7244, 0.0782], [0.0241, 0.1288, 0.8444]])
LMS2RGB = np.array([[4.4687, -3.5887, 0.1196], [-1.2197, 2.3831, -0.1626], [0.0585, -0.2611, 1.2057]])
LOG2LAB = np.diag([1 / np.sqrt(3), 1 / np.sqrt(6), 1 / np.sqrt(2)]) @ np.array([[1, 1, 1], [1, 1, -2], [1, -1, 0]])
LAB2LOG = np.array([[1, 1, 1], [1, 1, -1], [1, -2, 0]]) @ np.diag([np.sqrt(3) / 3, np.sqrt(6) / 6, np.sqrt(2) / 2])
# target statistics of the reference's normalization example (src/normalization/examples/normTest.cpp:34-35)
MEAN_T = [-0.176, -0.0033, 0.0143]
STD_T = [0.2317, 0.0491, 0.0156]


def lab64(bgr):
    rgb = bgr[..., ::-1].astype(np.float64) / 255
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.log10(rgb @ RGB2LMS.T) @ LOG2LAB.T


def bgr64(lab):
    with np.errstate(invalid="ignore", over="ignore"):
        lms = 10.0 ** (lab @ LAB2LOG.T)
    lms[~np.isfinite(lms)] = 0
    rgb = lms @ LMS2RGB.T * 255
    return np.clip(np.floor(rgb + 0.5), 0, 255)[..., ::-1]


@pytest.fixture(scope="module")
def colours():
    return all_colours()


def test_foreground_discriminant_over_every_colour(colours):
    """d2 - d1 = (97 r - 852 g + 621 b - 4169) / 1000 exactly: wherever that numerator is not zero
    (|d2 - d1| >= 0.001, far above the float rounding of values near 100) the mask is its sign."""
    fg = o_norm_fg(colours)
    b, g, r = (colours[..., k].astype(np.int64) for k in range(3))
    num = 97 * r - 852 * g + 621 * b - 4169
    assert np.array_equal(fg[num != 0] == 255, num[num != 0] > 0)
    assert set(np.unique(fg).tolist()) <= {0, 255}
    ties = int((num == 0).sum())
    assert 0 < ties < 2000  # exact ties: decided by the float roundings, kept as the reference computes them


def test_lab_against_float64(colours):
    lab = o_norm_lab(colours)
    ref = lab64(colours)
    black = (colours == 0).all(-1)
    assert np.isneginf(lab[black][:, 0]).all() and np.isnan(lab[black][:, 1:]).all()
    assert np.allclose(lab[~black], ref[~black], rtol=0, atol=4e-6)


def test_round_trip_through_both_matrices_is_close_to_identity(colours):
    """lab2BGR(bgr2Lab(x)): the 4-digit LMS->RGB matrix inverts RGB->LMS to about 1e-4, so every colour
    comes back within one level; a transposed or mis-signed matrix would move most colours."""
    back = o_norm_bgr(o_norm_lab(colours)).astype(np.int32)
    d = np.abs(back - colours.astype(np.int32))
    black = (colours == 0).all(-1)
    assert (back[black] == 0).all()
    assert d.max() <= 1
    ref = bgr64(lab64(colours))
    ref[black] = 0
    assert np.abs(back - ref).max() <= 1
    assert (back != ref).mean() < 1e-3


def synth_tile(seed, h=256, w=320):
    from nscale_amd.synth import synth_he
    return synth_he(h, w, seed)


@pytest.mark.parametrize("seed", [1, 2])
def test_target_parameters_against_float64(seed):
    bgr = synth_tile(seed)
    m, s, n = o_norm_target(bgr)
    fg = o_norm_fg(bgr) == 255
    lab = o_norm_lab(bgr)
    sel = fg & ((lab[..., 0] * lab[..., 1]) * lab[..., 2] > 0)
    assert n == int(sel.sum()) and n > 1000
    ref = lab64(bgr)[sel]
    assert np.allclose(m, ref.mean(0), rtol=1e-5, atol=1e-6)
    assert np.allclose(s, ref.std(0), rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("seed", [3, 4])
def test_normalization_against_float64(seed):
    bgr = synth_tile(seed)
    out = o_normalize(bgr, MEAN_T, STD_T).astype(np.int32)
    m, s, _ = o_norm_target(bgr)
    lab = lab64(bgr)
    fg = o_norm_fg(bgr) == 255
    lab[fg] = (lab[fg] - m.astype(np.float64)) * (np.float64(STD_T) / s) + np.float64(MEAN_T)
    ref = bgr64(lab)
    ref[(bgr == 0).all(-1)] = 0
    d = np.abs(out - ref)
    assert d.max() <= 1 and (d > 0).mean() < 1e-3
    assert fg.mean() > 0.05 and (out[fg] != bgr[fg]).any(-1).mean() > 0.5  # the transfer moved the foreground


def test_own_statistics_are_a_round_trip():
    """Normalizing to the image's own statistics changes each foreground LAB value by float rounding
    only: the result equals the plain LAB round trip to within one level."""
    bgr = synth_tile(5)
    m, s, _ = o_norm_target(bgr)
    out = o_normalize(bgr, m, s).astype(np.int32)
    rt = o_norm_bgr(o_norm_lab(bgr)).astype(np.int32)
    assert np.abs(out - rt).max() <= 1
    assert (out != rt).mean() < 1e-3


def test_transfer_moves_statistics_to_the_target():
    """The defining property of the transfer: over the pixels the statistics were taken from, the
    normalized image's LAB values have the target's mean and standard deviation, up to the u8
    quantisation of the output."""
    bgr = synth_tile(6, 512, 512)
    m, s, _ = o_norm_target(bgr)
    tm = m + np.float32([-0.03, 0.004, -0.003])  # a target close enough that no channel clips
    ts = s * np.float32([0.8, 1.1, 0.9])
    out = o_normalize(bgr, tm, ts)
    lab = o_norm_lab(bgr)
    sel = (o_norm_fg(bgr) == 255) & ((lab[..., 0] * lab[..., 1]) * lab[..., 2] > 0)
    assert ((out[sel] == 0) | (out[sel] == 255)).mean() < 1e-3
    got = lab64(out)[sel]
    assert np.allclose(got.mean(0), tm, rtol=0, atol=2e-4), (got.mean(0), tm)
    assert np.allclose(got.std(0), ts, rtol=5e-3), (got.std(0), ts)


def test_uniform_and_empty_foreground():
    """No foreground: every pixel takes the plain round trip.  Black stays black."""
    bgr = np.zeros((40, 50, 3), np.uint8)
    bgr[:, :25] = (230, 230, 230)  # white background
    m, s, n = o_norm_target(bgr)
    assert n == 0 and np.isnan(m).all()
    out = o_normalize(bgr, MEAN_T, STD_T)
    assert np.array_equal(out, o_norm_bgr(o_norm_lab(bgr)))
    assert (out[:, 25:] == 0).all()
