"""HIP colour normalization (Normalization.cpp) against the oracle: the statistics to double
rounding, the normalized image to one level on a small fraction of channel values (the oracle's
log10f / powf are the C library's float functions, the kernels round double results once), over
synthetic H&E tiles up to 4096^2, every 24-bit colour, pitched and in-place buffers, odd sizes and
an image without foreground."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MEAN_T = [-0.632356, -0.0516004, 0.0376543]  # segment/test/src/nu-segmentTest.cpp:185-187
STD_T = [0.26235, 0.0514831, 0.0114217]


def _close(got, exp, frac=1e-6):
    d = np.abs(got.astype(np.int32) - exp.astype(np.int32))
    print("channel values one level apart: %d of %d" % (int((d > 0).sum()), d.size), flush=True)
    assert d.max() <= 1, d.max()
    assert (d > 0).mean() <= frac, (d > 0).mean()


@pytest.mark.parametrize("h,w,seed", [(512, 512, 1), (333, 1500, 2), (1, 1, 3), (7, 3, 4), (4096, 4096, 2026)])
def test_target_parameters_and_normalization_match_the_oracle(h, w, seed):
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he
    from tests._lib import o_norm_target, o_normalize
    bgr = synth_he(h, w, seed)
    print("normalization", h, w, seed, flush=True)
    dev = torch.from_numpy(bgr).cuda()
    m, s, n = gpu.normalization_target_parameters(dev)
    om, os_, on = o_norm_target(bgr)
    if on == 0:
        assert n == 0 and np.isnan(m).all()
    else:
        assert abs(n - on) <= max(1, on * 1e-5)
        assert np.allclose(m, om, rtol=1e-6, atol=1e-9) and np.allclose(s, os_, rtol=1e-5, atol=1e-9)
    out = gpu.normalization(dev, MEAN_T, STD_T)
    torch.cuda.synchronize()
    _close(out.cpu().numpy(), o_normalize(bgr, MEAN_T, STD_T))


def test_every_colour():
    """All 2^24 colours in one 4096^2 image: every discriminant tie and every LAB value occurs."""
    from nscale_amd import gpu
    from tests._lib import all_colours, o_norm_target, o_normalize
    bgr = all_colours()
    dev = torch.from_numpy(bgr).cuda()
    m, s, n = gpu.normalization_target_parameters(dev)
    om, os_, on = o_norm_target(bgr)
    assert abs(n - on) <= 16 and np.allclose(m, om, rtol=1e-6) and np.allclose(s, os_, rtol=1e-5)
    out = gpu.normalization(dev, MEAN_T, STD_T)
    torch.cuda.synchronize()
    _close(out.cpu().numpy(), o_normalize(bgr, MEAN_T, STD_T))


def test_pitched_and_in_place():
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he
    bgr = synth_he(300, 401, 7)
    dense = torch.from_numpy(bgr).cuda()
    ref = gpu.normalization(dense, MEAN_T, STD_T).cpu().numpy()
    big = torch.zeros(300, 512, 3, dtype=torch.uint8, device="cuda")
    big[:, :401] = dense
    view = big[:, :401]  # rows 512 * 3 bytes apart
    assert np.allclose(gpu.normalization_target_parameters(view)[0], gpu.normalization_target_parameters(dense)[0])
    outbig = torch.zeros(300, 640, 3, dtype=torch.uint8, device="cuda")
    gpu.normalization(view, MEAN_T, STD_T, out=outbig[:, 100:501])
    torch.cuda.synchronize()
    assert np.array_equal(outbig[:, 100:501].cpu().numpy(), ref)
    assert (outbig[:, :100] == 0).all() and (outbig[:, 501:] == 0).all()
    gpu.normalization(view, MEAN_T, STD_T, out=view)  # in place
    torch.cuda.synchronize()
    assert np.array_equal(view.cpu().numpy(), ref)
    assert (big[:, 401:] == 0).all()


def test_no_foreground_and_overlap_rejected():
    from nscale_amd import gpu
    from nscale_amd._lib import NscError
    from tests._lib import o_norm_bgr, o_norm_lab
    bgr = np.zeros((40, 50, 3), np.uint8)
    bgr[:, :25] = (230, 230, 230)
    dev = torch.from_numpy(bgr).cuda()
    m, s, n = gpu.normalization_target_parameters(dev)
    assert n == 0 and np.isnan(m).all() and np.isnan(s).all()
    out = gpu.normalization(dev, MEAN_T, STD_T).cpu().numpy()
    assert np.array_equal(out, o_norm_bgr(o_norm_lab(bgr)))
    buf = torch.zeros(41, 50, 3, dtype=torch.uint8, device="cuda")
    with pytest.raises(NscError):
        gpu.normalization(buf[:40], MEAN_T, STD_T, out=buf[1:])
