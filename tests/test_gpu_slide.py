"""Whole-slide path: the procedural slide is deterministic and overlap-consistent, its tiles are
segmented like any tile (oracle parity on crops), and the served canonical labels carry boxes and
areas that match the labels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_slide_tiles_agree_on_overlaps_and_repeat():
    from nscale_amd import gpu
    a = gpu.synth_slide_tile(5000, 7000, 700, 500, seed=11)
    b = gpu.synth_slide_tile(5636, 7000, 700, 500, seed=11)  # 64-px overlap to the right
    c = gpu.synth_slide_tile(5000, 7000, 700, 500, seed=11)
    torch.cuda.synchronize()
    assert torch.equal(a, c)
    assert torch.equal(a[:, 636:], b[:, :64])
    d = gpu.synth_slide_tile(5000, 7436, 700, 500, seed=11)  # 64-px overlap below
    assert torch.equal(a[436:], d[:64])
    e = gpu.synth_slide_tile(5000, 7000, 700, 500, seed=12)
    assert not torch.equal(a, e)
    # H&E-like content: mostly background, some dark nuclei
    h = a.cpu().numpy()
    bg = (h[..., 0] > 200) & (h[..., 2] > 200)
    assert 0.4 < bg.mean() < 0.98


@pytest.mark.parametrize("x0,y0", [(0, 0), (40000, 33000), (99000, 79000)])
def test_slide_crops_segment_like_the_oracle(x0, y0):
    from nscale_amd import gpu
    from tests._lib import o_segment
    w = min(512, 100000 - x0)
    h = min(384, 80000 - y0)
    tile = gpu.synth_slide_tile(x0, y0, w, h, seed=2026)
    st, lab, cc, _, _ = gpu.segment_nuclei(tile)
    torch.cuda.synchronize()
    ost, olab, _ = o_segment(tile.cpu().numpy())
    print("slide crop", x0, y0, st, cc, flush=True)
    assert st == ost
    if st == 0:
        assert np.array_equal(lab.cpu().numpy(), olab)


def test_served_slide_tiles_with_boxes():
    from nscale_amd import gpu
    tiles = [gpu.synth_slide_tile(x, y, 1024, 1024, seed=5) for x, y in [(0, 0), (960, 0), (0, 960), (960, 960)]]
    srv = gpu.Server(slots=4, hbm_budget=32 << 30, max_w=1024, max_h=1024)
    res = [srv.wait(t) for t in [srv.submit(t, canonical=True, max_labels=4096) for t in tiles]]
    srv.close()
    for t, (st, n, lab, bbox, area, r) in zip(tiles, res):
        assert r.error == 0 and st == 0 and n > 0
        ast, alab, acc, _, _ = gpu.segment_nuclei(t)
        torch.cuda.synchronize()
        assert n == acc
        lab = lab.cpu().numpy()
        ids = lab[lab > 0]
        assert np.array_equal(area.cpu().numpy(), np.bincount(ids, minlength=n + 1)[1:])
        ys, xs = np.nonzero(lab)
        ll = lab[ys, xs]
        b = bbox.cpu().numpy()
        mnx = np.full(n + 1, 1 << 30); np.minimum.at(mnx, ll, xs)
        assert np.array_equal(b[1], mnx[1:])
