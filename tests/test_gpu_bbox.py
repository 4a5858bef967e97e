"""Bounding boxes and areas of canonical labels on the GPU (nsc_bounding_boxes, ccl.hip k_bbox)
against numpy min/max/bincount per label -- what the reference's ConnComponents::boundingBox
returns (ConnComponents.cpp:241-284).  Shapes cover row wraps inside a 64-pixel window, windows
that end mid-run, one-pixel-wide and one-row images and labels that touch every border."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _expect(L, n):
    ys, xs = np.nonzero((L >= 1) & (L <= n))
    ids = L[ys, xs]
    mnx = np.full(n + 1, 1 << 30); np.minimum.at(mnx, ids, xs)
    mxx = np.full(n + 1, -1); np.maximum.at(mxx, ids, xs)
    mny = np.full(n + 1, 1 << 30); np.minimum.at(mny, ids, ys)
    mxy = np.full(n + 1, -1); np.maximum.at(mxy, ids, ys)
    return np.stack([np.arange(1, n + 1), mnx[1:], mxx[1:], mny[1:], mxy[1:]]), np.bincount(ids, minlength=n + 1)[1:]


def _cases():
    rng = np.random.default_rng(41)
    for h, w in [(1, 1), (70, 1), (1, 70), (9, 63), (7, 64), (8, 65), (23, 100), (257, 333)]:
        yield "noise%dx%d" % (h, w), rng.integers(0, 7, (h, w)).astype(np.int32), 6
        blobs = (rng.random(((h + 3) // 4, (w + 3) // 4)) < 0.6).repeat(4, 0).repeat(4, 1)[:h, :w]
        yield "blobs%dx%d" % (h, w), blobs.astype(np.int32) * rng.integers(1, 3, (h, w)).astype(np.int32), 2
    # many disc components (labels 1..n in raster order of their centres), some on the borders
    h, w = 777, 1000
    yy, xx = np.mgrid[0:h, 0:w]
    L = np.zeros((h, w), np.int32)
    cy, cx = rng.integers(0, h, 400), rng.integers(0, w, 400)
    order = np.lexsort((cx, cy))
    for k, j in enumerate(order, 1):
        r = int(rng.integers(2, 15))
        y0, y1, x0, x1 = max(0, cy[j] - r), min(h, cy[j] + r + 1), max(0, cx[j] - r), min(w, cx[j] + r + 1)
        sub = (yy[y0:y1, x0:x1] - cy[j]) ** 2 + (xx[y0:y1, x0:x1] - cx[j]) ** 2 <= r * r
        L[y0:y1, x0:x1][sub] = k
    yield "discs", L, 400


@pytest.mark.parametrize("case", list(range(17)))
def test_bounding_boxes_match_numpy(case):
    from nscale_amd import gpu
    name, L, n = list(_cases())[case]
    print("bbox", name, L.shape, n, flush=True)
    box, area = gpu.bounding_boxes(torch.from_numpy(L).cuda(), n)
    torch.cuda.synchronize()
    ebox, earea = _expect(L, n)
    present = earea > 0  # labels with no pixel keep the initial (empty) box
    assert np.array_equal(box.cpu().numpy()[:, present], ebox[:, present]), name
    assert np.array_equal(area.cpu().numpy(), earea), name
