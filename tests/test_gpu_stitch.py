"""Whole-slide stitching (nsc_stitch_tile) and the tile-file outputs on the device:

* exact: tiles cropped from one label image of nuclei smaller than half the overlap keep every
  nucleus exactly once, with its slide box and area, and the tiles' owned labels paste back into
  the slide's components without conflicts;
* segmented: a 3 x 3 tile region of the procedural slide, tiles segmented on their own and
  stitched, against the same region segmented as one image;
* files: the file-level segmentNuclei writes the label image as imwrite stores it (saturated to
  8 bits), equal to the oracle's labels saturated the same way; the slide driver's outputs."""
import os

import numpy as np
import pytest
from scipy import ndimage as ndi

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _blobs(h, w, seed, spacing=40, rmax=11):
    rng = np.random.default_rng(seed)
    m = np.zeros((h, w), bool)
    yy, xx = np.mgrid[0:h, 0:w]
    for cy in range(15, h - 5, spacing):
        for cx in range(15, w - 5, spacing):
            if rng.random() < 0.25:
                continue
            y, x = cy + rng.integers(-6, 7), cx + rng.integers(-6, 7)
            ry, rx = rng.integers(2, rmax + 1, 2)
            m |= ((yy - y) / ry) ** 2 + ((xx - x) / rx) ** 2 <= 1
    return m


def _canon(mask):
    lab, n = ndi.label(mask, structure=np.ones((3, 3)))
    return lab.astype(np.int32), n


@pytest.mark.parametrize("W,H,T,OV", [(1700, 1300, 512, 64), (1000, 1000, 300, 40), (777, 555, 777, 32)])
def test_cropped_labels_stitch_exactly(W, H, T, OV):
    from nscale_amd.slide import Stitcher, tile_grid
    from tests._lib import boxes_numpy
    mask = _blobs(H, W, W + H, rmax=min(11, OV // 4 - 1))
    glab, gn = _canon(mask)
    gbox, garea = boxes_numpy(glab, bg=0)
    expect = {(int(gbox[1, i]), int(gbox[2, i]), int(gbox[3, i]), int(gbox[4, i]), int(garea[i])) for i in range(gn)}
    assert max(int((gbox[2] - gbox[1]).max()), int((gbox[4] - gbox[3]).max())) < OV // 2
    st = Stitcher(W, H, OV)
    slide = np.zeros((H, W), np.int64)
    for (x0, y0, w, h) in tile_grid(W, H, T, OV):
        lab, n = _canon(mask[y0:y0 + h, x0:x0 + w])
        box, area = boxes_numpy(lab, bg=0)
        if n == 0:
            box = np.zeros((5, 0), np.int32)
        assert n == 0 or (box[0] == np.arange(1, n + 1)).all()
        ids, own = st.add(x0, y0, torch.from_numpy(lab).cuda(), torch.from_numpy(np.ascontiguousarray(box)).cuda(),
                          torch.from_numpy(area).cuda() if n else None, owned=True)
        own = own.cpu().numpy()
        reg = slide[y0:y0 + h, x0:x0 + w]
        clash = (reg > 0) & (own > 0) & (reg != own)
        assert not clash.any()
        reg[own > 0] = own[own > 0]
    t = st.nuclei()
    print("slide", W, H, "tiles", len(tile_grid(W, H, T, OV)), "nuclei", gn, "kept", len(t), flush=True)
    assert st.truncated == 0
    assert np.array_equal(t[:, 0], np.arange(1, len(t) + 1))
    got = {tuple(int(v) for v in r[1:6]) for r in t}
    assert len(got) == len(t) == gn and got == expect
    # the pasted owned labels are the slide's components, one id each
    assert np.array_equal(slide > 0, glab > 0)
    pairs = np.unique(np.stack([glab[glab > 0], slide[glab > 0]], 1), axis=0)
    assert len(pairs) == gn and len(np.unique(pairs[:, 1])) == gn


def test_single_tile_keeps_everything_and_empty_tiles():
    from nscale_amd.slide import Stitcher
    from tests._lib import boxes_numpy
    lab, n = _canon(_blobs(300, 200, 3))
    box, area = boxes_numpy(lab, bg=0)
    st = Stitcher(200, 300, 64)
    ids, _ = st.add(0, 0, torch.from_numpy(lab).cuda(), torch.from_numpy(np.ascontiguousarray(box)).cuda())
    assert np.array_equal(ids.cpu().numpy(), np.arange(1, n + 1))
    st2 = Stitcher(200, 300, 64)
    ids2, own = st2.add(0, 0, torch.zeros(300, 200, dtype=torch.int32, device="cuda"),
                        torch.zeros(5, 0, dtype=torch.int32, device="cuda"), owned=True)
    assert ids2.numel() == 0 and not own.any() and st2.next_id == 1


def test_segmented_tiles_stitch_like_the_whole_region():
    """The chain is not translation invariant (reconstructions and hole filling see the tile border),
    so tiles do not reproduce the whole region's nuclei exactly; the stitched set must still count
    the same nuclei, claim no pixel twice, cut none, and cover nearly the same pixels."""
    from nscale_amd import gpu
    from nscale_amd.slide import segment_slide, tile_grid
    OX, OY, T, OV = 31000, 27000, 1024, 64
    W = H = 3 * T - 2 * OV
    assert len(tile_grid(W, H, T, OV)) == 9

    def src(x0, y0, w, h, out):
        gpu.synth_slide_tile(OX + x0, OY + y0, w, h, 2026, out.view(h, w, 3))

    raster = torch.zeros(H, W, dtype=torch.int32, device="cuda")
    st, rep = segment_slide(src, W, H, T, OV, slots=3, budget_gb=16, raster=raster)
    assert all(r[2] == 0 for r in rep), rep
    t = st.nuclei()
    whole = gpu.synth_slide_tile(OX, OY, W, H, 2026)
    s, lab, cc, nuc, box, area, _ = gpu.segment_nuclei_io(whole, canonical=True, bbox_mode=gpu.BBOX_ALL,
                                                          max_boxes=1 << 16, with_area=True)
    torch.cuda.synchronize()
    ref = lab.cpu().numpy()
    got = raster.cpu().numpy()
    fg_agree = float(((ref > 0) == (got > 0)).mean())
    # nuclei of the whole region whose pixels are exactly one stitched nucleus
    pairs, cnt = np.unique(np.stack([ref[ref > 0], got[ref > 0]], 1), axis=0, return_counts=True)
    ref_area = np.bincount(ref.ravel())
    got_area = np.bincount(got.ravel())
    exact = sum(1 for (a, b), c in zip(pairs, cnt) if b > 0 and c == ref_area[a] == got_area[b])
    print("whole region nuclei", cc, "stitched", len(t), "exactly equal", exact, "foreground agreement %.5f" % fg_agree,
          "cut", st.truncated, "conflicts", st.conflicts, flush=True)
    assert st.conflicts == 0 and st.truncated == 0
    assert abs(len(t) - cc) <= 0.02 * cc
    assert len(np.unique(got)) - 1 == len(t)
    assert fg_agree >= 0.99 and exact >= 0.5 * cc


def test_file_segmentation_writes_the_saturated_labels(tmp_path):
    from nscale_amd import tiles
    from nscale_amd.synth import synth_he
    from tests._lib import o_segment_bbox
    bgr = synth_he(384, 512, 41)
    tiles.write_png(tmp_path / "t-0-0.png", bgr)
    st, cc, bbox = tiles.segment_file(str(tmp_path / "t-0-0.png"), str(tmp_path / "t-0-0.mask.png"))
    ost, olab, occ, obox = o_segment_bbox(bgr)
    assert st == ost == 0 and cc == occ
    assert (bbox is None and occ == 0) or np.array_equal(bbox, obox)
    assert (olab > 255).any() and (olab == -1).any()  # both saturations occur
    got = tiles.read_png(tmp_path / "t-0-0.mask.png")[..., 0]
    assert np.array_equal(got, tiles.mask_of_labels(olab))
    # unreadable input: the reference's INVALID_IMAGE, nothing written
    (tmp_path / "bad.png").write_bytes(b"xx")
    assert tiles.segment_file(str(tmp_path / "bad.png"), str(tmp_path / "bad.mask.png"))[0] == -1
    assert not os.path.exists(tmp_path / "bad.mask.png")
    # directory mode over the tiles alone
    d = tmp_path / "in"
    d.mkdir()
    os.replace(tmp_path / "t-0-0.png", d / "t-0-0.png")
    os.replace(tmp_path / "bad.png", d / "bad.png")
    res = tiles.segment_directory(str(d), str(tmp_path / "out"))
    assert [(os.path.basename(r[0]), r[1]) for r in res] == [("bad.png", -1), ("t-0-0.png", 0)]
    assert np.array_equal(tiles.read_png(tmp_path / "out" / "t-0-0.mask.png")[..., 0], tiles.mask_of_labels(olab))


def test_tiff_tiles_in_directory_mode(tmp_path):
    """The reference's .tif tiles (decoded through PIL when installed) take the same chain."""
    pil = pytest.importorskip("PIL.Image")
    from nscale_amd import tiles
    from nscale_amd.synth import synth_he
    from tests._lib import o_segment
    bgr = synth_he(300, 420, 43)
    d = tmp_path / "in"
    d.mkdir()
    pil.fromarray(bgr[..., ::-1].copy()).save(d / "s-0-0.tif")
    (d / "broken.tiff").write_bytes(b"II*\x00junk")
    assert np.array_equal(tiles.read_image(d / "s-0-0.tif"), bgr)
    res = tiles.segment_directory(str(d), str(tmp_path / "out"))
    assert [(os.path.basename(r[0]), r[1]) for r in res] == [("broken.tiff", -1), ("s-0-0.tif", 0)]
    ost, olab, _ = o_segment(bgr)
    assert ost == 0
    assert np.array_equal(tiles.read_png(tmp_path / "out" / "s-0-0.mask.png")[..., 0], tiles.mask_of_labels(olab))


def test_slide_driver_outputs(tmp_path):
    from nscale_amd import gpu, tiles
    from nscale_amd.slide import segment_slide, tile_grid

    def src(x0, y0, w, h, out):
        gpu.synth_slide_tile(5000 + x0, 9000 + y0, w, h, 2026, out.view(h, w, 3))

    st, rep = segment_slide(src, 1500, 900, 512, 64, name="s1", out_dir=str(tmp_path), slots=2, budget_gb=8)
    names = sorted(f for f in os.listdir(tmp_path) if f.endswith(".mask.png"))
    assert names == sorted("s1-%d-%d.mask.png" % (x, y) for x, y, _, _ in tile_grid(1500, 900, 512, 64))
    rows = open(tmp_path / "s1.nuclei.csv").read().splitlines()
    assert rows[0] == "id,minx,maxx,miny,maxy,area,tile_x,tile_y" and len(rows) - 1 == len(st.nuclei()) > 0
    # each mask is the reference's label image of that tile, saturated to 8 bits; the stitched set
    # equals the one the driver finds without writing files
    for x, y, w, h in tile_grid(1500, 900, 512, 64)[:3]:
        tile = gpu.synth_slide_tile(5000 + x, 9000 + y, w, h, 2026)
        s0, lab, _, _, _ = gpu.segment_nuclei(tile)
        torch.cuda.synchronize()
        m = tiles.read_png(tmp_path / ("s1-%d-%d.mask.png" % (x, y)))[..., 0]
        assert np.array_equal(m, tiles.mask_of_labels(lab.cpu().numpy()))
    st2, _ = segment_slide(src, 1500, 900, 512, 64, slots=2, budget_gb=8)
    assert np.array_equal(st.nuclei(), st2.nuclei())
