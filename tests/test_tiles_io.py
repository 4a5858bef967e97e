"""PNG tile files on the host (no GPU): the library's codec against PIL in both directions for
every colour type it reads, the imwrite conversion of label images, and nu-segment's output names."""
import os

import numpy as np
import pytest

from nscale_amd import tiles

PIL = pytest.importorskip("PIL.Image")


def test_write_then_pil_reads_the_same_pixels(tmp_path):
    rng = np.random.default_rng(1)
    bgr = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    grey = rng.integers(0, 256, (19, 7), dtype=np.uint8)
    tiles.write_png(tmp_path / "c.png", bgr)
    tiles.write_png(tmp_path / "g.png", grey)
    rgb = np.asarray(PIL.open(tmp_path / "c.png"))
    assert np.array_equal(rgb[..., ::-1], bgr)
    assert np.array_equal(np.asarray(PIL.open(tmp_path / "g.png")), grey)
    assert np.array_equal(tiles.read_png(tmp_path / "c.png"), bgr)
    assert np.array_equal(tiles.read_png(tmp_path / "g.png"), np.repeat(grey[..., None], 3, 2))


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "I;16", "1", "P4"])
def test_read_what_pil_writes(tmp_path, mode):
    """Grey, grey+alpha, RGB, RGBA, palette and 16-bit grey, with PIL's adaptive scanline filters."""
    rng = np.random.default_rng(2)
    h, w = 41, 67
    if mode in ("RGB", "RGBA"):
        arr = rng.integers(0, 256, (h, w, len(mode)), dtype=np.uint8)
        img = PIL.fromarray(arr, mode)
        exp = arr[..., 2::-1]
    elif mode in ("L", "LA"):
        arr = rng.integers(0, 256, (h, w, len(mode)), dtype=np.uint8)
        img = PIL.fromarray(arr if mode == "LA" else arr[..., 0], mode)
        exp = np.repeat(arr[..., :1], 3, 2)
    elif mode == "1":
        bits = rng.integers(0, 2, (h, w), dtype=np.uint8).astype(bool)
        img = PIL.fromarray(bits)
        exp = np.repeat((bits * 255).astype(np.uint8)[..., None], 3, 2)
    elif mode in ("P", "P4"):
        idx = rng.integers(0, 16 if mode == "P4" else 200, (h, w), dtype=np.uint8)
        pal = rng.integers(0, 256, (256, 3), dtype=np.uint8)
        img = PIL.fromarray(idx, "P")
        img.putpalette(pal.ravel().tolist())
        exp = pal[idx][..., ::-1]
    else:
        arr = rng.integers(0, 65536, (h, w), dtype=np.uint16)
        img = PIL.fromarray(arr)
        exp = np.repeat((arr >> 8).astype(np.uint8)[..., None], 3, 2)  # 16-bit keeps the high byte
    p = tmp_path / ("x_%s.png" % mode.replace(";", ""))
    img.save(p, optimize=True, **({"bits": 4} if mode == "P4" else {}))
    assert np.array_equal(tiles.read_png(p), exp)


def test_unreadable_files_are_errors(tmp_path):
    from nscale_amd._lib import NscError
    (tmp_path / "bad.png").write_bytes(b"not a png")
    with pytest.raises(NscError):
        tiles.read_png(tmp_path / "bad.png")
    with pytest.raises(NscError):
        tiles.read_png(tmp_path / "missing.png")


def test_label_image_conversion_is_saturation():
    lab = np.array([[-1, 0, 1, 254], [255, 256, 70000, -5]], np.int32)
    assert tiles.mask_of_labels(lab).tolist() == [[0, 0, 1, 254], [255, 255, 255, 0]]


def test_output_names_follow_get_files(tmp_path):
    d = tmp_path / "in" / "sub"
    d.mkdir(parents=True)
    for f in ("a-0-0.tif", "b.png", "c.txt"):
        (d / f).write_bytes(b"")
    imgs = tiles.list_images(str(tmp_path / "in"))
    assert [os.path.basename(i) for i in imgs] == ["a-0-0.tif", "b.png"]
    mask, bounds = tiles.output_names(imgs[0], str(tmp_path / "in"), str(tmp_path / "out"))
    assert mask == str(tmp_path / "out" / "sub" / "a-0-0.mask.png")
    assert bounds == str(tmp_path / "out" / "sub" / "a-0-0.bounds.csv")
    assert os.path.isdir(tmp_path / "out" / "sub")
