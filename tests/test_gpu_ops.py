"""GPU parity of every operator (through the C ABI) against the oracle, bit-exact
(float32 distance maps: exact as well, the tolerance 1e-5 is asserted on top)."""
import os

import numpy as np
import pytest
from scipy import ndimage as ndi

from tests._lib import (CROSS3, DISK19, o_bwareaopen2, o_bwselect, o_cc_label, o_cv_watershed, o_edt, o_imfill,
                        o_imhmin, o_local_extrema, o_morph, o_recon_binary, o_u8_unary, o_watershed2, oracle, ptr)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def G():
    from nscale_amd import gpu
    return gpu


SHAPES = [(1, 1), (3, 200), (200, 3), (64, 64), (65, 130), (257, 300)]


def rand_bin(rng, shape, p=0.5):
    return ((rng.random(shape) < p) * 255).astype(np.uint8)


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("shape", SHAPES)
def test_cc_label_and_relabel(conn, shape):
    rng = np.random.default_rng(sum(shape) + conn)
    img = (rng.integers(0, 3, shape) * 127).astype(np.uint8)  # equal-value merging matters
    print("cc_label", shape, conn, flush=True)
    lab = G().cc_label(dev(img), conn)
    exp = o_cc_label(img, conn)
    assert np.array_equal(host(lab), exp)
    n = G().cc_relabel(lab)
    e2 = exp.copy()
    ne = oracle().orc_cc_relabel(shape[1], shape[0], ptr(e2), -1)
    assert n == ne and np.array_equal(host(lab), e2)


@pytest.mark.parametrize("conn", [4, 8])
def test_cc_label_large_components(conn):
    # few huge, winding components spanning many tiles
    rng = np.random.default_rng(conn)
    img = (ndi.binary_dilation(rng.random((900, 1100)) < 0.004, iterations=6) * 255).astype(np.uint8)
    img[:, 500] = 255
    print("cc_label large", conn, flush=True)
    assert np.array_equal(host(G().cc_label(dev(img), conn)), o_cc_label(img, conn))


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("lohi", [(1, 2), (11, 1000), (30, 2 ** 31 - 1), (0, 5)])
def test_bwareaopen2(conn, lohi):
    rng = np.random.default_rng(conn + lohi[0])
    img = rand_bin(rng, (300, 411), 0.45)
    print("bwareaopen2", conn, lohi, flush=True)
    got, n = G().bwareaopen2(dev(img), lohi[0], lohi[1], conn)
    exp, ne = o_bwareaopen2(img, lohi[0], lohi[1], conn)
    assert n == ne and np.array_equal(host(got), exp)


def test_bwareaopen2_size_phantom():
    ph = np.load(os.path.join(GOLD, "size_phantom_1024.npz"))
    img = ph["img"]
    for lo, hi in ((1, 30), (30, 1000), (1000, 1 << 30)):
        got, n = G().bwareaopen2(dev(img), lo, hi, 8)
        exp, ne = o_bwareaopen2(img, lo, hi, 8)
        assert n == ne and np.array_equal(host(got), exp)


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("shape", SHAPES)
def test_binary_family(conn, shape):
    rng = np.random.default_rng(7 * sum(shape) + conn)
    mask = np.where(rng.random(shape) < 0.55, rng.integers(1, 256, shape), 0).astype(np.uint8)
    marker = np.where(rng.random(shape) < 0.05, rng.integers(1, 256, shape), 0).astype(np.uint8)
    print("binary family", shape, conn, flush=True)
    assert np.array_equal(host(G().imreconstruct_binary(dev(marker), dev(mask), conn)),
                          o_recon_binary(marker, mask, conn))
    bw = rand_bin(rng, shape, 0.5)
    seeds = rand_bin(rng, shape, 0.03)
    assert np.array_equal(host(G().bwselect(dev(bw), dev(seeds), conn)), o_bwselect(bw, seeds, conn))
    assert np.array_equal(host(G().imfill_holes(dev(bw), conn)), o_u8_unary("orc_imfill_holes_u8", bw, conn))
    assert np.array_equal(host(G().imfill(dev(bw), dev(seeds), conn)), o_imfill(bw, seeds, conn))


def test_imfill_matlab_doc_example():
    d = np.load(os.path.join(GOLD, "imfill_bw1.npz"))
    seeds = np.zeros_like(d["bw1"])
    seeds[2, 2] = 255
    assert np.array_equal(host(G().imfill(dev(d["bw1"]), dev(seeds), 8)), d["bw2"])


@pytest.mark.parametrize("kernel", ["disk19", "cross3", "rect5", "holes7", "rand3", "disk31", "holes11"])
@pytest.mark.parametrize("shape", [(1, 1), (5, 300), (130, 129), (300, 257), (67, 256)])
def test_morphology(kernel, shape):
    rng = np.random.default_rng(sum(shape))
    img = rng.integers(0, 256, shape).astype(np.uint8)
    k = {"disk19": DISK19, "cross3": CROSS3, "rect5": np.ones((5, 5), np.uint8)}.get(kernel)
    if kernel == "holes7":
        k = (np.random.default_rng(3).random((7, 7)) < 0.5).astype(np.uint8)
        k[3, 3] = 1
    if kernel == "rand3":  # 3 x 3 without its centre: the direct 3 x 3 kernel, any taps
        k = np.array([[1, 0, 1], [0, 0, 1], [1, 1, 0]], np.uint8)
    if kernel == "disk31":  # the largest element: 16-bit tables above 64 KiB of LDS
        yy, xx = np.mgrid[-15:16, -15:16]
        k = (yy * yy + xx * xx <= 15.5 ** 2).astype(np.uint8)
    if kernel == "holes11":  # many runs: too large for the 16-bit tables, byte tables above 64 KiB
        k = (np.random.default_rng(11).random((11, 11)) < 0.75).astype(np.uint8)
    print("morph", kernel, shape, flush=True)
    for outside in (0, 255, 40):
        assert np.array_equal(host(G().erode(dev(img), k, outside)), o_morph("erode", img, k, outside))
        assert np.array_equal(host(G().dilate(dev(img), k, outside)), o_morph("dilate", img, k, outside))
    assert np.array_equal(host(G().morph_open(dev(img), k)), o_morph("open", img, k))


@pytest.mark.parametrize("case", ["blobs", "sparse_zeros", "no_zero_rows", "single_zero", "big", "tallest"])
def test_distance_transform(case):
    rng = np.random.default_rng(len(case))
    if case == "blobs":
        img = (ndi.binary_dilation(rng.random((311, 277)) < 0.01, iterations=5) * 255).astype(np.uint8)
    elif case == "sparse_zeros":
        img = ((rng.random((200, 600)) > 0.0005) * 255).astype(np.uint8)
        img[0, 0] = 0
    elif case == "no_zero_rows":
        img = np.full((150, 180), 255, np.uint8)
        img[10, :] = 0
        img[100:, 40] = 0
    elif case == "single_zero":
        img = np.full((120, 1500), 255, np.uint8)  # distances beyond the search cap
        img[60, 1499] = 0
    elif case == "tallest":  # the largest vertical distance (65533): squared distances near 2^32
        img = np.full((65534, 3), 255, np.uint8)
        img[0, 0] = 0
    else:
        from nscale_amd.synth import synth_he
        bgr = synth_he(1024, 1024, 9)
        img = ((bgr[:, :, 2] < 150) * 255).astype(np.uint8)
    print("edt", case, flush=True)
    got = host(G().dist_transform(dev(img)))
    exp = o_edt(img)
    assert np.max(np.abs(got - exp)) <= 1e-5
    assert np.array_equal(got, exp)


def test_distance_transform_all_foreground():
    img = np.full((40, 50), 255, np.uint8)
    got = host(G().dist_transform(dev(img)))
    assert np.array_equal(got, o_edt(img))


def test_sqrt_is_correctly_rounded():
    # the transform evaluates sqrtf(float(dx^2) + float(g^2)); check sqrtf on every integer below
    # 2^22 through single-zero images is too slow -- check the GPU sqrt directly via torch?  No:
    # use the library on rows with one zero at column 0: value at column x is exactly x
    img = np.full((1, 4096), 255, np.uint8)
    img[0, 0] = 0
    got = host(G().dist_transform(dev(img)))
    assert np.array_equal(got[0], np.arange(4096, dtype=np.float32))


@pytest.mark.parametrize("conn", [4, 8])
def test_imhmin_and_extrema(conn):
    rng = np.random.default_rng(50 + conn)
    seg = ndi.binary_dilation(rng.random((260, 300)) < 0.01, iterations=4)
    dist = ndi.distance_transform_edt(seg).astype(np.float32)
    img = np.where(seg, -dist, -np.finfo(np.float32).max).astype(np.float32)
    print("imhmin", conn, flush=True)
    d2 = host(G().imhmin(dev(img), 1.0, conn))
    assert np.array_equal(d2, o_imhmin(img, 1.0, conn))
    assert np.array_equal(host(G().local_minima(dev(d2), conn)), o_local_extrema("minima", d2, conn))
    q = (rng.integers(0, 6, (90, 77)) * np.float32(0.5)).astype(np.float32)
    assert np.array_equal(host(G().local_maxima(dev(q), conn)), o_local_extrema("maxima", q, conn))


@pytest.mark.parametrize("seed", range(4))
def test_cv_watershed_random(seed):
    rng = np.random.default_rng(seed)
    h, w = int(rng.integers(5, 120)), int(rng.integers(5, 120))
    bgr = ndi.uniform_filter(rng.integers(0, 256, (h, w, 3)).astype(np.uint8), size=(3, 3, 1))
    m = np.where(rng.random((h, w)) < 0.02, rng.integers(1, 9, (h, w)), 0).astype(np.int32)
    m[rng.random((h, w)) < 0.05] = -1
    print("cv_watershed", seed, h, w, flush=True)
    got = host(G().cv_watershed(dev(bgr), dev(m)))
    assert np.array_equal(got, o_cv_watershed(bgr, m))


def test_watershed2_reference_image():
    from PIL import Image
    rgb = np.asarray(Image.open(os.path.join(GOLD, "watershed.png")).convert("RGB"))
    bgr = np.ascontiguousarray(rgb[:, :, ::-1])
    gray = bgr.astype(np.float32).mean(axis=2)
    seg = gray < np.percentile(gray, 40)
    dist = ndi.distance_transform_edt(seg).astype(np.float32)
    img = np.where(seg, -dist, -np.finfo(np.float32).max).astype(np.float32)
    print("watershed2 reference image", flush=True)
    got = host(G().watershed2(dev(bgr), dev(img), 8))
    assert np.array_equal(got, o_watershed2(bgr, img, 8))


def test_ccl_family_repeatable_on_chain_images():
    """Union-find under contention: the same label/area/fill results on every repetition, equal to
    the oracle, on the chain's own masks (huge background component, many nuclei)."""
    from nscale_amd.synth import synth_he
    from tests._lib import o_segment
    bgr = synth_he(1024, 1024, 2026)
    st, _, b = o_segment(bgr, stages=(21, 23))
    nonov, seg = b[21], b[23]
    exp_lab8 = o_cc_label(seg, 8)
    exp_open = o_bwareaopen2(nonov, 21, 1000, 4)
    inv = (255 - seg).astype(np.uint8)
    exp_inv4 = o_cc_label(inv, 4)
    print("ccl repeatable", flush=True)
    for rep in range(12):
        assert np.array_equal(host(G().cc_label(dev(seg), 8)), exp_lab8), rep
        assert np.array_equal(host(G().cc_label(dev(inv), 4)), exp_inv4), rep
        got, _ = G().bwareaopen2(dev(nonov), 21, 1000, 4)
        assert np.array_equal(host(got), exp_open[0] if isinstance(exp_open, tuple) else exp_open), rep
