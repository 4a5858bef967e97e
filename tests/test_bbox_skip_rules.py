"""CPU check of the atomic-skip rules of the bounding-box pass (nscale_amd/csrc/ccl.hip, k_bbox).

The pass reads 64 consecutive pixels per wave, splits them into runs of equal labels (a row start
also starts a run) and issues a run's x0/x1/y0/y1 atomics only where no pixel of its component is
known to beat it.  This test restates the window, run and skip rules on numpy label images (row
wraps inside a window, windows ending mid-run, one-pixel and one-row images) and checks that the
boxes equal brute-force min/max per label -- the result the reference's
ConnComponents::boundingBox gives (ConnComponents.cpp:241-284).  A rule with a plausible mistake
(a skip that also points at the row below for x0, so that two rows can skip each other) must fail.
It checks the rules, not the kernel; the GPU tests compare the kernel with numpy.
"""
import numpy as np
import pytest

NONE = -(2 ** 31)


def brute_boxes(L, n):
    h, w = L.shape
    box = np.array([[w, -1, h, -1]] * n, np.int64)  # x0, x1, y0, y1
    for y in range(h):
        for x in range(w):
            r = L[y, x]
            if 1 <= r <= n:
                b = box[r - 1]
                b[0], b[1], b[2], b[3] = min(b[0], x), max(b[1], x), min(b[2], y), max(b[3], y)
    return box


def rule_boxes(L, n, bad_x0=False):
    """The kernel's rules, one 64-pixel window at a time (no atomics skipped that it issues)."""
    h, w = L.shape
    total = w * h
    flat = L.ravel()
    box = np.array([[w, -1, h, -1]] * n, np.int64)

    def lab(y, x):
        return int(L[y, x]) if 0 <= y < h and 0 <= x < w else NONE

    for base in range(0, total, 64):
        lanes = [i for i in range(base, min(base + 64, total))]
        pos = [(i // w, i % w) for i in lanes]
        r = [int(flat[i]) for i in lanes]
        head = [k == 0 or r[k - 1] != r[k] or pos[k][1] == 0 for k in range(len(lanes))]
        for s in range(len(lanes)):
            if not head[s] or not (1 <= r[s] <= n):
                continue
            e = s + 1
            while e < len(lanes) and not head[e]:
                e += 1
            c = r[s]
            y, x = pos[s]
            xe = pos[e - 1][1]
            up = any(lab(y - 1, pos[k][1]) == c for k in range(s, e))
            down = any(lab(y + 1, pos[k][1]) == c for k in range(s, e))
            if bad_x0:  # the pixels above and below the first pixel as witnesses: two rows can skip each other
                x0_beaten = lab(y - 1, x) == c or lab(y + 1, x) == c or any(lab(y + dy, x - 1) == c for dy in (-1, 0, 1))
            else:
                x0_beaten = lab(y - 1, x) == c or any(lab(y + dy, x - 1) == c for dy in (-1, 0, 1))
            x1_beaten = lab(y - 1, xe) == c or any(lab(y + dy, xe + 1) == c for dy in (-1, 0, 1))
            b = box[c - 1]
            if not x0_beaten:
                b[0] = min(b[0], x)
            if not x1_beaten:
                b[1] = max(b[1], xe)
            if not up:
                b[2] = min(b[2], y)
            if not down:
                b[3] = max(b[3], y)
    return box


def _images():
    rng = np.random.default_rng(17)
    for w, h in [(1, 1), (1, 70), (70, 1), (63, 9), (64, 7), (65, 8), (100, 23), (37, 31)]:
        yield (rng.random((h, w)) < 0.5).astype(np.int32) * rng.integers(1, 4, (h, w)).astype(np.int32), 3
        blobs = (rng.random(((h + 3) // 4, (w + 3) // 4)) < 0.6).repeat(4, 0).repeat(4, 1)[:h, :w]
        yield blobs.astype(np.int32) * int(rng.integers(1, 3)), 2
        yield rng.integers(0, 7, (h, w)).astype(np.int32), 6
    # discs and rings of distinct labels, some touching the borders
    yy, xx = np.mgrid[0:40, 0:130]
    L = np.zeros((40, 130), np.int32)
    for k, (cy, cx, rad) in enumerate([(10, 10, 8), (20, 60, 15), (39, 120, 12), (0, 90, 6)], 1):
        d = (yy - cy) ** 2 + (xx - cx) ** 2
        L[d <= rad * rad] = k
    L[((yy - 20) ** 2 + (xx - 60) ** 2 <= 100) & ((yy - 20) ** 2 + (xx - 60) ** 2 >= 36)] = 5
    yield L, 5


def test_skip_rules_keep_every_extreme():
    for L, n in _images():
        assert np.array_equal(rule_boxes(L, n), brute_boxes(L, n)), L.shape


def test_skip_rule_check_catches_a_cyclic_witness():
    # a vertical bar two pixels tall at the component's leftmost column: with the row below as a
    # witness as well as the row above, both rows skip their x0 atomic
    L = np.zeros((4, 6), np.int32)
    L[1:3, 1:4] = 1
    assert not np.array_equal(rule_boxes(L, 1, bad_x0=True), brute_boxes(L, 1))
    assert np.array_equal(rule_boxes(L, 1), brute_boxes(L, 1))


@pytest.mark.parametrize("seed", range(4))
def test_skip_rules_random_components(seed):
    # components of a random labelling: each label one 8-connected blob grown from a seed
    rng = np.random.default_rng(100 + seed)
    h, w = int(rng.integers(5, 30)), int(rng.integers(5, 140))
    L = np.zeros((h, w), np.int32)
    n = int(rng.integers(1, 9))
    for k in range(1, n + 1):
        y, x = int(rng.integers(0, h)), int(rng.integers(0, w))
        for _ in range(int(rng.integers(1, 60))):
            L[y, x] = k
            y = min(h - 1, max(0, y + int(rng.integers(-1, 2))))
            x = min(w - 1, max(0, x + int(rng.integers(-1, 2))))
    assert np.array_equal(rule_boxes(L, n), brute_boxes(L, n))
