"""CPU check of the union rules of the CCL tile pass (nscale_amd/csrc/ccl.hip, ccl_local).

The tile pass starts every foreground pixel at the first pixel of its run of left-joined pixels
and unites a vertical or diagonal join only where no join further left already links the same two
runs.  This test restates those rules on a numpy tile and checks that the components equal those
of uniting every join (brute force), for 4- and 8-connectivity, binary and multi-valued tiles (the
reference's ConnComponents::label joins neighbours of equal value, ConnComponents.cpp:52-86).
It checks the rules, not the kernel; the GPU tests compare the kernel with the oracle.
"""
import numpy as np
import pytest


def _find(par, x):
    while par[x] != x:
        par[x] = par[par[x]]
        x = par[x]
    return x


def _unite(par, a, b):
    a, b = _find(par, a), _find(par, b)
    if a != b:
        par[max(a, b)] = min(a, b)


def _labels(par, fg):
    h, w = fg.shape
    return np.array([_find(par, i) if fg.flat[i] else -1 for i in range(h * w)]).reshape(h, w)


def brute_force(val, fg, conn):
    h, w = fg.shape
    par = list(range(h * w))
    nb = [(0, -1), (-1, 0)] + ([(-1, -1), (-1, 1)] if conn == 8 else [])
    for y in range(h):
        for x in range(w):
            if not fg[y, x]:
                continue
            for dy, dx in nb:
                yy, xx = y + dy, x + dx
                if 0 <= yy < h and 0 <= xx < w and fg[yy, xx] and val[y, x] == val[yy, xx]:
                    _unite(par, y * w + x, yy * w + xx)
    return _labels(par, fg)


def tile_rules(val, fg, conn, drop=None):
    """The tile pass: run starts from the row's left joins, then the skip rules.  `drop` names a
    rule to leave out (to show that the checks notice a missing rule)."""
    h, w = fg.shape
    jl = np.zeros((h, w), bool)  # pixel x joins pixel x - 1
    jl[:, 1:] = fg[:, 1:] & fg[:, :-1] & (val[:, 1:] == val[:, :-1])
    par = list(range(h * w))
    for y in range(h):
        for x in range(w):
            if fg[y, x]:
                s = x
                while jl[y, s]:
                    s -= 1
                par[y * w + x] = y * w + s
    for y in range(1, h):
        ju = fg[y] & fg[y - 1] & (val[y] == val[y - 1])  # pixel x joins the pixel above
        for x in range(w):
            if not fg[y, x]:
                continue
            k = y * w + x
            ju_left = x > 0 and ju[x - 1]
            if drop != "up" and ju[x] and not (jl[y, x] and ju_left and jl[y - 1, x]):
                _unite(par, k, k - w)
            if conn == 8:
                jul = x > 0 and fg[y - 1, x - 1] and val[y, x] == val[y - 1, x - 1]
                jur = x < w - 1 and fg[y - 1, x + 1] and val[y, x] == val[y - 1, x + 1]
                jl_upright = x < w - 1 and jl[y - 1, x + 1]
                jl_right = x < w - 1 and jl[y, x + 1]
                ju_right = x < w - 1 and ju[x + 1]
                if drop != "up_left" and jul and not ((ju[x] and jl[y - 1, x]) or (jl[y, x] and ju_left)):
                    _unite(par, k, k - w - 1)
                if drop != "up_right" and jur and not ((ju[x] and jl_upright) or (jl_right and ju_right)):
                    _unite(par, k, k - w + 1)
    return _labels(par, fg)


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("kind", ["binary", "values", "blobs"])
def test_tile_rules_match_all_joins(conn, kind):
    rng = np.random.default_rng(7 + conn + len(kind))
    for _ in range(12):
        p = rng.uniform(0.2, 0.9)
        if kind == "binary":
            val = (rng.random((32, 64)) < p).astype(np.int32)
        elif kind == "values":
            val = rng.integers(0, 4, (32, 64)).astype(np.int32)
        else:
            val = (rng.random((8, 16)) < p).repeat(4, 0).repeat(4, 1).astype(np.int32) * int(rng.integers(1, 3))
        fg = val != 0
        assert np.array_equal(tile_rules(val, fg, conn), brute_force(val, fg, conn))


def test_exhaustive_3x4_binary():
    # every binary 3 x 4 tile, 8-connectivity: the diagonal rules' corner cases
    for m in range(1 << 12):
        val = np.array([(m >> b) & 1 for b in range(12)], np.int32).reshape(3, 4)
        fg = val != 0
        assert np.array_equal(tile_rules(val, fg, 8), brute_force(val, fg, 8)), val


@pytest.mark.parametrize("drop", ["up", "up_left", "up_right"])
def test_rule_check_catches_a_dropped_rule(drop):
    # the checks have teeth: without any one of the three unions some 3 x 4 tile is labelled wrong
    bad = 0
    for m in range(1 << 12):
        val = np.array([(m >> b) & 1 for b in range(12)], np.int32).reshape(3, 4)
        fg = val != 0
        bad += not np.array_equal(tile_rules(val, fg, 8, drop=drop), brute_force(val, fg, 8))
    assert bad > 0
