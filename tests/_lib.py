"""ctypes access to the CPU oracle (test infrastructure), the reference's own ConnComponents
build (oracle/_ref, when present) and numpy helpers shared by the tests."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_P = ctypes.c_void_p


def ptr(a):
    return a.ctypes.data_as(_P)


class SegParams(ctypes.Structure):
    _fields_ = [("blue", ctypes.c_uint8), ("green", ctypes.c_uint8), ("red", ctypes.c_uint8),
                ("T1", ctypes.c_double), ("T2", ctypes.c_double), ("G1", ctypes.c_uint8),
                ("minSize", ctypes.c_int), ("maxSize", ctypes.c_int), ("G2", ctypes.c_uint8),
                ("minSizePl", ctypes.c_int), ("minSizeSeg", ctypes.c_int), ("maxSizeSeg", ctypes.c_int),
                ("fillHolesConnectivity", ctypes.c_int), ("reconConnectivity", ctypes.c_int),
                ("watershedConnectivity", ctypes.c_int)]


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        _oracle = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        _oracle.orc_cc_relabel.restype = ctypes.c_int
        _oracle.orc_bwareaopen2_u8.restype = ctypes.c_int
        _oracle.orc_cc_area_threshold_labeled.restype = ctypes.c_int
        _oracle.orc_cc_bounding_box_literal.restype = ctypes.c_int
        _oracle.orc_segment_nuclei.restype = ctypes.c_int
        _oracle.orc_segment_nuclei_bbox.restype = ctypes.c_int
        _oracle.orc_norm_target_parameters.restype = ctypes.c_longlong
        _oracle.orc_feat_area.argtypes = [_P, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, _P]
        _oracle.orc_feat_ellipse.argtypes = [_P, _P, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, _P, _P, _P]
        _oracle.orc_feat_perimeter.argtypes = [_P, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, _P]
        _oracle.orc_feat_extent_circ.argtypes = [_P, _P, _P, ctypes.c_int, _P, _P]
        _oracle.orc_feat_intensity.argtypes = [_P, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int, _P]
        _oracle.orc_feat_gradient.argtypes = [_P, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int, _P]
        _oracle.orc_gradient_u8.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int]
        _oracle.orc_imhmin_f32.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int]
        _oracle.orc_get_background.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_uint8, ctypes.c_uint8, ctypes.c_uint8]
        _oracle.orc_get_rbc.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    return _oracle


def ref_cc():
    """The reference's ConnComponents.cpp compiled from its sources (None if not built)."""
    p = os.path.join(ROOT, "oracle", "_ref", "libcc_ref.so")
    if not os.path.exists(p):
        return None
    lib = ctypes.CDLL(p)
    for f in ("ref_cc_relabel", "ref_cc_area_threshold_labeled", "ref_cc_area_threshold", "ref_cc_bounding_box"):
        getattr(lib, f).restype = ctypes.c_int
    return lib


# ---------------------------------------------------------------- oracle wrappers (numpy)
def o_recon(marker, mask, conn):
    h, w = mask.shape
    out = np.empty_like(mask)
    fn = {np.uint8: "orc_imreconstruct_u8", np.float32: "orc_imreconstruct_f32",
          np.int32: "orc_imreconstruct_i32"}[mask.dtype.type]
    getattr(oracle(), fn)(ptr(np.ascontiguousarray(marker)), ptr(np.ascontiguousarray(mask)), ptr(out), w, h, conn)
    return out


def o_recon_binary(marker, mask, conn):
    h, w = mask.shape
    out = np.empty_like(mask)
    oracle().orc_imreconstruct_binary_u8(ptr(marker), ptr(mask), ptr(out), w, h, conn)
    return out


def o_u8_unary(name, img, *args):
    h, w = img.shape[:2]
    out = np.empty((h, w), np.uint8)
    getattr(oracle(), name)(ptr(np.ascontiguousarray(img)), ptr(out), w, h, *args)
    return out


def o_imfill(img, seeds, conn):
    h, w = img.shape
    out = np.empty_like(img)
    oracle().orc_imfill_u8(ptr(img), ptr(seeds), ptr(out), w, h, conn)
    return out


def o_bwselect(bw, seeds, conn):
    h, w = bw.shape
    out = np.empty_like(bw)
    oracle().orc_bwselect_u8(ptr(bw), ptr(seeds), ptr(out), w, h, conn)
    return out


def o_cc_label(img, conn, bgval=-1):
    h, w = img.shape
    lab = np.empty((h, w), np.int32)
    oracle().orc_cc_label(ptr(np.ascontiguousarray(img)), w, h, ptr(lab), bgval, conn)
    return lab


def o_bwareaopen2(img, lo, hi, conn):
    h, w = img.shape
    out = np.empty_like(img)
    n = oracle().orc_bwareaopen2_u8(ptr(img), ptr(out), w, h, lo, hi, conn)
    return out, n


def o_morph(kind, img, kernel, outside=None):
    h, w = img.shape
    k = kernel.shape[0]
    out = np.empty_like(img)
    kern = np.ascontiguousarray(kernel.astype(np.uint8))
    if kind == "open":
        oracle().orc_morph_open_u8(ptr(img), ptr(out), w, h, ptr(kern), k)
    else:
        getattr(oracle(), "orc_%s_u8" % kind)(ptr(img), ptr(out), w, h, ptr(kern), k, outside)
    return out


def o_edt(img):
    h, w = img.shape
    out = np.empty((h, w), np.float32)
    oracle().orc_dist_transform_u8(ptr(np.ascontiguousarray(img)), ptr(out), w, h)
    return out


def o_imhmin(img, hval, conn):
    h, w = img.shape
    out = np.empty_like(img)
    oracle().orc_imhmin_f32(ptr(img), ptr(out), w, h, hval, conn)
    return out


def o_local_extrema(kind, img, conn):
    h, w = img.shape
    out = np.empty((h, w), np.uint8)
    getattr(oracle(), "orc_local_%s_f32" % kind)(ptr(np.ascontiguousarray(img)), ptr(out), w, h, conn)
    return out


def o_cv_watershed(bgr, markers):
    h, w = markers.shape
    m = np.ascontiguousarray(markers.astype(np.int32)).copy()
    oracle().orc_cv_watershed(ptr(np.ascontiguousarray(bgr)), ptr(m), w, h)
    return m


def o_watershed2(bgr, img, conn):
    h, w = img.shape
    out = np.empty((h, w), np.int32)
    oracle().orc_watershed2(ptr(np.ascontiguousarray(bgr)), ptr(np.ascontiguousarray(img)), ptr(out), w, h, conn)
    return out


def o_norm_fg(bgr):
    h, w = bgr.shape[:2]
    out = np.zeros((h, w), np.uint8)
    oracle().orc_norm_seg_fg(ptr(np.ascontiguousarray(bgr)), ptr(out), w, h)
    return out


def o_norm_lab(bgr):
    h, w = bgr.shape[:2]
    lab = np.zeros((h, w, 3), np.float32)
    oracle().orc_norm_bgr2lab(ptr(np.ascontiguousarray(bgr)), ptr(lab), w, h)
    return lab


def o_norm_bgr(lab):
    h, w = lab.shape[:2]
    out = np.zeros((h, w, 3), np.uint8)
    oracle().orc_norm_lab2bgr(ptr(np.ascontiguousarray(lab, np.float32)), ptr(out), w, h)
    return out


def o_norm_target(bgr):
    h, w = bgr.shape[:2]
    m, s = np.zeros(3, np.float32), np.zeros(3, np.float32)
    n = oracle().orc_norm_target_parameters(ptr(np.ascontiguousarray(bgr)), w, h, ptr(m), ptr(s))
    return m, s, n


def o_normalize(bgr, mean, std):
    h, w = bgr.shape[:2]
    out = np.zeros((h, w, 3), np.uint8)
    oracle().orc_norm_normalization(ptr(np.ascontiguousarray(bgr)), w, h, ptr(np.asarray(mean, np.float32)),
                                    ptr(np.asarray(std, np.float32)), ptr(out))
    return out


def all_colours():
    """4096 x 4096 BGR image holding every 24-bit colour once (b = low byte of the pixel index)."""
    i = np.arange(1 << 24, dtype=np.uint32)
    return np.stack([i & 255, (i >> 8) & 255, i >> 16], -1).astype(np.uint8).reshape(4096, 4096, 3)


def default_params():
    p = SegParams()
    oracle().orc_seg_params_default(ctypes.byref(p))
    return p


STAGES = {1: (np.uint8, 1), 2: (np.uint8, 1), 3: (np.uint8, 1), 4: (np.uint8, 1), 5: (np.uint8, 1),
          6: (np.uint8, 1), 7: (np.uint8, 1), 8: (np.uint8, 1), 9: (np.uint8, 1), 10: (np.uint8, 1),
          11: (np.uint8, 1), 12: (np.uint8, 1), 13: (np.uint8, 1), 14: (np.uint8, 1), 15: (np.uint8, 1),
          16: (np.float32, 1), 17: (np.float32, 1), 18: (np.float32, 1), 19: (np.uint8, 3),
          20: (np.int32, 1), 21: (np.uint8, 1), 23: (np.uint8, 1), 24: (np.uint8, 1)}


def o_segment(bgr, params=None, stages=()):
    """Run the oracle chain; returns (status, labels, {stage_id: array})."""
    h, w = bgr.shape[:2]
    p = params or default_params()
    lab = np.full((h, w), -7, np.int32)
    arr = (ctypes.c_void_p * 32)()
    bufs = {}
    for s in stages:
        dt, ch = STAGES[s]
        bufs[s] = np.zeros((h, w, ch) if ch > 1 else (h, w), dt)
        arr[s] = bufs[s].ctypes.data
    st = oracle().orc_segment_nuclei(ptr(np.ascontiguousarray(bgr)), w, h, ctypes.byref(p), ptr(lab), arr, None)
    return st, lab, bufs


def o_segment_bbox(bgr, params=None, cap=1 << 16):
    """Oracle chain plus its compcount and bbox outputs exactly as HistologicalEntities.cpp:589-591
    computes them (ConnComponents::boundingBox(labels, bgval 0)): (status, labels, compcount, bbox
    as a 5 x compcount array)."""
    h, w = bgr.shape[:2]
    p = params or default_params()
    lab = np.full((h, w), -7, np.int32)
    cc = ctypes.c_int(0)
    box = np.zeros(5 * cap, np.int32)
    st = oracle().orc_segment_nuclei_bbox(ptr(np.ascontiguousarray(bgr)), w, h, ctypes.byref(p), ptr(lab),
                                          ctypes.byref(cc), ptr(box), cap, None, None)
    n = cc.value
    return st, lab, n, box[:5 * n].reshape(5, n)


def ref_test_params():
    """The parameter set of the reference's own chain test, nu-segmentTest.cpp:207:
    segmentNuclei(img, mask, 220, 220, 220, 4.0, 3.5, 44, 5, 9000, 31, 20, 5, 1150, 4, 4, 4)."""
    p = SegParams()
    p.blue, p.green, p.red = 220, 220, 220
    p.T1, p.T2 = 4.0, 3.5
    p.G1, p.minSize, p.maxSize = 44, 5, 9000
    p.G2, p.minSizePl, p.minSizeSeg, p.maxSizeSeg = 31, 20, 5, 1150
    p.fillHolesConnectivity, p.reconConnectivity, p.watershedConnectivity = 4, 4, 4
    return p


def boxes_numpy(lab, bg=-1):
    """[label | minx | maxx | miny | maxy] (5 x n) and areas of every label != bg, ascending label order."""
    ys, xs = np.nonzero(lab != bg)
    v = lab[ys, xs]
    labels, inv, area = np.unique(v, return_inverse=True, return_counts=True)
    n = len(labels)
    out = np.zeros((5, n), np.int64)
    out[0] = labels
    out[1] = np.full(n, 1 << 40); np.minimum.at(out[1], inv, xs)
    out[2] = np.full(n, -1); np.maximum.at(out[2], inv, xs)
    out[3] = np.full(n, 1 << 40); np.minimum.at(out[3], inv, ys)
    out[4] = np.full(n, -1); np.maximum.at(out[4], inv, ys)
    return out.astype(np.int32), area.astype(np.int32)


# ---------------------------------------------------------------- structuring elements
DISK19 = np.array([[0, 0, 0, 0] + [1] * 11 + [0, 0, 0, 0],
                   [0, 0, 0] + [1] * 13 + [0, 0, 0],
                   [0, 0] + [1] * 15 + [0, 0],
                   [0] + [1] * 17 + [0]] + [[1] * 19] * 11 +
                  [[0] + [1] * 17 + [0],
                   [0, 0] + [1] * 15 + [0, 0],
                   [0, 0, 0] + [1] * 13 + [0, 0, 0],
                   [0, 0, 0, 0] + [1] * 11 + [0, 0, 0, 0]], np.uint8)
CROSS3 = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], np.uint8)
assert DISK19.shape == (19, 19)


def structure(conn):
    return np.ones((3, 3), bool) if conn == 8 else CROSS3.astype(bool)


# ---------------------------------------------------------------- long propagation paths
def serpentine(h=124, w=62, top=50, bottom=70, value=200):
    """A one-pixel path of vertical segments at the odd columns, rows top..bottom, joined alternately
    at the top and bottom rows: with 62x62 tiles every segment crosses the seam at row 62, so the
    front needs about one pass per segment.  Returns (marker, mask, expected reconstruction): the
    path is one component holding the marker value, so the whole path takes it."""
    mask = np.zeros((h, w), np.uint8)
    xs = list(range(1, w, 2))
    for k, x in enumerate(xs):
        mask[top:bottom + 1, x] = value
        if k + 1 < len(xs):
            mask[top if k % 2 == 0 else bottom, x + 1] = value
    marker = np.zeros_like(mask)
    marker[bottom, xs[0]] = value
    return marker, mask, mask.copy()


def boustrophedon(n=1024, value=200):
    """A one-pixel maze: corridors along the even rows joined alternately at the right and left ends,
    one path of about n*n/2 pixels that crosses every vertical tile seam once per corridor.
    Returns (marker, mask, expected reconstruction)."""
    mask = np.zeros((n, n), np.uint8)
    mask[0::2, :] = value
    for k, y in enumerate(range(1, n - 1, 2)):
        mask[y, n - 1 if k % 2 == 0 else 0] = value
    marker = np.zeros_like(mask)
    marker[0, 0] = value
    return marker, mask, mask.copy()
