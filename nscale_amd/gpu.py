"""Operator surface of nscale:: / nscale::gpu:: (MorphologicOperations.h, HistologicalEntities.h)
over torch device tensors, calling libnscale_hip.so through its C ABI on torch's current stream.

torch is used for device memory and streams only; every operator runs the library's HIP kernels.
"""
import ctypes

import numpy as np

import torch

from nscale_amd._lib import (BBOX_ALL, BBOX_NONE, BBOX_REFERENCE_CPU, ObjFeaturesOut, ReconDesc, Request, Result,
                             SegIO, SegParams, ServerConfig, call, check, lib)

_P = ctypes.c_void_p


def _ptr(t):
    assert t.is_cuda and t.is_contiguous(), "expects a contiguous device tensor"
    return _P(t.data_ptr())


def _rows(t):
    """Device pointer and row step in bytes of a 2-D (h, w) or BGR (h, w, 3) tensor whose rows are
    contiguous but may be spaced further apart (a pitched view, like a GpuMat's step)."""
    assert t.is_cuda
    es = t.element_size()
    if t.dim() == 3:
        assert t.stride(2) == 1 and t.stride(1) == t.shape[2], "BGR pixels must be interleaved"
    else:
        assert t.stride(1) == 1, "rows must be contiguous"
    return _P(t.data_ptr()), int(t.stride(0) * es)


def _stream():
    return _P(torch.cuda.current_stream().cuda_stream)


def _hw(t):
    return int(t.shape[0]), int(t.shape[1])


def imreconstruct(marker, mask, conn=8):
    """nscale::imreconstruct<T>(seeds, image, conn) for uint8 / float32 / int32 planes."""
    h, w = _hw(mask)
    out = torch.empty_like(mask)
    fn = {torch.uint8: "nsc_imreconstruct_u8", torch.float32: "nsc_imreconstruct_f32",
          torch.int32: "nsc_imreconstruct_i32"}[mask.dtype]
    call(fn, _ptr(marker), _ptr(mask), _ptr(out), w, h, conn, _stream())
    return out


def imreconstruct_batched(pairs, conns):
    """n independent reconstructions (uint8 or float32 (marker, mask) pairs of any sizes) in one
    launch per pass; conns: one connectivity per pair.  Returns the outputs."""
    n = len(pairs)
    outs = [torch.empty_like(mask) for _, mask in pairs]
    arr = (ReconDesc * n)()
    dt = pairs[0][1].dtype
    for i, ((marker, mask), o) in enumerate(zip(pairs, outs)):
        assert mask.dtype == dt and marker.shape == mask.shape
        h, w = _hw(mask)
        arr[i] = ReconDesc(_ptr(marker), _ptr(mask), _ptr(o), w, h, int(conns[i]))
    fn = {torch.uint8: "nsc_imreconstruct_u8_batched", torch.float32: "nsc_imreconstruct_f32_batched"}[dt]
    call(fn, arr, n, _stream())
    return outs


class ReconGraph:
    """A batched reconstruction captured as a hipGraph of `passes` window passes per replay: the
    (marker, mask) pairs and the outputs are fixed at creation (their contents may change), run()
    recomputes every output (replaying the graph until no tile is left marked) and returns the
    number of passes launched."""

    def __init__(self, pairs, conns, passes=16):
        n = len(pairs)
        self.dtype = pairs[0][1].dtype
        self.pairs = pairs
        self.outs = [torch.empty_like(mask) for _, mask in pairs]
        arr = (ReconDesc * n)()
        for i, ((marker, mask), o) in enumerate(zip(pairs, self.outs)):
            assert mask.dtype == self.dtype and marker.shape == mask.shape
            h, w = _hw(mask)
            arr[i] = ReconDesc(_ptr(marker), _ptr(mask), _ptr(o), w, h, int(conns[i]))
        self.handle = _P()
        call("nsc_recon_graph_create", arr, n, 1 if self.dtype == torch.float32 else 0, int(passes),
             ctypes.byref(self.handle))

    def run(self):
        k = ctypes.c_int(0)
        call("nsc_recon_graph_run", self.handle, ctypes.byref(k), _stream())
        return k.value

    def close(self):
        if self.handle:
            lib().nsc_recon_graph_destroy(self.handle)
            self.handle = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_recon(pairs, conns):
    """Pick the reconstruction schedule (windows per workgroup side, passes per host check) by timing
    each candidate on this batch; later calls of the same element type use the choice.
    Returns {"region": G, "passes": P, "timings_ms": [...]} (candidates (1,4), (1,8), (2,4), (2,8) and the
    device-side work queue, reported as passes 0)."""
    n = len(pairs)
    arr = (ReconDesc * n)()
    outs = [torch.empty_like(mask) for _, mask in pairs]
    for i, ((marker, mask), o) in enumerate(zip(pairs, outs)):
        h, w = _hw(mask)
        arr[i] = ReconDesc(_ptr(marker), _ptr(mask), _ptr(o), w, h, int(conns[i]))
    g, p = ctypes.c_int(0), ctypes.c_int(0)
    t = (ctypes.c_float * 5)()
    call("nsc_recon_autotune", arr, n, 1 if pairs[0][1].dtype == torch.float32 else 0, _stream(), ctypes.byref(g),
         ctypes.byref(p), t)
    return {"region": g.value, "passes": p.value, "timings_ms": [round(x, 3) for x in t]}


def set_recon_schedule(is_float=False, region=0, passes=0, queue=-1):
    """Sets the calling thread's reconstruction schedule (see nsc_recon_set_schedule)."""
    call("nsc_recon_set_schedule", 1 if is_float else 0, int(region), int(passes), int(queue))


def imhmin_batched(imgs, hval, conns):
    """nscale::imhmin<float> on n float32 images in one batch."""
    n = len(imgs)
    outs = [torch.empty_like(im) for im in imgs]
    arr = (ReconDesc * n)()
    for i, (im, o) in enumerate(zip(imgs, outs)):
        h, w = _hw(im)
        arr[i] = ReconDesc(None, _ptr(im), _ptr(o), w, h, int(conns[i]))
    call("nsc_imhmin_f32_batched", arr, n, ctypes.c_float(hval), _stream())
    return outs


def synth_slide_tile(x0, y0, w, h, seed=2026, out=None):
    """Tile (x0, y0, w, h) of the procedural whole-slide H&E image, rendered on the device (BGR)."""
    if out is None:
        out = torch.empty(h, w, 3, dtype=torch.uint8, device="cuda")
    call("nsc_synth_slide_tile", _ptr(out), int(x0), int(y0), int(w), int(h), int(seed), _stream())
    return out


def _u8out(t):
    return torch.empty(t.shape[0], t.shape[1], dtype=torch.uint8, device=t.device)


def _i32out(t):
    return torch.empty(t.shape[0], t.shape[1], dtype=torch.int32, device=t.device)


def imreconstruct_binary(marker, mask, conn=8):
    """nscale::imreconstructBinary<uchar>(seeds, image, conn)."""
    h, w = _hw(mask)
    out = _u8out(mask)
    call("nsc_imreconstruct_binary_u8", _ptr(marker), _ptr(mask), _ptr(out), w, h, conn, _stream())
    return out


def imfill(img, seeds, conn=8):
    """nscale::imfill<uchar>(image, seeds, binary=true, conn)."""
    h, w = _hw(img)
    out = _u8out(img)
    call("nsc_imfill_u8", _ptr(img), _ptr(seeds), _ptr(out), w, h, conn, _stream())
    return out


def imfill_holes(img, conn=4):
    """nscale::imfillHoles<uchar>(image, binary=true, conn)."""
    h, w = _hw(img)
    out = _u8out(img)
    call("nsc_imfill_holes_u8", _ptr(img), _ptr(out), w, h, conn, _stream())
    return out


def bwselect(bw, seeds, conn=8):
    """nscale::bwselect<uchar>(binaryImage, seeds, conn)."""
    h, w = _hw(bw)
    out = _u8out(bw)
    call("nsc_bwselect_u8", _ptr(bw), _ptr(seeds), _ptr(out), w, h, conn, _stream())
    return out


def imhmin(img, hval=1.0, conn=8):
    """nscale::imhmin<float>(image, h, conn)."""
    h, w = _hw(img)
    out = torch.empty_like(img)
    lib().nsc_imhmin_f32.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int, _P]
    call("nsc_imhmin_f32", _ptr(img), _ptr(out), w, h, float(hval), conn, _stream())
    return out


def cc_label(img, conn=8, bgval=-1):
    """ConnComponents::label (equal-value union-find, root = smallest raster index)."""
    h, w = _hw(img)
    out = _i32out(img)
    call("nsc_cc_label_u8", _ptr(img), _ptr(out), w, h, bgval, conn, _stream())
    return out


def bwlabel2(img, conn=8):
    """nscale::bwlabel2(binaryImage, conn, relab): labels = smallest raster index, bg -1."""
    h, w = _hw(img)
    out = _i32out(img)
    call("nsc_bwlabel2_u8", _ptr(img), _ptr(out), w, h, conn, _stream())
    return out


def cc_relabel(labels, bgval=-1):
    """ConnComponents::relabel in place; returns the number of components."""
    h, w = _hw(labels)
    n = ctypes.c_int(0)
    call("nsc_cc_relabel", _ptr(labels), w, h, bgval, ctypes.byref(n), _stream())
    return n.value


def bwareaopen2(img, min_size, max_size, conn=8):
    """nscale::bwareaopen2(img, false, true, min, max, conn, count) -> (mask, count)."""
    h, w = _hw(img)
    out = _u8out(img)
    n = ctypes.c_int(0)
    call("nsc_bwareaopen2_u8", _ptr(img), _ptr(out), w, h, min_size, max_size, conn, ctypes.byref(n), _stream())
    return out, n.value


def bounding_boxes(canon_labels, n):
    """[label | minx | maxx | miny | maxy] (5 x n) and per-label area for canonical labels 1..n."""
    h, w = _hw(canon_labels)
    out5 = torch.empty(5 * max(n, 1), dtype=torch.int32, device=canon_labels.device)
    area = torch.empty(max(n, 1), dtype=torch.int32, device=canon_labels.device)
    call("nsc_bounding_boxes", _ptr(canon_labels), w, h, n, _ptr(out5), _ptr(area), _stream())
    return out5[:5 * n].view(5, n), area[:n]


def _kernel_arg(kernel):
    import numpy as np
    k = np.ascontiguousarray(np.asarray(kernel, dtype=np.uint8))
    return k, k.ctypes.data_as(_P), int(k.shape[0])


def erode(img, kernel, outside=255):
    h, w = _hw(img)
    out = _u8out(img)
    k, kp, ks = _kernel_arg(kernel)
    call("nsc_erode_u8", _ptr(img), _ptr(out), w, h, kp, ks, outside, _stream())
    return out


def dilate(img, kernel, outside=0):
    h, w = _hw(img)
    out = _u8out(img)
    k, kp, ks = _kernel_arg(kernel)
    call("nsc_dilate_u8", _ptr(img), _ptr(out), w, h, kp, ks, outside, _stream())
    return out


def morph_open(img, kernel):
    """nscale::morphOpen<uchar>(image, kernel)."""
    h, w = _hw(img)
    out = _u8out(img)
    k, kp, ks = _kernel_arg(kernel)
    call("nsc_morph_open_u8", _ptr(img), _ptr(out), w, h, kp, ks, _stream())
    return out


def dist_transform(img):
    """cv::distanceTransform(img, CV_DIST_L2, CV_DIST_MASK_PRECISE) as used by the chain."""
    h, w = _hw(img)
    out = torch.empty(h, w, dtype=torch.float32, device=img.device)
    call("nsc_dist_transform_u8", _ptr(img), _ptr(out), w, h, _stream())
    return out


def local_maxima(img, conn=8):
    h, w = _hw(img)
    out = _u8out(img)
    call("nsc_local_maxima_f32", _ptr(img), _ptr(out), w, h, conn, _stream())
    return out


def local_minima(img, conn=8):
    h, w = _hw(img)
    out = _u8out(img)
    call("nsc_local_minima_f32", _ptr(img), _ptr(out), w, h, conn, _stream())
    return out


def cv_watershed(bgr, markers):
    """cv::watershed(image, markers) in place on a copy; returns the markers."""
    h, w = _hw(markers)
    m = markers.clone()
    call("nsc_cv_watershed", _ptr(bgr), _ptr(m), w, h, _stream())
    return m


def watershed2(bgr, img, conn=8):
    """nscale::watershed2(origImage, image, conn)."""
    h, w = _hw(img)
    out = _i32out(img)
    call("nsc_watershed2", _ptr(bgr), _ptr(img), _ptr(out), w, h, conn, _stream())
    return out


class Workspace:
    """Device workspace for segment_nuclei on tiles up to max_w x max_h."""

    def __init__(self, max_w, max_h):
        self.max_w, self.max_h = int(max_w), int(max_h)
        self.handle = _P()
        call("nsc_workspace_create", max_w, max_h, ctypes.byref(self.handle))

    def fits(self, w, h):
        return (w <= self.max_w and h <= self.max_h) or (w <= self.max_h and h <= self.max_w)

    def __del__(self):
        try:
            if self.handle:
                lib().nsc_workspace_destroy(self.handle)
        except Exception:
            pass


def default_params():
    p = SegParams()
    lib().nsc_seg_params_default(ctypes.byref(p))
    return p


STAGE_TYPES = {16: torch.float32, 17: torch.float32, 18: torch.float32, 20: torch.int32, 25: torch.int32}


def segment_nuclei(bgr, params=None, ws=None, stages=(), timing=False):
    """HistologicalEntities::segmentNuclei on a device BGR tile.

    Returns (status, labels, compcount, {stage_id: tensor}, {stage_id: ms})."""
    h, w = int(bgr.shape[0]), int(bgr.shape[1])
    p = params or default_params()
    ws = ws or Workspace(w, h)
    labels = torch.empty(h, w, dtype=torch.int32, device=bgr.device)
    arr = (_P * 32)()
    bufs = {}
    for s in stages:
        if s in (0, 19):
            bufs[s] = torch.zeros(h, w, 3, dtype=torch.uint8, device=bgr.device)
        else:
            bufs[s] = torch.zeros(h, w, dtype=STAGE_TYPES.get(s, torch.uint8), device=bgr.device)
        arr[s] = bufs[s].data_ptr()
    ms = (ctypes.c_double * 32)()
    cc = ctypes.c_int(0)
    L = lib()
    L.nsc_segment_nuclei.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, _P, _P]
    st = L.nsc_segment_nuclei(_ptr(bgr), w, h, ctypes.byref(p), _ptr(labels), ctypes.byref(cc), ws.handle,
                              arr if stages else None, ms if timing else None, _stream())
    if st < -2:
        from nscale_amd._lib import check
        check(st, "nsc_segment_nuclei")
    return st, labels, cc.value, bufs, {i: ms[i] for i in range(32) if ms[i] > 0}


def segment_nuclei_io(bgr, labels=None, params=None, ws=None, canonical=False, bbox_mode=BBOX_NONE, max_boxes=0,
                      with_area=False, stages=()):
    """HistologicalEntities::segmentNuclei with pitched images and the reference's compcount/bbox.

    bgr and labels may be pitched views (rows contiguous, spaced by their row stride).  bbox_mode:
    BBOX_REFERENCE_CPU = ConnComponents::boundingBox(labels, 0) as the reference CPU chain calls it,
    BBOX_ALL = one box per nucleus (nscale::gpu::boundingBox2 on the returned labels).
    Returns (status, labels, compcount, nuclei, bbox (5 x compcount) or None, area or None, stages)."""
    h, w = int(bgr.shape[0]), int(bgr.shape[1])
    p = params or default_params()
    ws = ws or Workspace(w, h)
    if labels is None:
        labels = torch.empty(h, w, dtype=torch.int32, device=bgr.device)
    bp, bstep = _rows(bgr)
    lp, lstep = _rows(labels)
    box = area = None
    if bbox_mode != BBOX_NONE:
        box = torch.empty(5 * max(max_boxes, 1), dtype=torch.int32, device=bgr.device)
        if with_area:
            area = torch.empty(max(max_boxes, 1), dtype=torch.int32, device=bgr.device)
    io = SegIO(bstep, lstep, 1 if canonical else 0, bbox_mode, _P(box.data_ptr()) if box is not None else None,
               _P(area.data_ptr()) if area is not None else None, max_boxes, 0, 0)
    arr = (_P * 32)()
    bufs = {}
    for s_ in stages:
        if s_ in (0, 19):
            bufs[s_] = torch.zeros(h, w, 3, dtype=torch.uint8, device=bgr.device)
        else:
            bufs[s_] = torch.zeros(h, w, dtype=STAGE_TYPES.get(s_, torch.uint8), device=bgr.device)
        arr[s_] = bufs[s_].data_ptr()
    st = lib().nsc_segment_nuclei_io(bp, w, h, ctypes.byref(p), lp, ctypes.byref(io), ws.handle,
                                     arr if stages else None, None, _stream())
    if st < -2:
        check(st, "nsc_segment_nuclei_io")
    n = io.compcount
    if box is not None:
        box = box[:5 * n].view(5, n)
        if area is not None:
            area = area[:n]
    return st, labels, n, io.nuclei, box, area, bufs


def pl_find_nucleus_candidates(bgr, params=None, ws=None, out=None):
    """HistologicalEntities::plFindNucleusCandidates (S1): (status, seg_norbc).  status 4 (CONTINUE)
    means the chain goes on; pitched views are accepted."""
    h, w = int(bgr.shape[0]), int(bgr.shape[1])
    p = params or default_params()
    ws = ws or Workspace(w, h)
    out = out if out is not None else torch.empty(h, w, dtype=torch.uint8, device=bgr.device)
    bp, bstep = _rows(bgr)
    op, ostep = _rows(out)
    st = lib().nsc_pl_find_nucleus_candidates(bp, bstep, w, h, ctypes.byref(p), op, ostep, ws.handle, _stream())
    if st < -2:
        check(st, "nsc_pl_find_nucleus_candidates")
    return st, out


def pl_separate_nuclei(bgr, seg_open, min_size_pl=30, conn=8, ws=None, out=None):
    """HistologicalEntities::plSeparateNuclei (A4): (status, seg_nonoverlap)."""
    h, w = int(bgr.shape[0]), int(bgr.shape[1])
    ws = ws or Workspace(w, h)
    out = out if out is not None else torch.empty(h, w, dtype=torch.uint8, device=bgr.device)
    bp, bstep = _rows(bgr)
    sp, sstep = _rows(seg_open)
    op, ostep = _rows(out)
    st = lib().nsc_pl_separate_nuclei(bp, bstep, sp, sstep, w, h, int(min_size_pl), int(conn), op, ostep, ws.handle,
                                      _stream())
    if st < -2:
        check(st, "nsc_pl_separate_nuclei")
    return st, out


def object_features(labels, bbox, img=None, channel=0, intensity=True, gradient=True):
    """nscale::ObjFeatures on the device: labels (h, w) int32, bbox (5, n) int32 [label | minx | maxx |
    miny | maxy], img an interleaved (h, w, c) or (h, w) uint8 image whose plane `channel` is measured.
    Returns a dict of device tensors: area, perimeter, major_axis, minor_axis, eccentricity,
    extent_ratio, circularity, and intensity (n, 8) / gradient (n, 6) when an image is given."""
    h, w = _hw(labels)
    n = int(bbox.shape[1])
    box = bbox.contiguous()
    dev = labels.device
    out = {"area": torch.empty(max(n, 1), dtype=torch.int32, device=dev)}
    for k in ("perimeter", "major_axis", "minor_axis", "eccentricity", "extent_ratio", "circularity"):
        out[k] = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
    chans = 1
    if img is not None:
        chans = 1 if img.dim() == 2 else int(img.shape[2])
        if intensity:
            out["intensity"] = torch.empty(max(n, 1), 8, dtype=torch.float32, device=dev)
        if gradient:
            out["gradient"] = torch.empty(max(n, 1), 6, dtype=torch.float32, device=dev)
    F = ObjFeaturesOut(*[_P(out[k].data_ptr()) if k in out else None for k, _ in ObjFeaturesOut._fields_])
    call("nsc_object_features", _ptr(labels), w, h, _ptr(box), n, _ptr(img) if img is not None else None, chans,
         int(channel), ctypes.byref(F), _stream())
    return {k: v[:n] for k, v in out.items()}


def normalization_target_parameters(bgr):
    """nscale::Normalization::targetParameters on a device BGR (h, w, 3) uint8 image (rows may be
    pitched): (mean[3], std[3], pixel count) of the foreground's l-alpha-beta values."""
    h, w = int(bgr.shape[0]), int(bgr.shape[1])
    bp, ld = _rows(bgr)
    m = (ctypes.c_float * 3)()
    s = (ctypes.c_float * 3)()
    n = ctypes.c_int64()
    call("nsc_normalization_target_parameters", bp, w, h, ld, m, s, ctypes.byref(n), _stream())
    return np.array(m, np.float32), np.array(s, np.float32), int(n.value)


def normalization(bgr, target_mean, target_std, out=None):
    """nscale::Normalization::normalization on the device: the BGR image with its foreground moved to
    the target l-alpha-beta statistics.  out may be bgr itself (in place) or None (a new tensor)."""
    h, w = int(bgr.shape[0]), int(bgr.shape[1])
    out = out if out is not None else torch.empty(h, w, 3, dtype=torch.uint8, device=bgr.device)
    bp, ld = _rows(bgr)
    op, old = _rows(out)
    m = (ctypes.c_float * 3)(*[float(v) for v in target_mean])
    s = (ctypes.c_float * 3)(*[float(v) for v in target_std])
    call("nsc_normalization", bp, w, h, ld, m, s, op, old, _stream())
    return out


def watershed_tier_runs():
    """Phase-1 runs per LDS tier (<=512, 1024, 2048, 4096, 6400 px, global bucket), process-wide sums."""
    arr = (ctypes.c_longlong * 6)()
    lib().nsc_watershed_tier_runs(arr, 6)
    return list(arr)


def watershed_stats(with_reason=False):
    """(decomposed-flood calls, sequential-fallback calls[, last fallback reason bits])."""
    a, b, c = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    lib().nsc_watershed_stats(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return (a.value, b.value, c.value) if with_reason else (a.value, b.value)


class Server:
    """Continuous-batching segmentNuclei server (nsc_server_*): requests are admitted when a slot is
    free and their footprint fits the HBM budget, and run concurrently on their own streams.

    submit() takes a device BGR tile (kept alive by the server object until the request is waited
    on) and returns a ticket; wait() returns (status, compcount, labels, bbox5, area, result)."""

    def __init__(self, slots=8, hbm_budget=200 << 30, max_w=4096, max_h=8192):
        self.handle = _P()
        cfg = ServerConfig(slots, hbm_budget, max_w, max_h)
        call("nsc_server_create", ctypes.byref(cfg), ctypes.byref(self.handle))
        self._live = {}

    def submit(self, bgr, canonical=False, max_labels=0, params=None, stage=0, seg_open=None):
        """stage 0: the whole chain (labels out); 1: plFindNucleusCandidates (S1, seg_norbc out);
        2: plSeparateNuclei (A4, seg_open in, seg_nonoverlap out)."""
        h, w = int(bgr.shape[0]), int(bgr.shape[1])
        assert bgr.is_contiguous() and bgr.dtype == torch.uint8 and bgr.dim() == 3 and bgr.shape[2] == 3
        dev = bgr.device if bgr.is_cuda else torch.device("cuda", torch.cuda.current_device())
        # outputs always live on the device, also for host inputs (the worker copies those in)
        labels = bbox = area = mask = None
        if stage == 0:
            labels = torch.empty(h, w, dtype=torch.int32, device=dev)
            if max_labels:
                bbox = torch.empty(5 * max_labels, dtype=torch.int32, device=dev)
                area = torch.empty(max_labels, dtype=torch.int32, device=dev)
        else:
            mask = torch.empty(h, w, dtype=torch.uint8, device=dev)
        if stage == 2:
            assert seg_open is not None and seg_open.is_cuda and seg_open.is_contiguous()
        req = Request(_P(bgr.data_ptr()), w, h, 0 if bgr.is_cuda else 1,
                      _P(labels.data_ptr()) if labels is not None else None, 1 if canonical else 0,
                      _P(bbox.data_ptr()) if bbox is not None else None,
                      _P(area.data_ptr()) if area is not None else None, max_labels,
                      ctypes.cast(ctypes.byref(params), _P) if params is not None else None, int(stage),
                      _P(mask.data_ptr()) if mask is not None else None,
                      _P(seg_open.data_ptr()) if seg_open is not None else None)
        # inputs must be ready on the device before another stream reads them
        if bgr.is_cuda:
            torch.cuda.current_stream().synchronize()
        t = ctypes.c_int64(0)
        call("nsc_server_submit", self.handle, ctypes.byref(req), ctypes.byref(t))
        self._live[t.value] = (bgr, labels, bbox, area, mask, seg_open)
        return t.value

    def wait(self, ticket):
        """(status, compcount, labels, bbox5, area, result) for the chain; (status, 0, mask, None,
        None, result) for a stage request."""
        res = Result()
        rc = lib().nsc_server_wait(self.handle, ticket, ctypes.byref(res))
        bgr, labels, bbox, area, mask, _ = self._live.pop(ticket)
        check(rc, "nsc_server_wait")
        if mask is not None:
            return res.status, 0, mask, None, None, res
        n = res.compcount
        if bbox is not None:
            bbox = bbox[:5 * n].view(5, n) if n else bbox[:0].view(5, 0)
            area = area[:n]
        return res.status, n, labels, bbox, area, res

    def stats(self):
        """(queued, running, device bytes the slots hold, budget)"""
        q, r = ctypes.c_int(0), ctypes.c_int(0)
        b, bud = ctypes.c_size_t(0), ctypes.c_size_t(0)
        call("nsc_server_stats", self.handle, ctypes.byref(q), ctypes.byref(r), ctypes.byref(b), ctypes.byref(bud))
        return q.value, r.value, b.value, bud.value

    def close(self):
        if self.handle:
            lib().nsc_server_destroy(self.handle)
            self.handle = _P()
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_bytes(w, h):
    return int(lib().nsc_request_bytes(w, h))
