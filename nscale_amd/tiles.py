"""Tile files of nu-segment (segment/nu-segment.cpp): PNG input tiles, the "<tile>.mask.png" label
outputs and their names, the file-level HistologicalEntities::segmentNuclei(in, out, ...) and the
directory driver, over the library's C ABI (PNG codec on the host, chain on the device)."""
import ctypes
import os

import numpy as np

from nscale_amd._lib import NscError, call, lib

IMAGE_EXTS = (".tif", ".tiff", ".png")  # nu-segment.cpp:127-130


def read_png(path):
    """cv::imread(path) (IMREAD_COLOR) for PNG files: (h, w, 3) uint8 BGR."""
    p = ctypes.c_void_p()
    w, h = ctypes.c_int(), ctypes.c_int()
    call("nsc_png_read", os.fsencode(path), ctypes.byref(p), ctypes.byref(w), ctypes.byref(h))
    try:
        buf = (ctypes.c_uint8 * (3 * w.value * h.value)).from_address(p.value)
        return np.frombuffer(buf, np.uint8).reshape(h.value, w.value, 3).copy()
    finally:
        lib().nsc_free_host(p)


def write_png(path, img):
    """cv::imwrite(path, img) for an (h, w) grey or (h, w, 3) BGR uint8 image."""
    img = np.ascontiguousarray(img, np.uint8)
    ch = 1 if img.ndim == 2 else img.shape[2]
    call("nsc_png_write", os.fsencode(path), img.ctypes.data_as(ctypes.c_void_p), img.shape[1], img.shape[0], ch, 0)


def mask_of_labels(labels):
    """The 8-bit image imwrite stores for an int32 label image: saturate_cast<uchar> of each label."""
    return np.clip(labels, 0, 255).astype(np.uint8)


def list_images(image_dir):
    """FileUtils::traverseDirectory over the tile extensions, recursive and sorted (nu-segment.cpp:132);
    a single file is a list of one."""
    if os.path.isfile(image_dir):
        return [image_dir]
    out = []
    for root, _, files in os.walk(image_dir):
        out += [os.path.join(root, f) for f in files if f.lower().endswith(IMAGE_EXTS)]
    return sorted(out)


def output_names(image, image_dir, out_dir):
    """getFiles (nu-segment.cpp:123-158): the tile's extension replaced by .mask.png / .bounds.csv and
    its directory prefix image_dir replaced by out_dir (created)."""
    base = image_dir if os.path.isdir(image_dir) else os.path.dirname(image_dir)
    rel = os.path.relpath(os.path.splitext(image)[0], base)
    stem = os.path.join(out_dir, rel)
    os.makedirs(os.path.dirname(stem) or ".", exist_ok=True)
    return stem + ".mask.png", stem + ".bounds.csv"


def segment_file(in_path, out_path, params=None, want_bbox=True):
    """HistologicalEntities::segmentNuclei(in, out, compcount, bbox, ...): the chain on the PNG tile
    `in`; on success the label image is written to `out`.  Returns (status, compcount, bbox (5, n)
    int32 array or None)."""
    cc = ctypes.c_int()
    box = ctypes.c_void_p()
    pp = ctypes.byref(params) if params is not None else None
    st = lib().nsc_segment_nuclei_file(os.fsencode(in_path), os.fsencode(out_path), pp, ctypes.byref(cc),
                                       ctypes.byref(box) if want_bbox else None, None)
    if st < -2:
        raise NscError("nsc_segment_nuclei_file failed (%d): %s" % (st, lib().nsc_last_error().decode()))
    bbox = None
    if box.value:
        n = cc.value
        bbox = np.frombuffer((ctypes.c_int32 * (5 * n)).from_address(box.value), np.int32).reshape(5, n).copy()
        lib().nsc_free_host(box)
    return st, cc.value, bbox


def read_image(path):
    """cv::imread(path) (IMREAD_COLOR) as (h, w, 3) uint8 BGR: PNG by the library's codec, other
    formats (the reference's .tif / .tiff tiles) through PIL when it is installed; None when the
    file cannot be decoded (imread's empty Mat)."""
    if str(path).lower().endswith(".png"):
        try:
            return read_png(path)
        except NscError:
            return None
    try:
        from PIL import Image
    except ImportError:
        return None
    try:
        with Image.open(path) as im:
            rgb = np.asarray(im.convert("RGB"))
    except Exception:
        return None
    return np.ascontiguousarray(rgb[..., ::-1])


def segment_image(bgr, out_path, params=None, ws=None):
    """The chain on a host BGR tile (the file-level segmentNuclei after imread): on success writes
    the label image to out_path as imwrite stores it.  Returns (status, compcount)."""
    import torch
    from nscale_amd import gpu
    from nscale_amd._lib import BBOX_REFERENCE_CPU
    dev = torch.from_numpy(np.ascontiguousarray(bgr)).cuda()
    h, w = bgr.shape[:2]
    st, lab, cc, _, _, _, _ = gpu.segment_nuclei_io(dev, params=params, ws=ws, bbox_mode=BBOX_REFERENCE_CPU,
                                                   max_boxes=1 << 16)
    if st != 0:
        return st, 0
    mask = torch.empty(h, w, dtype=torch.uint8, device=dev.device)
    call("nsc_labels_to_u8", ctypes.c_void_p(lab.data_ptr()), 0, w, h, ctypes.c_void_p(mask.data_ptr()), 0,
         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    write_png(out_path, mask.cpu().numpy())
    return st, cc


def segment_directory(image_dir, out_dir, params=None):
    """nu-segment's serial mode (nu-segment.cpp:505-528): every tile under image_dir segmented, the
    label image written to the name output_names gives (PNG tiles through segment_file, other
    formats decoded by read_image).  Returns [(image, status, compcount)]; an undecodable tile gives
    the reference's INVALID_IMAGE (-1)."""
    res = []
    ws = None
    for img in list_images(image_dir):
        mask, _ = output_names(img, image_dir, out_dir)
        if img.lower().endswith(".png"):
            st, cc, _ = segment_file(img, mask, params, want_bbox=False)
        else:
            bgr = read_image(img)
            if bgr is None:
                st, cc = -1, 0
            else:
                if ws is None or not ws.fits(bgr.shape[1], bgr.shape[0]):
                    from nscale_amd import gpu
                    ws = gpu.Workspace(bgr.shape[1], bgr.shape[0])
                st, cc = segment_image(bgr, mask, params, ws)
        res.append((img, st, cc))
    return res
