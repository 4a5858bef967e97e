"""Whole-slide segmentation: a slide cut into overlapping tiles, the tiles streamed through the
tile server, each tile's nuclei stitched into one slide-wide set (nsc_stitch_tile: a nucleus is
kept by the tile whose core holds its box centre), and the outputs nu-segment writes per tile
("<slide>-<x>-<y>.mask.png", the tile naming of pipeline/nu-features.cpp:411-428) plus one table
of the slide's nuclei ("<slide>.nuclei.csv": id, box in slide coordinates, area, owning tile)."""
import ctypes
import os

import numpy as np
import torch

from nscale_amd import gpu, tiles
from nscale_amd._lib import call

_P = ctypes.c_void_p


def tile_grid(width, height, tile=4096, overlap=64):
    """Tiles (x0, y0, w, h) of a width x height slide on the grid of stride tile - overlap, the last
    row and column clipped to the slide."""
    stride = tile - overlap
    xs = list(range(0, max(1, width - overlap), stride))
    ys = list(range(0, max(1, height - overlap), stride))
    return [(x, y, min(tile, width - x), min(tile, height - y)) for y in ys for x in xs]


class Stitcher:
    """Slide-wide ids for the nuclei of overlapping tiles, tile by tile (any order)."""

    def __init__(self, width, height, overlap):
        self.width, self.height, self.overlap = int(width), int(height), int(overlap)
        self.next_id = 1
        self.rows = []       # per tile: (ids, x0, y0, bbox5 (5, n) numpy, area numpy or None)
        self.truncated = 0
        self.conflicts = 0

    def add(self, x0, y0, labels, bbox, area=None, owned=False):
        """labels: (h, w) canonical int32 device labels; bbox: (5, n) int32 device box block.
        Returns (ids (n,) device int32: slide id or 0, owned (h, w) device labels as slide ids or None)."""
        h, w = int(labels.shape[0]), int(labels.shape[1])
        n = int(bbox.shape[1])
        ids = torch.zeros(max(n, 1), dtype=torch.int32, device=labels.device)
        own = torch.empty(h, w, dtype=torch.int32, device=labels.device) if owned else None
        box = bbox.contiguous()
        kept, trunc = ctypes.c_int(0), ctypes.c_int(0)
        stream = _P(torch.cuda.current_stream().cuda_stream)
        call("nsc_stitch_tile", _P(labels.data_ptr()), int(labels.stride(0) * 4), _P(box.data_ptr()), n, int(x0),
             int(y0), w, h, self.width, self.height, self.overlap, self.next_id, _P(ids.data_ptr()),
             _P(own.data_ptr()) if own is not None else None, int(own.stride(0) * 4) if own is not None else 0,
             ctypes.byref(kept), ctypes.byref(trunc), stream)
        self.next_id += kept.value
        self.truncated += trunc.value
        self.rows.append((ids[:n].cpu().numpy(), int(x0), int(y0), box.cpu().numpy(),
                          area.cpu().numpy()[:n] if area is not None else None))
        return ids[:n], own

    def nuclei(self):
        """(N, 8) int64 table sorted by id: id, minx, maxx, miny, maxy (slide coordinates), area
        (-1 if unknown), tile x0, tile y0."""
        out = []
        for ids, x0, y0, box, area in self.rows:
            k = ids > 0
            if not k.any():
                continue
            a = area[k] if area is not None else np.full(int(k.sum()), -1)
            out.append(np.stack([ids[k], box[1][k] + x0, box[2][k] + x0, box[3][k] + y0, box[4][k] + y0, a,
                                 np.full(a.shape, x0), np.full(a.shape, y0)], 1))
        if not out:
            return np.zeros((0, 8), np.int64)
        t = np.concatenate(out).astype(np.int64)
        return t[np.argsort(t[:, 0], kind="stable")]


def write_nuclei_csv(path, table):
    with open(path, "w") as f:
        f.write("id,minx,maxx,miny,maxy,area,tile_x,tile_y\n")
        for r in table:
            f.write(",".join(str(int(v)) for v in r) + "\n")


def segment_slide(tile_source, width, height, tile=4096, overlap=64, name="slide", out_dir=None, slots=8,
                  budget_gb=64, params=None, max_labels=65536, tiles_subset=None, raster=None):
    """Segments a width x height slide tile by tile.  tile_source(x0, y0, w, h, out) fills the device
    BGR tensor `out` (h, w, 3) with that region.  Writes "<name>-<x0>-<y0>.mask.png" per tile and
    "<name>.nuclei.csv" under out_dir when given.  raster (optional, (height, width) int32 device
    tensor of zeros) receives every tile's kept nuclei as slide ids; pixels that two tiles' kept
    nuclei both claim are counted in Stitcher.conflicts.  Returns (Stitcher, per-tile [(x0, y0,
    status, nuclei)])."""
    grid = tile_grid(width, height, tile, overlap)
    if tiles_subset is not None:
        grid = [grid[i] for i in tiles_subset]
    st = Stitcher(width, height, overlap)
    srv = gpu.Server(slots=slots, hbm_budget=budget_gb << 30, max_w=tile, max_h=tile)
    ring = [torch.empty(tile * tile * 3, dtype=torch.uint8, device="cuda") for _ in range(2 * slots)]
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
    inflight, report = [], []

    # Writing masks needs the reference's labels (smallest raster index, background -1: the file
    # is their 8-bit saturation); the stitcher then relabels them canonically here.  Without
    # outputs the server relabels and boxes them itself.
    write = bool(out_dir)

    def finish(entry):
        tk, (x0, y0, w, h) = entry
        status, n, lab, bbox, area, res = srv.wait(tk)
        if res.error != 0:  # e.g. more nuclei than max_labels: reported, not stitched
            status = res.error
        if write:
            mask = torch.empty(h, w, dtype=torch.uint8, device=lab.device)
            if status == 0:  # the chain's labels, as imwrite stores them
                call("nsc_labels_to_u8", _P(lab.data_ptr()), 0, w, h, _P(mask.data_ptr()), 0,
                     _P(torch.cuda.current_stream().cuda_stream))
            else:
                mask.zero_()
            tiles.write_png(os.path.join(out_dir, "%s-%d-%d.mask.png" % (name, x0, y0)), mask.cpu().numpy())
            if status == 0:
                n = gpu.cc_relabel(lab)
                bbox, area = gpu.bounding_boxes(lab, n)
        report.append((x0, y0, status, n))
        if status == 0:
            _, own = st.add(x0, y0, lab, bbox, area, owned=raster is not None)
            if raster is not None:
                reg = raster[y0:y0 + h, x0:x0 + w]
                st.conflicts += int(((reg > 0) & (own > 0) & (reg != own)).sum())
                reg.copy_(torch.where(own > 0, own, reg))

    try:
        for i, (x0, y0, w, h) in enumerate(grid):
            if len(inflight) == len(ring):
                finish(inflight.pop(0))
            buf = ring[i % len(ring)][:h * w * 3].view(h, w, 3)
            tile_source(x0, y0, w, h, buf)
            inflight.append((srv.submit(buf, canonical=not write, max_labels=0 if write else max_labels, params=params),
                             (x0, y0, w, h)))
        for e in inflight:
            finish(e)
    finally:
        srv.close()
    if out_dir:
        write_nuclei_csv(os.path.join(out_dir, "%s.nuclei.csv" % name), st.nuclei())
    return st, report
