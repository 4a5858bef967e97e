"""Seeded synthetic inputs (numpy, host) from the library's generators."""
import ctypes

import numpy as np

from nscale_amd._lib import call


def synth_mask(h, w, seed):
    """Grayscale mask tile: background 220, ellipses U[60,120], noise sigma 8."""
    out = np.empty((h, w), np.uint8)
    call("nsc_synth_mask", out.ctypes.data_as(ctypes.c_void_p), w, h, seed)
    return out


def synth_marker(mask):
    """marker = mask - 10 (saturating)."""
    out = np.empty_like(mask)
    call("nsc_synth_marker", mask.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), mask.size)
    return out


def synth_he(h, w, seed):
    """BGR H&E-like tile (background, nuclei, red-blood-cell discs)."""
    out = np.empty((h, w, 3), np.uint8)
    call("nsc_synth_he", out.ctypes.data_as(ctypes.c_void_p), w, h, seed)
    return out


def serve_sizes(n, seed):
    """Tile sizes (h, w) of the batched / served workloads: short side log-uniform in [512, 4096],
    long side = short * U[1, 2], random orientation."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = int(round(float(np.exp(rng.uniform(np.log(512), np.log(4096))))))
        b = int(round(a * float(rng.uniform(1.0, 2.0))))
        out.append((a, b) if rng.random() < 0.5 else (b, a))
    return out
