"""Operator surface of nscale:: / nscale::gpu:: (MorphologicOperations.h, HistologicalEntities.h)
over torch device tensors, calling libnscale_hip.so through its C ABI on torch's current stream.

torch is used for device memory and streams only; every operator runs the library's HIP kernels.
"""
import ctypes

import torch

from nscale_amd._lib import SegParams, call, lib

_P = ctypes.c_void_p


def _ptr(t):
    assert t.is_cuda and t.is_contiguous(), "expects a contiguous device tensor"
    return _P(t.data_ptr())


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
