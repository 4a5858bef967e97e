"""nscale_amd: MI355X (gfx950) implementation of nscale's nuclei-segmentation morphology chain.

The compute lives in libnscale_hip.so (hand-written HIP kernels behind the C ABI declared in
include/nscale_hip.h).  This package holds the Python-side binding used by the tests and the
benchmark: `nscale_amd.gpu` mirrors the reference's nscale:: / nscale::gpu:: operator names on
torch device tensors, `nscale_amd.synth` produces the seeded synthetic inputs.
"""
from nscale_amd._lib import LIB_PATH, NscError, SegParams, lib  # noqa: F401

__all__ = ["lib", "LIB_PATH", "NscError", "SegParams"]
