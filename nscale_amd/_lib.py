"""ctypes binding of libnscale_hip.so (the C ABI declared in include/nscale_hip.h).

The library is built in-tree by __graft_entry__.build() (make -C nscale_amd/csrc).  Loading
fails loudly when it is missing: there is no fallback implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnscale_hip.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_lib = None


class NscError(RuntimeError):
    pass


class SegParams(ctypes.Structure):
    """nsc_seg_params: the arguments of HistologicalEntities::segmentNuclei."""
    _fields_ = [("blue", ctypes.c_uint8), ("green", ctypes.c_uint8), ("red", ctypes.c_uint8),
                ("T1", ctypes.c_double), ("T2", ctypes.c_double), ("G1", ctypes.c_uint8),
                ("minSize", _I), ("maxSize", _I), ("G2", ctypes.c_uint8),
                ("minSizePl", _I), ("minSizeSeg", _I), ("maxSizeSeg", _I),
                ("fillHolesConnectivity", _I), ("reconConnectivity", _I), ("watershedConnectivity", _I)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NscError("libnscale_hip.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.nsc_version.restype = ctypes.c_char_p
        L.nsc_last_error.restype = ctypes.c_char_p
        L.nsc_synth_mask.argtypes = [_P, _I, _I, ctypes.c_uint64]
        L.nsc_synth_he.argtypes = [_P, _I, _I, ctypes.c_uint64]
        L.nsc_synth_marker.argtypes = [_P, _P, ctypes.c_size_t]
        _lib = L
    return _lib


def check(rc, what=""):
    if rc < 0:
        msg = lib().nsc_last_error().decode(errors="replace")
        raise NscError("%s failed (%d): %s" % (what, rc, msg))
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)
