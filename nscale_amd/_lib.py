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


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", _I), ("hbm_budget", ctypes.c_size_t), ("max_w", _I), ("max_h", _I)]


class Request(ctypes.Structure):
    _fields_ = [("bgr", _P), ("w", _I), ("h", _I), ("host_input", _I), ("labels", _P), ("canonical", _I),
                ("bbox5", _P), ("area", _P), ("max_labels", _I), ("params", _P), ("stage", _I), ("mask_out", _P),
                ("mask_in", _P)]


STAGE_CHAIN, STAGE_FIND_CANDIDATES, STAGE_SEPARATE_NUCLEI = 0, 1, 2


class SegIO(ctypes.Structure):
    """nsc_seg_io: row steps, label form and the bounding-box outputs of nsc_segment_nuclei_io."""
    _fields_ = [("bgr_step", ctypes.c_size_t), ("labels_step", ctypes.c_size_t), ("canonical", _I),
                ("bbox_mode", _I), ("bbox5", _P), ("area", _P), ("max_boxes", _I),
                ("compcount", _I), ("nuclei", _I)]


BBOX_NONE, BBOX_REFERENCE_CPU, BBOX_ALL = 0, 1, 2


class ObjFeaturesOut(ctypes.Structure):
    """nsc_object_features_out: device arrays of the per-nucleus features."""
    _fields_ = [(k, _P) for k in ("area", "perimeter", "major_axis", "minor_axis", "eccentricity", "extent_ratio",
                                  "circularity", "intensity", "gradient")]


class ReconDesc(ctypes.Structure):
    _fields_ = [("marker", _P), ("mask", _P), ("out", _P), ("w", _I), ("h", _I), ("conn", _I)]


class Result(ctypes.Structure):
    _fields_ = [("status", _I), ("compcount", _I), ("error", _I), ("slot", _I),
                ("queue_ms", ctypes.c_double), ("run_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("start_ms", ctypes.c_double), ("end_ms", ctypes.c_double)]


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
        L.nsc_synth_slide_tile.argtypes = [_P, _I, _I, _I, _I, ctypes.c_uint64, _P]
        L.nsc_request_bytes.restype = ctypes.c_size_t
        L.nsc_request_bytes.argtypes = [_I, _I]
        L.nsc_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]
        L.nsc_server_submit.argtypes = [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_int64)]
        L.nsc_server_wait.argtypes = [_P, ctypes.c_int64, ctypes.POINTER(Result)]
        L.nsc_server_poll.argtypes = [_P, ctypes.c_int64, ctypes.POINTER(Result)]
        L.nsc_server_stats.argtypes = [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.POINTER(ctypes.c_size_t)]
        L.nsc_server_destroy.argtypes = [_P]
        L.nsc_imreconstruct_u8_batched.argtypes = [ctypes.POINTER(ReconDesc), _I, _P]
        L.nsc_imreconstruct_f32_batched.argtypes = [ctypes.POINTER(ReconDesc), _I, _P]
        L.nsc_imhmin_f32_batched.argtypes = [ctypes.POINTER(ReconDesc), _I, ctypes.c_float, _P]
        L.nsc_segment_nuclei_io.argtypes = [_P, _I, _I, _P, _P, ctypes.POINTER(SegIO), _P, _P, _P, _P]
        L.nsc_pl_find_nucleus_candidates.argtypes = [_P, ctypes.c_size_t, _I, _I, _P, _P, ctypes.c_size_t, _P, _P]
        L.nsc_pl_separate_nuclei.argtypes = [_P, ctypes.c_size_t, _P, ctypes.c_size_t, _I, _I, _I, _I, _P,
                                             ctypes.c_size_t, _P, _P]
        L.nsc_object_features.argtypes = [_P, _I, _I, _P, _I, _P, _I, _I, ctypes.POINTER(ObjFeaturesOut), _P]
        L.nsc_normalization_target_parameters.argtypes = [_P, _I, _I, ctypes.c_size_t, _P, _P, _P, _P]
        L.nsc_normalization.argtypes = [_P, _I, _I, ctypes.c_size_t, _P, _P, _P, ctypes.c_size_t, _P]
        L.nsc_png_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(_P), ctypes.POINTER(_I), ctypes.POINTER(_I)]
        L.nsc_png_write.argtypes = [ctypes.c_char_p, _P, _I, _I, _I, ctypes.c_size_t]
        L.nsc_free_host.argtypes = [_P]
        L.nsc_free_host.restype = None
        L.nsc_labels_to_u8.argtypes = [_P, ctypes.c_size_t, _I, _I, _P, ctypes.c_size_t, _P]
        L.nsc_segment_nuclei_file.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _P, ctypes.POINTER(_I),
                                              ctypes.POINTER(_P), _P]
        L.nsc_stitch_tile.argtypes = [_P, ctypes.c_size_t, _P, _I, _I, _I, _I, _I, _I, _I, _I, ctypes.c_int32, _P, _P,
                                      ctypes.c_size_t, ctypes.POINTER(_I), ctypes.POINTER(_I), _P]
        L.nsc_recon_graph_create.argtypes = [ctypes.POINTER(ReconDesc), _I, _I, _I, ctypes.POINTER(_P)]
        L.nsc_recon_graph_run.argtypes = [_P, ctypes.POINTER(_I), _P]
        L.nsc_recon_graph_destroy.argtypes = [_P]
        L.nsc_watershed_tier_runs.argtypes = [ctypes.POINTER(ctypes.c_longlong), _I]
        L.nsc_recon_autotune.argtypes = [ctypes.POINTER(ReconDesc), _I, _I, _P, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                         ctypes.POINTER(ctypes.c_float)]
        _lib = L
    return _lib


def check(rc, what=""):
    if rc < 0:
        msg = lib().nsc_last_error().decode(errors="replace")
        raise NscError("%s failed (%d): %s" % (what, rc, msg))
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)
