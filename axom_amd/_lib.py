"""ctypes binding of libaxsq.so (the C-ABI declared in include/axsq.h).

The library is the only compute path: if it is missing, or no GPU is
visible when a compute call is made, these functions raise.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("AXSQ_LIB_PATH") or os.path.join(HERE, "lib", "libaxsq.so")
HEADERS = [os.path.join(os.path.dirname(HERE), "include", h) for h in ("axsq.h", "axsq_quest.h")]

AXSQ_OK = 0
AXSQ_FP64 = 0
AXSQ_FP32_TRAVERSAL = 1
QUERY_POINTS, QUERY_BOXES, QUERY_RAYS = 0, 1, 2

ERRORS = {-1: "AXSQ_ERR_INVALID_ARG", -2: "AXSQ_ERR_HIP", -3: "AXSQ_ERR_NOMEM",
          -4: "AXSQ_ERR_STACK_OVERFLOW", -5: "AXSQ_ERR_STATE"}
ERRORS_BY_NAME = {v: k for k, v in ERRORS.items()}
AXSQ_ALLOC_DEVICE, AXSQ_ALLOC_MANAGED, AXSQ_ALLOC_PINNED_HOST = 0, 1, 2

_lib = None

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_dbl = ctypes.c_double
_flt = ctypes.c_float
_dbl3 = ctypes.c_double * 3


class AxsqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_SIGS = {
    "axsq_last_error": (ctypes.c_char_p, []),
    "axsq_version": (_i32, []),
    "axsq_free": (_i32, [_vp]),
    "axsq_bvh_build": (_i32, [ctypes.POINTER(_vp), _vp, _i64, _dbl, _i32, _vp]),
    "axsq_bvh_destroy": (_i32, [_vp]),
    "axsq_bvh_build2d": (_i32, [ctypes.POINTER(_vp), _vp, _i64, _dbl, _i32, _vp]),
    "axsq_bvh_dimension": (_i32, [_vp]),
    "axsq_bvh_rebuild": (_i32, [_vp, _vp, _i64, _dbl, _i32, _vp]),
    "axsq_bvh_device_arrays": (_i32, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                      ctypes.POINTER(ctypes.c_int32)]),
    "axsq_bvh_bounds": (_i32, [_vp, _dbl3, _dbl3]),
    "axsq_bvh_num_leaves": (_i64, [_vp]),
    "axsq_bvh_export": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "axsq_bvh_find_count": (_i32, [_vp, _i32, _vp, _vp, _vp, _i64, _i64, _dbl, _vp, _vp,
                                   ctypes.POINTER(_i64), _vp]),
    "axsq_bvh_find_fill": (_i32, [_vp, _i32, _vp, _vp, _vp, _i64, _i64, _dbl, _vp, _vp, _vp,
                                  _vp]),
    "axsq_bvh_find_into": (_i32, [_vp, _i32, _vp, _vp, _vp, _i64, _i64, _dbl, _vp, _vp, _vp,
                                  _i64, ctypes.POINTER(_i64), _vp]),
    "axsq_bvh_find_into_noalloc": (_i32, [_vp, _i32, _vp, _vp, _vp, _i64, _i64, _dbl, _vp, _vp,
                                          _vp, _i64, ctypes.POINTER(_i64), _vp]),
    "axsq_bvh_build_f32": (_i32, [ctypes.POINTER(_vp), _vp, _i64, _i32, _flt, _vp]),
    "axsq_bvh_rebuild_f32": (_i32, [_vp, _vp, _i64, _flt, _vp]),
    "axsq_bvh_precision": (_i32, [_vp]),
    "axsq_bvh_find_count_f32": (_i32, [_vp, _i32, _vp, _vp, _vp, _i64, _i64, _flt, _vp, _vp,
                                       ctypes.POINTER(_i64), _vp]),
    "axsq_bvh_find_fill_f32": (_i32, [_vp, _i32, _vp, _vp, _vp, _i64, _i64, _flt, _vp, _vp, _vp,
                                      _vp]),
    "axsq_bvh_find_into_f32": (_i32, [_vp, _i32, _vp, _vp, _vp, _i64, _i64, _flt, _vp, _vp, _vp,
                                      _i64, ctypes.POINTER(_i64), _vp]),
    "axsq_bvh_find_f32": (_i32, [_vp, _i32, _vp, _vp, _vp, _i64, _i64, _flt, _vp, _vp,
                                 ctypes.POINTER(_vp), ctypes.POINTER(_i64), _vp]),
    "axsq_bvh_set_allocator": (_i32, [_vp, _i32]),
    "axsq_bvh_stats_enable": (_i32, [_vp, _i32]),
    "axsq_bvh_stats_get": (_i32, [_vp, ctypes.POINTER(_i64)]),
    "axsq_bvh_get_allocator": (_i32, [_vp]),
    "axsq_bvh_write_vtk": (_i32, [_vp, ctypes.c_char_p]),
    "axsq_host_free": (_i32, [_vp]),
    "axsq_release_workspace": (_i32, []),
    "axsq_stl_read": (_i32, [ctypes.c_char_p, ctypes.POINTER(_vp), ctypes.POINTER(_i64),
                             ctypes.POINTER(_i64)]),
    "axsq_sdf_create_from_stl": (_i32, [ctypes.POINTER(_vp), ctypes.c_char_p, _i32, _i32, _i32,
                                        _vp]),
    "axsq_sdf_eval_point": (_i32, [_vp, _dbl3, _i32, ctypes.POINTER(_dbl), _dbl3, _dbl3]),
    "axsq_bvh_find": (_i32, [_vp, _i32, _vp, _vp, _vp, _i64, _i64, _dbl, _vp, _vp,
                             ctypes.POINTER(_vp), ctypes.POINTER(_i64), _vp]),
    "axsq_sdf_create": (_i32, [ctypes.POINTER(_vp), _vp, _vp, _vp, _i64, _vp, _i64, _i32, _i32,
                               _i32, _vp]),
    "axsq_sdf_create_cells": (_i32, [ctypes.POINTER(_vp), _vp, _vp, _vp, _i64, _vp, _vp, _i64,
                                     _i32, _i32, _i32, _vp]),
    "axsq_sdf_destroy": (_i32, [_vp]),
    "axsq_sdf_mesh_bounds": (_i32, [_vp, _dbl3, _dbl3]),
    "axsq_sdf_update_vertices": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "axsq_sdf_eval": (_i32, [_vp, _vp, _vp, _vp, _i64, _i64, _i32, _vp, _vp, _vp, _vp, _vp]),
    "axsq_sdf_backward": (_i32, [_vp, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "axsq_prof_enable": (_i32, [_i32]),
    "axsq_prof_reset": (_i32, []),
    "axsq_prof_get": (_i32, [ctypes.c_char_p, ctypes.POINTER(_i64), ctypes.POINTER(_dbl)]),
    "axsq_sdf_stats_enable": (_i32, [_vp, _i32]),
    "axsq_sdf_stats_get": (_i32, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "axsq_sdf_last_suspended": (_i32, [_vp, ctypes.POINTER(_i64)]),
    "axsq_sdf_last_early": (_i32, [_vp, ctypes.POINTER(_i64)]),
    "axsq_tri_tri_intersect": (_i32, [_vp, _vp, _i64, _i32, _dbl, _vp, _vp]),
    "axsq_mesh_candidates": (_i32, [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _i64, _vp,
                                    _i64, _i32, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                    ctypes.POINTER(_i64), _vp]),
    # quest signed-distance interface (include/axsq_quest.h)
    "QUEST_signed_distance_init_serial": (_i32, [ctypes.c_char_p]),
    "QUEST_signed_distance_init_serial_bufferify": (_i32, [ctypes.c_char_p, _i32]),
    "QUEST_signed_distance_initialized": (ctypes.c_bool, []),
    "QUEST_signed_distance_get_mesh_bounds": (None, [_dbl3, _dbl3]),
    "QUEST_signed_distance_set_dimension": (None, [_i32]),
    "QUEST_signed_distance_set_closed_surface": (None, [ctypes.c_bool]),
    "QUEST_signed_distance_set_compute_signs": (None, [ctypes.c_bool]),
    "QUEST_signed_distance_set_allocator": (None, [_i32]),
    "QUEST_signed_distance_set_verbose": (None, [ctypes.c_bool]),
    "QUEST_signed_distance_use_shared_memory": (None, [ctypes.c_bool]),
    "QUEST_signed_distance_set_execution_space": (None, [_i32]),
    "QUEST_signed_distance_evaluate_0": (_dbl, [_dbl, _dbl, _dbl]),
    "QUEST_signed_distance_evaluate_1": (_dbl, [_dbl, _dbl, _dbl] + [ctypes.POINTER(_dbl)] * 6),
    "QUEST_signed_distance_finalize": (None, []),
    "axsq_signed_distance_init_mesh": (_i32, [_vp, _vp, _vp, _i64, _vp, _i64]),
    "axsq_signed_distance_evaluate_batch": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "axsq_signed_distance_set_precision": (None, [_i32]),
    "axsq_mesh_self_intersections": (_i32, [_vp, _vp, _vp, _i64, _vp, _i64, _dbl,
                                            ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                            ctypes.POINTER(_i64), ctypes.POINTER(_vp),
                                            ctypes.POINTER(_i64), _vp]),
}


def header_symbols():
    """Names of the entry points declared in include/axsq.h and axsq_quest.h."""
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names.update(re.findall(r"\b((?:axsq|QUEST)_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(names)


def lib():
    """Loads libaxsq.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run axom_amd/build.py (or "
                               "__graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != AXSQ_OK:
        msg = lib().axsq_last_error()
        raise AxsqError(rc, msg.decode() if msg else "")
    return rc


def last_error():
    """Text of the calling thread's last error ("" after a call that succeeded)."""
    msg = lib().axsq_last_error()
    return msg.decode() if msg else ""
