"""ctypes binding of lib/libsph_hip.so (the C ABI declared in include/sph_hip.h).

The library is the product: if it is not built this module raises, it never
falls back to anything else.
"""
import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libsph_hip.so")

# status codes (include/sph_hip.h)
SPH_OK, SPH_ERR_INVALID, SPH_ERR_CAPACITY, SPH_ERR_HIP, SPH_ERR_OOM, SPH_ERR_NO_DEVICE, SPH_ERR_STATE, \
    SPH_ERR_FAULT = range(8)
PHASE_GRID, PHASE_DENSITY, PHASE_FORCES, PHASE_COLLISIONS, PHASE_INTEGRATE = range(5)
FLAG_DIAGNOSTICS, FLAG_PROFILE, FLAG_REFERENCE_ORDER = 1, 2, 4
NEIGHBOUR_LISTS_OFF = 0xFFFFFFFF
(DEBUG_CELL_PRESORT, DEBUG_PERM, DEBUG_CELL, DEBUG_CELL_START, DEBUG_NEIGHBOURS, DEBUG_DENSITY,
 DEBUG_PRESSURE, DEBUG_ACCEL, DEBUG_IDS) = range(9)
STAGES = ["hash", "sort", "cells", "reorder", "density", "forces", "collisions", "integrate"]
EMIT_FOUNTAIN, EMIT_DISC = 0, 1
NUM_VARIANTS = 21
DEFAULT_VARIANT = 12
NEIGHBOURS_WALK, NEIGHBOURS_LISTS, NEIGHBOURS_MASKS = 0, 1, 2
GROUP_SLOTS = 256

# Every entry point of include/sph_hip.h with its ctypes signature.
_vp = C.c_void_p
_f = C.c_float
_u32 = C.c_uint32
SIGNATURES = {
    "sph_default_params": (None, [_vp, _f, _u32]),
    "sph_grid_res": (C.c_int, [_f]),
    "sph_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "sph_destroy": (None, [_vp]),
    "sph_upload": (C.c_int, [_vp, _vp, _vp, _vp, _u32]),
    "sph_append": (C.c_int, [_vp, _vp, _vp, _vp, _u32]),
    "sph_set_gravity": (C.c_int, [_vp, _f, _f, _f]),
    "sph_phase": (C.c_int, [_vp, C.c_int, _vp]),
    "sph_step": (C.c_int, [_vp, _u32, _vp]),
    "sph_step_noalloc": (C.c_int, [_vp, _vp, _u32, _u32, _vp]),
    "sph_download": (C.c_int, [_vp, _vp, _vp, _vp, _u32, C.c_int]),
    "sph_debug": (C.c_int, [_vp, C.c_int, _vp, C.c_size_t]),
    "sph_count": (C.c_int, [_vp, C.POINTER(_u32)]),
    "sph_grid_dims": (C.c_int, [_vp, _vp, _vp]),
    "sph_device_name": (C.c_int, [C.c_char_p, C.c_size_t]),
    "sph_sync": (C.c_int, [_vp]),
    "sph_stage_times": (C.c_int, [_vp, _vp, _vp, C.c_int]),
    "sph_last_error": (C.c_char_p, []),
    "sph_scene_dam_break": (C.c_int, [_f, _vp, _u32, C.POINTER(_u32)]),
    "sph_scene_lattice": (C.c_int, [_u32, _u32, _u32, _vp, _f, C.c_uint64, _vp]),
    "sph_scene_block": (C.c_int, [_vp, _u32, _u32, _u32, _u32, _f, C.c_uint64, _vp]),
    "sph_default_scene": (None, [_vp, _f, _u32]),
    "sph_batch_create": (C.c_int, [_vp, _vp, _u32, C.POINTER(_vp)]),
    "sph_scene_add": (C.c_int, [_vp, _vp, C.POINTER(_u32)]),
    "sph_scene_remove": (C.c_int, [_vp, _u32]),
    "sph_scene_upload": (C.c_int, [_vp, _u32, _vp, _vp, _vp, _u32]),
    "sph_scene_append": (C.c_int, [_vp, _u32, _vp, _vp, _vp, _u32]),
    "sph_scene_download": (C.c_int, [_vp, _u32, _vp, _vp, _vp, _u32, C.c_int]),
    "sph_scene_set_gravity": (C.c_int, [_vp, _u32, _f, _f, _f]),
    "sph_scene_count": (C.c_int, [_vp, _u32, C.POINTER(_u32)]),
    "sph_scene_debug": (C.c_int, [_vp, _u32, C.c_int, _vp, C.c_size_t]),
    "sph_scene_grid": (C.c_int, [_vp, _u32, _vp, _vp, _vp, _vp]),
    "sph_batch_info": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "sph_scene_set_emitter": (C.c_int, [_vp, _u32, _vp]),
    "sph_set_graphs": (C.c_int, [_vp, C.c_int]),
    "sph_set_variant": (C.c_int, [_vp, C.c_int, C.c_int]),
    "sph_get_variant": (C.c_int, [_vp, _vp, _vp]),
    "sph_variant_shape": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "sph_autotune": (C.c_int, [_vp, _u32, _vp]),
    "sph_slot_bytes": (C.c_uint64, [_u32]),
    "sph_scene_readback": (C.c_int, [_vp, _u32, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_u32)]),
    "sph_set_event_stride": (C.c_int, [_vp, _u32]),
    "sph_profile_events": (C.c_int, [_vp, _vp, _u32, C.POINTER(_u32), C.c_int]),
    "sph_write_logs": (C.c_int, [_vp, _u32, C.c_char_p, C.c_char_p]),
    "sph_export_logs": (C.c_int, [_vp, C.c_char_p, C.c_char_p]),
    "sph_server_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "sph_server_destroy": (None, [_vp]),
    "sph_server_submit": (C.c_int, [_vp, _vp, C.POINTER(C.c_uint64)]),
    "sph_server_job": (C.c_int, [_vp, C.c_uint64, _vp]),
    "sph_server_wait": (C.c_int, [_vp, C.c_uint64, C.c_double, _vp]),
    "sph_server_result": (C.c_int, [_vp, C.c_uint64, _vp, _vp, _u32]),
    "sph_server_stats_get": (C.c_int, [_vp, _vp]),
}


class sph_params(C.Structure):
    _fields_ = [("box", C.c_float * 3), ("gravity", C.c_float * 3), ("dt", C.c_float),
                ("capacity", C.c_uint32), ("flags", C.c_uint32), ("max_neighbours", C.c_uint32)]


class sph_device_state(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("px", "py", "pz", "vx", "vy", "vz", "id")]


NOALLOC_MAX_PARTICLES = 65536
NOALLOC_MAX_CELLS = 262144


class sph_scene_desc(C.Structure):
    _fields_ = [("box", C.c_float * 3), ("gravity", C.c_float * 3), ("dt", C.c_float), ("capacity", C.c_uint32)]


class sph_batch_params(C.Structure):
    _fields_ = [("pool", C.c_uint32), ("max_scenes", C.c_uint32), ("flags", C.c_uint32),
                ("max_neighbours", C.c_uint32)]


class sph_emitter(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("rate", C.c_uint32), ("limit", C.c_uint32), ("centre", C.c_float * 2),
                ("radius", C.c_float), ("speed", C.c_float), ("spread", C.c_float), ("seed", C.c_uint64)]


class sph_profile_event(C.Structure):
    _fields_ = [("iteration", C.c_uint64), ("fps", C.c_double), ("update_grid", C.c_double),
                ("update_density_pressure", C.c_double), ("update_forces", C.c_double),
                ("update_collisions", C.c_double), ("integrate", C.c_double)]


class SphError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("sph status %d: %s" % (status, msg))
        self.status = status


_lib = None


def load():
    """Load lib/libsph_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsph_hip.so is not built (%s); run `make -C %s` or "
                           "__graft_entry__.build()" % (LIB_PATH, PKG_ROOT))
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != SPH_OK:
        raise SphError(status, load().sph_last_error().decode(errors="replace"))


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


def readback(lib, ctx, scene, with_velocity=True):
    """sph_scene_readback: positions (and velocities) by id, copied out of the
    library's pinned buffer."""
    pp, vp, n = C.c_void_p(), C.c_void_p(), C.c_uint32(0)
    check(lib.sph_scene_readback(ctx, int(scene), int(bool(with_velocity)), C.byref(pp), C.byref(vp), C.byref(n)))
    k = n.value
    if k == 0:
        return np.zeros((0, 3), np.float32), (np.zeros((0, 3), np.float32) if with_velocity else None)
    pos = np.ctypeslib.as_array(C.cast(pp, C.POINTER(C.c_float)), shape=(k, 3)).copy()
    vel = np.ctypeslib.as_array(C.cast(vp, C.POINTER(C.c_float)), shape=(k, 3)).copy() if with_velocity else None
    return pos, vel


def profile_events(lib, ctx, reset=False):
    """sph_profile_events as a list of dicts (sProfilingEvent fields)."""
    n = C.c_uint32(0)
    check(lib.sph_profile_events(ctx, None, 0, C.byref(n), 0))
    arr = (sph_profile_event * max(n.value, 1))()
    check(lib.sph_profile_events(ctx, arr, n.value, C.byref(n), int(bool(reset))))
    return [{f: getattr(arr[i], f) for f, _ in sph_profile_event._fields_} for i in range(n.value)]


def write_logs(events, stem, label):
    """sph_write_logs for a list of event dicts (exportLogs format)."""
    arr = (sph_profile_event * max(len(events), 1))()
    for i, e in enumerate(events):
        for f, _ in sph_profile_event._fields_:
            setattr(arr[i], f, e.get(f, 0))
    check(load().sph_write_logs(arr, len(events), str(stem).encode(), str(label).encode()))


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)
