"""sph_step_noalloc (the entry point that allocates nothing) against sph_step.

Both run the same kernels on the same state, so after the same steps the
arrays must be bitwise equal in array order: positions, velocities and the
permuted ids. Two calls follow each other on the same GPU, the second with a
smaller scene in another box, so the second finds the first's contents in the
shared scratch array; an odd step count exercises the final id copy.
Device memory for the caller's arrays comes from hipMalloc in the test (the
caller owns it); the entry point itself allocates nothing.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import lattice_numpy

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip_runtime():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMalloc.restype = C.c_int
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipFree.restype = C.c_int
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpy.restype = C.c_int
    hip.hipDeviceSynchronize.argtypes = []
    hip.hipDeviceSynchronize.restype = C.c_int
    return hip


class CallerArrays:
    """Seven device arrays of `cap` elements (px, py, pz, vx, vy, vz, id), owned by the caller."""

    NAMES = ("px", "py", "pz", "vx", "vy", "vz", "id")

    def __init__(self, hip, cap):
        self.hip, self.cap, self.ptrs = hip, cap, []
        for _ in self.NAMES:
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), C.c_size_t(4 * cap)) == 0
            self.ptrs.append(p)

    def state(self, _lib):
        st = _lib.sph_device_state()
        for name, p in zip(self.NAMES, self.ptrs):
            setattr(st, name, p)
        return st

    def put(self, pos, vel, ids):
        n = len(pos)
        assert n <= self.cap
        cols = [pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2]]
        for p, a in zip(self.ptrs, cols + [ids]):
            a = np.ascontiguousarray(a)
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(4 * n), H2D) == 0

    def get(self, n):
        out = []
        for p, dt in zip(self.ptrs, [np.float32] * 6 + [np.uint32]):
            a = np.empty(n, dt)
            assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, C.c_size_t(4 * n), D2H) == 0
            out.append(a)
        return np.stack(out[:3], 1), np.stack(out[3:6], 1), out[6]

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def step_with_context(hip_mod, box, pos, vel, ids, steps):
    """sph_create + sph_upload + sph_step + sph_download in array order."""
    _lib = hip_mod._lib
    lib = _lib.load()
    n = len(pos)
    p = _lib.sph_params()
    lib.sph_default_params(C.byref(p), box, n)
    ctx = C.c_void_p()
    _lib.check(lib.sph_create(C.byref(p), C.byref(ctx)))
    try:
        _lib.check(lib.sph_upload(ctx, _lib.ptr(pos), _lib.ptr(vel), _lib.ptr(ids), n))
        _lib.check(lib.sph_step(ctx, steps, None))
        po, vo, io = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32), np.empty(n, np.uint32)
        _lib.check(lib.sph_download(ctx, _lib.ptr(po), _lib.ptr(vo), _lib.ptr(io), n, 0))
    finally:
        lib.sph_destroy(ctx)
    return po, vo, io


def step_without_alloc(hip_mod, arrays, box, pos, vel, ids, steps):
    _lib = hip_mod._lib
    lib = _lib.load()
    p = _lib.sph_params()
    lib.sph_default_params(C.byref(p), box, len(pos))
    arrays.put(pos, vel, ids)
    st = arrays.state(_lib)
    _lib.check(lib.sph_step_noalloc(C.byref(p), C.byref(st), len(pos), steps, None))
    assert arrays.hip.hipDeviceSynchronize() == 0
    return arrays.get(len(pos))


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


def test_noalloc_equals_sph_step_on_two_calls(hip):
    hrt = _hip_runtime()
    arrays = CallerArrays(hrt, 4096)
    try:
        cases = []
        # 1: the 4,096-particle lattice at rest, 7 steps (odd: the ids end in scratch and are copied back)
        pos1 = lattice_numpy(16, 16, 16, 0.4, 0.01, 0)
        cases.append((0.4, pos1, np.zeros_like(pos1), np.arange(len(pos1), dtype=np.uint32), 7))
        # 2: a smaller block in a larger box, moving, with shuffled ids, 4 steps
        rng = np.random.default_rng(11)
        pos2 = lattice_numpy(12, 14, 10, 0.6, 0.01, 5)
        vel2 = rng.uniform(-0.3, 0.3, pos2.shape).astype(np.float32)
        ids2 = rng.permutation(len(pos2)).astype(np.uint32)
        cases.append((0.6, pos2, vel2, ids2, 4))
        assert len(pos2) < len(pos1)
        for box, pos, vel, ids, steps in cases:
            print("noalloc case: %d particles, box %.1f, %d steps" % (len(pos), box, steps), flush=True)
            ref = step_with_context(hip, box, pos, vel, ids, steps)
            got = step_without_alloc(hip, arrays, box, pos, vel, ids, steps)
            assert np.array_equal(got[2], ref[2]), "ids differ"
            assert bits_equal(got[0], ref[0]), "positions differ"
            assert bits_equal(got[1], ref[1]), "velocities differ"
            assert not bits_equal(got[0], pos), "the state did not advance"
    finally:
        arrays.free()


def test_noalloc_rejects_sizes_beyond_the_maximum(hip):
    _lib = hip._lib
    lib = _lib.load()
    hrt = _hip_runtime()
    arrays = CallerArrays(hrt, 8)
    try:
        p = _lib.sph_params()
        lib.sph_default_params(C.byref(p), 0.4, 8)
        st = arrays.state(_lib)
        assert lib.sph_step_noalloc(C.byref(p), C.byref(st), _lib.NOALLOC_MAX_PARTICLES + 1, 1, None) == \
            _lib.SPH_ERR_CAPACITY
        assert b"maximum" in lib.sph_last_error()
        lib.sph_default_params(C.byref(p), 3.2, 8)   # 71^3 cells
        assert lib.sph_step_noalloc(C.byref(p), C.byref(st), 8, 1, None) == _lib.SPH_ERR_CAPACITY
        assert b"cells" in lib.sph_last_error()
    finally:
        arrays.free()
