"""ctypes binding of oracle/libsph_oracle.so (the CPU restatement of
CCPUParticleSimulator, src/CCPUParticleSimulator.cpp). Test infrastructure only.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libsph_oracle.so"

DAM_BREAK, FOUNTAIN = 0, 1
ORDER_REFERENCE, ORDER_CANONICAL = 0, 1

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def lib_path():
    return os.path.join(HERE, _LIB_NAME)


def build(quiet=True):
    """Compile the oracle with its Makefile (gcc, seconds)."""
    out = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    return lib_path()


_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    src = os.path.join(HERE, "sph_oracle.c")
    if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(src):
        build()
    lib = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "orc_constants": (None, [_f64p]),
        "orc_grid_res": (C.c_int, [C.c_float]),
        "orc_dam_break_count": (C.c_uint32, [C.c_float]),
        "orc_create": (vp, [C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_count": (C.c_uint32, [vp]),
        "orc_max_count": (C.c_uint32, [vp]),
        "orc_grid_dims": (None, [vp, _i32p]),
        "orc_add_particle": (C.c_int, [vp] + [C.c_float] * 6),
        "orc_add_particles": (None, [vp, vp, vp, C.c_uint32]),
        "orc_setup_scene": (C.c_int, [vp]),
        "orc_set_gravity": (None, [vp, C.c_float, C.c_float, C.c_float]),
        "orc_set_order_mode": (None, [vp, C.c_int]),
        "orc_set_diagnostics": (None, [vp, C.c_int]),
        "orc_update_grid": (None, [vp]),
        "orc_update_density_pressure": (None, [vp]),
        "orc_update_forces": (None, [vp]),
        "orc_integrate": (None, [vp]),
        "orc_step": (None, [vp, C.c_uint32]),
        "orc_get_state": (None, [vp, vp, vp, vp, vp, vp]),
        "orc_set_state": (None, [vp, vp, vp]),
        "orc_get_cells": (None, [vp, _u32p]),
        "orc_get_buckets": (None, [vp, _u32p, _u32p]),
        "orc_get_neighbour_counts": (None, [vp, _u32p]),
        "orc_get_force_scale": (None, [vp, _f32p]),
        "orc_hash": (C.c_uint32, [C.c_float] * 6 + [C.c_int] * 3),
        "orc_hash_many": (None, [_f32p, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                 C.c_int, C.c_int, C.c_int, _u32p]),
        "orc_stable_perm": (None, [_u32p, C.c_uint32, C.c_uint32, _u32p]),
        "orc_cell_start": (None, [_u32p, C.c_uint32, C.c_uint32, _u32p]),
        "orc_wall_accel": (None, [_f32p, _f32p, C.c_float, C.c_float, C.c_float, _f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def constants():
    """{poly6, spiky, visc, hsq, h, mass, rest_density} as the reference computes them."""
    out = np.zeros(7, np.float64)
    _load().orc_constants(out)
    keys = ["poly6", "spiky", "visc", "hsq", "h", "mass", "rest_density"]
    return dict(zip(keys, out.tolist()))


def grid_res(box):
    return int(_load().orc_grid_res(float(box)))


def dam_break_count(box):
    return int(_load().orc_dam_break_count(float(box)))


def hash_positions(pos, box, res):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    out = np.zeros(len(pos), np.uint32)
    b = [float(x) for x in np.broadcast_to(np.asarray(box, np.float32), (3,))]
    r = [int(x) for x in np.broadcast_to(np.asarray(res), (3,))]
    _load().orc_hash_many(pos.reshape(-1), len(pos), *b, *r, out)
    return out


def stable_perm(keys, ncells):
    keys = np.ascontiguousarray(keys, np.uint32)
    out = np.zeros(len(keys), np.uint32)
    _load().orc_stable_perm(keys, len(keys), int(ncells), out)
    return out


def cell_start(sorted_keys, ncells):
    sk = np.ascontiguousarray(sorted_keys, np.uint32)
    out = np.zeros(int(ncells) + 1, np.uint32)
    _load().orc_cell_start(sk, len(sk), int(ncells), out)
    return out


def wall_accel(pos, vel, box):
    b = [float(x) for x in np.broadcast_to(np.asarray(box, np.float32), (3,))]
    out = np.zeros(3, np.float32)
    _load().orc_wall_accel(np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(vel, np.float32),
                           *b, out)
    return out


_GOLDEN = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1


def _splitmix64_at(state, k):
    """Output k (k >= 1) of the SplitMix64 stream that starts at `state`."""
    z = (state + k * _GOLDEN) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def _unit24(x):
    return np.float32((x >> 40) * 2.0 ** -24)


def disc_emission(seed, first_id, count, centre, radius, speed, spread, floor_y):
    """Particles first_id .. first_id+count-1 of a floor-disc particle source
    (the benchmark fountain; the reference's own fountain is the fixed 7-particle
    pattern of src/CBaseParticleSimulator.cpp:187-210). Particle g draws from the
    SplitMix64 stream at seed + 32*g*golden: up to 15 points (2u-1, 2u-1) of the
    unit square until one lies in the unit disc (else the centre), then
    v_y = speed * (1 + spread*(2u-1)) from draw 31. Float32 arithmetic, one IEEE
    operation at a time. Returns pos (count, 3) and vel (count, 3)."""
    f = np.float32
    one, two = f(1.0), f(2.0)
    cx, cz, r, sp, spr = f(centre[0]), f(centre[1]), f(radius), f(speed), f(spread)
    pos = np.zeros((count, 3), np.float32)
    vel = np.zeros((count, 3), np.float32)
    for q in range(count):
        g = first_id + q
        st = (int(seed) + 32 * g * _GOLDEN) & _M64
        a = b = f(0.0)
        for t in range(15):
            u = f(two * _unit24(_splitmix64_at(st, 2 * t + 1)) - one)
            w = f(two * _unit24(_splitmix64_at(st, 2 * t + 2)) - one)
            if f(f(u * u) + f(w * w)) <= one:
                a, b = u, w
                break
        pos[q] = (f(cx + f(r * a)), f(floor_y), f(cz + f(r * b)))
        vel[q, 1] = f(sp * f(one + f(spr * f(two * _unit24(_splitmix64_at(st, 31)) - one))))
    return pos, vel


class OracleSim:
    """The reference CPU simulator, restated (CCPUParticleSimulator)."""

    def __init__(self, box, capacity=0, scenario=DAM_BREAK):
        lib = _load()
        b = [float(x) for x in np.broadcast_to(np.asarray(box, np.float32), (3,))]
        self.box = np.array(b, np.float32)
        self._lib = lib
        self._h = lib.orc_create(*b, int(capacity), int(scenario))
        if not self._h:
            raise MemoryError("orc_create failed")

    def close(self):
        if self._h:
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def count(self):
        return int(self._lib.orc_count(self._h))

    @property
    def max_count(self):
        return int(self._lib.orc_max_count(self._h))

    @property
    def grid_dims(self):
        out = np.zeros(3, np.int32)
        self._lib.orc_grid_dims(self._h, out)
        return tuple(int(x) for x in out)

    def add_particle(self, x, y, z, vx=0.0, vy=0.0, vz=0.0):
        return self._lib.orc_add_particle(self._h, x, y, z, vx, vy, vz)

    def add_particles(self, pos, vel=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        self._lib.orc_add_particles(self._h, _ptr(pos), _ptr(vel), len(pos))

    def setup_scene(self):
        return self._lib.orc_setup_scene(self._h)

    def set_gravity(self, g):
        self._lib.orc_set_gravity(self._h, *[float(x) for x in g])

    def set_order_mode(self, mode):
        self._lib.orc_set_order_mode(self._h, int(mode))

    def set_diagnostics(self, on=True):
        self._lib.orc_set_diagnostics(self._h, int(bool(on)))

    def update_grid(self):
        self._lib.orc_update_grid(self._h)

    def update_density_pressure(self):
        self._lib.orc_update_density_pressure(self._h)

    def update_forces(self):
        self._lib.orc_update_forces(self._h)

    def integrate(self):
        self._lib.orc_integrate(self._h)

    def step(self, n=1):
        self._lib.orc_step(self._h, int(n))

    def state(self):
        n = self.count
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        acc = np.zeros((n, 3), np.float32)
        rho = np.zeros(n, np.float32)
        prs = np.zeros(n, np.float32)
        self._lib.orc_get_state(self._h, _ptr(pos), _ptr(vel), _ptr(acc), _ptr(rho), _ptr(prs))
        return dict(pos=pos, vel=vel, acc=acc, density=rho, pressure=prs)

    def set_state(self, pos=None, vel=None):
        p = None if pos is None else np.ascontiguousarray(pos, np.float32)
        v = None if vel is None else np.ascontiguousarray(vel, np.float32)
        self._lib.orc_set_state(self._h, _ptr(p), _ptr(v))

    def cells(self):
        out = np.zeros(self.count, np.uint32)
        self._lib.orc_get_cells(self._h, out)
        return out

    def buckets(self):
        dims = self.grid_dims
        nc = dims[0] * dims[1] * dims[2]
        start = np.zeros(nc + 1, np.uint32)
        ids = np.zeros(self.count, np.uint32)
        self._lib.orc_get_buckets(self._h, start, ids)
        return start, ids

    def neighbour_counts(self):
        out = np.zeros(self.count, np.uint32)
        self._lib.orc_get_neighbour_counts(self._h, out)
        return out

    def force_scale(self):
        out = np.zeros((self.count, 3), np.float32)
        self._lib.orc_get_force_scale(self._h, out.reshape(-1))
        return out
