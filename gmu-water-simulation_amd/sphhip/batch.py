"""SceneBatch: many independent grid-SPH scenes advanced together by
libsph_hip.so, one launch per stage for all of them (include/sph_hip.h, batched
scenes). The reference advances one scene per CBaseParticleSimulator
(src/CBaseParticleSimulator.cpp:116-144); each scene of a batch computes exactly
what it computes alone.
"""
import ctypes as C

import numpy as np

from . import _lib


def scene_desc(box, capacity, gravity=(0.0, -9.80665, 0.0), dt=0.01):
    d = _lib.sph_scene_desc()
    b = np.broadcast_to(np.asarray(box, np.float32), (3,))
    g = np.asarray(gravity, np.float32).reshape(3)
    for a in range(3):
        d.box[a] = float(b[a])
        d.gravity[a] = float(g[a])
    d.dt = float(np.float32(dt))
    d.capacity = int(capacity)
    return d


class SceneBatch:
    """A pool of particle slots shared by independent scenes.

    scenes: list of (box, capacity) or dicts with keys box, capacity[, gravity, dt],
    created with ids 0..len-1. pool: total slots (0 = the scenes' capacities, each
    rounded up to 256). max_scenes: scene ids available to add_scene later.
    """

    def __init__(self, scenes=(), pool=0, max_scenes=0, flags=0, max_neighbours=0):
        lib = _lib.load()
        self._lib = lib
        descs = []
        for sc in scenes:
            if isinstance(sc, dict):
                descs.append(scene_desc(sc["box"], sc["capacity"], sc.get("gravity", (0.0, -9.80665, 0.0)),
                                        sc.get("dt", 0.01)))
            else:
                descs.append(scene_desc(sc[0], sc[1]))
        arr = (_lib.sph_scene_desc * max(len(descs), 1))(*descs)
        bp = _lib.sph_batch_params()
        bp.pool = int(pool)
        bp.max_scenes = int(max_scenes or len(descs))
        bp.flags = int(flags)
        bp.max_neighbours = int(max_neighbours)
        ctx = C.c_void_p()
        _lib.check(lib.sph_batch_create(C.byref(bp), arr, len(descs), C.byref(ctx)))
        self._ctx = ctx

    # -- lifetime -------------------------------------------------------
    def close(self):
        if self._ctx is not None:
            self._lib.sph_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- scenes ---------------------------------------------------------
    def add_scene(self, box, capacity, gravity=(0.0, -9.80665, 0.0), dt=0.01):
        d = scene_desc(box, capacity, gravity, dt)
        sid = C.c_uint32(0)
        _lib.check(self._lib.sph_scene_add(self._ctx, C.byref(d), C.byref(sid)))
        return sid.value

    def remove_scene(self, scene):
        _lib.check(self._lib.sph_scene_remove(self._ctx, int(scene)))

    def upload(self, scene, pos, vel=None, ids=None):
        pos = _lib.f32(pos, (-1, 3))
        vel = None if vel is None else _lib.f32(vel, (-1, 3))
        ids = None if ids is None else _lib.u32(ids)
        _lib.check(self._lib.sph_scene_upload(self._ctx, int(scene), _lib.ptr(pos), _lib.ptr(vel), _lib.ptr(ids),
                                              len(pos)))

    def append(self, scene, pos, vel=None, ids=None):
        pos = _lib.f32(pos, (-1, 3))
        vel = None if vel is None else _lib.f32(vel, (-1, 3))
        ids = None if ids is None else _lib.u32(ids)
        _lib.check(self._lib.sph_scene_append(self._ctx, int(scene), _lib.ptr(pos), _lib.ptr(vel), _lib.ptr(ids),
                                              len(pos)))

    def count(self, scene):
        n = C.c_uint32(0)
        _lib.check(self._lib.sph_scene_count(self._ctx, int(scene), C.byref(n)))
        return n.value

    def download(self, scene, by_id=True):
        n = self.count(scene)
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        ids = np.zeros(n, np.uint32)
        _lib.check(self._lib.sph_scene_download(self._ctx, int(scene), _lib.ptr(pos), _lib.ptr(vel), _lib.ptr(ids),
                                                n, int(by_id)))
        return pos, vel, ids

    def readback(self, scene, with_velocity=True):
        """Positions (and velocities) of a scene by id via the device scatter into pinned memory."""
        return _lib.readback(self._lib, self._ctx, scene, with_velocity)

    def set_gravity(self, scene, g):
        g = np.asarray(g, np.float32).reshape(3)
        _lib.check(self._lib.sph_scene_set_gravity(self._ctx, int(scene), *[float(v) for v in g]))

    def grid(self, scene):
        xyz = np.zeros(3, np.int32)
        nc, first, slots = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
        _lib.check(self._lib.sph_scene_grid(self._ctx, int(scene), _lib.ptr(xyz), C.byref(nc), C.byref(first),
                                            C.byref(slots)))
        return tuple(int(v) for v in xyz), nc.value, first.value, slots.value

    def info(self):
        v = [C.c_uint32(0) for _ in range(4)]
        _lib.check(self._lib.sph_batch_info(self._ctx, *[C.byref(x) for x in v]))
        return dict(zip(("scenes", "pool", "free_slots", "active_slots"), (x.value for x in v)))

    def debug(self, scene, kind):
        n = self.count(scene)
        if kind == _lib.DEBUG_CELL_START:
            out = np.zeros(self.grid(scene)[1] + 1, np.uint32)
        elif kind in (_lib.DEBUG_DENSITY, _lib.DEBUG_PRESSURE):
            out = np.zeros(n, np.float32)
        elif kind == _lib.DEBUG_ACCEL:
            out = np.zeros((n, 3), np.float32)
        else:
            out = np.zeros(n, np.uint32)
        _lib.check(self._lib.sph_scene_debug(self._ctx, int(scene), kind, _lib.ptr(out), out.nbytes))
        return out

    def set_emitter(self, scene, kind=_lib.EMIT_DISC, rate=1024, limit=0, centre=(0.0, 0.0), radius=0.0,
                    speed=2.0, spread=0.1, seed=0):
        """Attach a device particle source (None-like kind=-1 detaches)."""
        if kind is None or kind < 0:
            _lib.check(self._lib.sph_scene_set_emitter(self._ctx, int(scene), None))
            return
        e = _lib.sph_emitter()
        e.kind = int(kind)
        e.rate = int(rate)
        e.limit = int(limit)
        e.centre[0], e.centre[1] = float(centre[0]), float(centre[1])
        e.radius = float(radius)
        e.speed = float(speed)
        e.spread = float(spread)
        e.seed = int(seed)
        _lib.check(self._lib.sph_scene_set_emitter(self._ctx, int(scene), C.byref(e)))

    # -- stepping -------------------------------------------------------
    def step(self, nsteps=1, timed=False):
        """Advance every scene by nsteps; returns device ms if timed."""
        if timed:
            ms = C.c_float(0.0)
            _lib.check(self._lib.sph_step(self._ctx, int(nsteps), C.byref(ms)))
            return ms.value
        _lib.check(self._lib.sph_step(self._ctx, int(nsteps), None))
        return None

    def phase(self, phase):
        _lib.check(self._lib.sph_phase(self._ctx, int(phase), None))

    def sync(self):
        _lib.check(self._lib.sph_sync(self._ctx))

    def set_graphs(self, enable):
        _lib.check(self._lib.sph_set_graphs(self._ctx, int(bool(enable))))

    def stage_times(self, reset=False):
        """{stage: (device ms summed, launches)} since the last reset, for all
        scenes together (a context made with FLAG_PROFILE)."""
        ms = np.zeros(8, np.float32)
        cnt = np.zeros(8, np.uint32)
        _lib.check(self._lib.sph_stage_times(self._ctx, _lib.ptr(ms), _lib.ptr(cnt), int(reset)))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(_lib.STAGES)}

    def set_variant(self, variant, mode=_lib.NEIGHBOURS_WALK):
        """Kernel shape `variant` and neighbour mode (NEIGHBOURS_WALK, _LISTS, _MASKS)."""
        _lib.check(self._lib.sph_set_variant(self._ctx, int(variant), int(mode)))

    def variant(self):
        v, m = C.c_int(0), C.c_int(0)
        _lib.check(self._lib.sph_get_variant(self._ctx, C.byref(v), C.byref(m)))
        return v.value, m.value

    def autotune(self, steps=5):
        """Time every (variant, neighbour mode) choice for `steps` steps from the
        current state, keep the fastest and leave the state unchanged. Returns
        {(variant, mode): ms_per_step} for the choices that ran."""
        ms = np.zeros(3 * _lib.NUM_VARIANTS, np.float32)
        _lib.check(self._lib.sph_autotune(self._ctx, int(steps), _lib.ptr(ms)))
        return {(v, m): float(ms[3 * v + m]) for v in range(_lib.NUM_VARIANTS) for m in (0, 1, 2)
                if ms[3 * v + m] >= 0}


def variant_shape(variant):
    """(threads per workgroup, candidates per lane per trip, mask-mode LDS words
    per lane, mask-mode drain stall percent, mask-mode word queue 0/1)."""
    t, u, c, st, q = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    _lib.check(_lib.load().sph_variant_shape(int(variant), C.byref(t), C.byref(u), C.byref(c), C.byref(st),
                                             C.byref(q)))
    return t.value, u.value, c.value, st.value, q.value
