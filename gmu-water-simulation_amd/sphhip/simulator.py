"""HIPGridParticleSimulator: the grid-SPH simulator behind the reference's
CBaseParticleSimulator interface (include/CBaseParticleSimulator.h:29-110), with
every phase running in libsph_hip.so on the GPU.

Method names and meanings follow the reference so that a caller of
CGPUParticleSimulator / CCPUParticleSimulator finds the same surface:
setupScene, step, updateGrid, updateDensityPressure, updateForces,
updateCollisions, integrate (each phase returns its device time in ms, as the
reference's phase methods return ms), setGravityVector, toggleGravity,
getSelectedDevice, getParticlesCount, getMaxParticlesCount.
"""
import ctypes as C

import numpy as np

from . import _lib
from . import scenes

DAM_BREAK, FOUNTAIN = 0, 1  # SimulationScenario, include/CBaseParticleSimulator.h:21-25
GRAVITY_ACCELERATION = np.float32(-9.80665)  # include/CBaseParticleSimulator.h:19
H = np.float32(0.0457)


class HIPGridParticleSimulator:
    """Drop-in for CGPUParticleSimulator (src/CGPUParticleSimulator.cpp)."""

    def __init__(self, box_size, scenario=DAM_BREAK, capacity=None, flags=0, fused=True, max_neighbours=0,
                 device_emitter=False):
        lib = _lib.load()
        self._lib = lib
        box = np.broadcast_to(np.asarray(box_size, np.float32), (3,)).astype(np.float32)
        self.box = box
        self.m_scenario = scenario
        self.gravity = np.array([0.0, GRAVITY_ACCELERATION, 0.0], np.float32)
        self.fused = fused
        self.flags = int(flags)
        self.max_neighbours = int(max_neighbours)
        # FOUNTAIN: add the 7 particles of generateParticles on the device
        # (sph_scene_set_emitter) instead of appending them from the host
        self.device_emitter = bool(device_emitter)
        self._ctx = None
        self._capacity = capacity
        self.m_particlesCount = 0
        self.m_maxParticlesCount = 0
        self.totalIteration = 0

    # -- lifetime -------------------------------------------------------
    def _create(self, capacity):
        if self._ctx is not None:
            self.close()
        p = _lib.sph_params()
        self._lib.sph_default_params(C.byref(p), float(self.box[0]), int(capacity))
        for a in range(3):
            p.box[a] = float(self.box[a])
            p.gravity[a] = float(self.gravity[a])
        p.flags = self.flags
        p.max_neighbours = self.max_neighbours
        ctx = C.c_void_p()
        _lib.check(self._lib.sph_create(C.byref(p), C.byref(ctx)))
        self._ctx = ctx
        self._capacity = capacity

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

    # -- CBaseParticleSimulator surface ---------------------------------
    def _calculated_count(self):
        """src/CBaseParticleSimulator.cpp:40-41."""
        half = float(np.float32(H / np.float32(2.0)))
        bz, by, bx = (float(v) for v in (self.box[2], self.box[1], self.box[0]))
        return int(np.ceil(bz / half) * np.ceil(by / half) * np.ceil(float(np.float32(self.box[0]) / np.float32(4)) / half))

    def setupScene(self):
        """CBaseParticleSimulator::setupScene + upload (src/CGPUBaseParticleSimulator.cpp:22-26)."""
        if self.m_scenario == DAM_BREAK:
            if not (self.box[0] == self.box[1] == self.box[2]):
                raise ValueError("the reference dam break is defined for a cubic box")
            pos = scenes.dam_break(float(self.box[0]))
            self._create(self._capacity or len(pos))
            self.upload(pos)
            self.m_maxParticlesCount = len(pos)
        else:
            cap = self._capacity or self._calculated_count()
            self._create(cap)
            self.m_maxParticlesCount = self._calculated_count()
            self.m_particlesCount = 0
            if self.device_emitter:
                e = _lib.sph_emitter()
                e.kind = _lib.EMIT_FOUNTAIN
                e.limit = self.m_maxParticlesCount
                _lib.check(self._lib.sph_scene_set_emitter(self._ctx, 0, C.byref(e)))

    def setGravityVector(self, g):
        self.gravity = np.asarray(g, np.float32).reshape(3)
        if self._ctx is not None:
            _lib.check(self._lib.sph_set_gravity(self._ctx, *[float(v) for v in self.gravity]))

    def toggleGravity(self):
        """src/CBaseParticleSimulator.cpp:101-109."""
        if float(np.linalg.norm(self.gravity.astype(np.float64))) > 0.0:
            self.setGravityVector([0.0, 0.0, 0.0])
        else:
            self.setGravityVector([0.0, GRAVITY_ACCELERATION, 0.0])

    def getSelectedDevice(self):
        buf = C.create_string_buffer(256)
        _lib.check(self._lib.sph_device_name(buf, 256))
        return buf.value.decode()

    def _refresh_count(self):
        if self.device_emitter and self._ctx is not None:
            n = C.c_uint32(0)
            _lib.check(self._lib.sph_count(self._ctx, C.byref(n)))
            self.m_particlesCount = n.value
        return self.m_particlesCount

    def getParticlesCount(self):
        return self._refresh_count()

    def getMaxParticlesCount(self):
        return self.m_maxParticlesCount

    def generateParticles(self):
        """The reference fountain: 7 particles per step at the floor centre
        (src/CBaseParticleSimulator.cpp:187-210), appended on the device."""
        if self.m_scenario != FOUNTAIN or self.device_emitter:
            return
        per = 7
        if self.m_particlesCount >= (self.m_maxParticlesCount - per) % (1 << 32):
            return
        half = np.float32(H / np.float32(2.0))
        q = np.float32(-H / np.float32(4))
        oy = np.float32(-self.box[1] / np.float32(2.0))
        vy = np.float32(self.box[1] * np.float32(3.2))
        pts = np.array([[0, oy, 0], [-half, oy, 0], [half, oy, 0], [q, oy, -half], [-q, oy, -half],
                        [q, oy, half], [-q, oy, half]], np.float32)
        vel = np.tile(np.array([0, vy, 0], np.float32), (per, 1))
        self.append(pts, vel)

    def step(self):
        """CBaseParticleSimulator::step (src/CBaseParticleSimulator.cpp:116-144)."""
        self.generateParticles()
        if self.m_particlesCount == 0 and not self.device_emitter:
            return
        if self.fused:
            _lib.check(self._lib.sph_step(self._ctx, 1, None))
        else:
            self.updateGrid()
            self.updateDensityPressure()
            self.updateForces()
            self.updateCollisions()
            self.integrate()
        self.totalIteration += 1

    def advance(self, nsteps):
        """nsteps fused steps in one call (no generation); returns device ms."""
        ms = C.c_float(0.0)
        _lib.check(self._lib.sph_step(self._ctx, int(nsteps), C.byref(ms)))
        self.totalIteration += nsteps
        return ms.value

    def _phase(self, phase):
        ms = C.c_float(0.0)
        _lib.check(self._lib.sph_phase(self._ctx, phase, C.byref(ms)))
        return ms.value

    def updateGrid(self):
        return self._phase(_lib.PHASE_GRID)

    def updateDensityPressure(self):
        return self._phase(_lib.PHASE_DENSITY)

    def updateForces(self):
        return self._phase(_lib.PHASE_FORCES)

    def updateCollisions(self):
        return self._phase(_lib.PHASE_COLLISIONS)

    def integrate(self):
        return self._phase(_lib.PHASE_INTEGRATE)

    # -- state transfer -------------------------------------------------
    def ensure_context(self, capacity):
        if self._ctx is None:
            self._create(capacity)

    def upload(self, pos, vel=None, ids=None):
        pos = _lib.f32(pos, (-1, 3))
        n = len(pos)
        if self._ctx is None:
            self._create(self._capacity or n)
        vel = None if vel is None else _lib.f32(vel, (-1, 3))
        ids = None if ids is None else _lib.u32(ids)
        _lib.check(self._lib.sph_upload(self._ctx, _lib.ptr(pos), _lib.ptr(vel), _lib.ptr(ids), n))
        self.m_particlesCount = n
        self.m_maxParticlesCount = max(self.m_maxParticlesCount, n)

    def append(self, pos, vel=None, ids=None):
        pos = _lib.f32(pos, (-1, 3))
        vel = None if vel is None else _lib.f32(vel, (-1, 3))
        ids = None if ids is None else _lib.u32(ids)
        _lib.check(self._lib.sph_append(self._ctx, _lib.ptr(pos), _lib.ptr(vel), _lib.ptr(ids), len(pos)))
        self.m_particlesCount += len(pos)

    def download(self, by_id=True):
        n = self._refresh_count()
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        ids = np.zeros(n, np.uint32)
        _lib.check(self._lib.sph_download(self._ctx, _lib.ptr(pos), _lib.ptr(vel), _lib.ptr(ids), n, int(by_id)))
        return pos, vel, ids

    def readback(self, with_velocity=True):
        """Positions (and velocities) by id, written by the device into pinned
        host memory (the render boundary of integrate())."""
        return _lib.readback(self._lib, self._ctx, 0, with_velocity)

    def profile_events(self, reset=False):
        """Logged sProfilingEvents (needs FLAG_PROFILE)."""
        return _lib.profile_events(self._lib, self._ctx, reset)

    def exportLogs(self, stem, label="HIP Grid"):
        """MainWindow::exportLogs format (src/mainwindow.cpp:310-368)."""
        _lib.check(self._lib.sph_export_logs(self._ctx, str(stem).encode(), str(label).encode()))

    def grid_dims(self):
        xyz = np.zeros(3, np.int32)
        nc = C.c_uint32(0)
        _lib.check(self._lib.sph_grid_dims(self._ctx, _lib.ptr(xyz), C.byref(nc)))
        return tuple(int(v) for v in xyz), nc.value

    def debug(self, kind):
        n = self._refresh_count()
        if kind == _lib.DEBUG_CELL_START:
            out = np.zeros(self.grid_dims()[1] + 1, np.uint32)
        elif kind in (_lib.DEBUG_DENSITY, _lib.DEBUG_PRESSURE):
            out = np.zeros(n, np.float32)
        elif kind == _lib.DEBUG_ACCEL:
            out = np.zeros((n, 3), np.float32)
        else:
            out = np.zeros(n, np.uint32)
        _lib.check(self._lib.sph_debug(self._ctx, kind, _lib.ptr(out), out.nbytes))
        return out

    def sync(self):
        _lib.check(self._lib.sph_sync(self._ctx))

    def stage_times(self, reset=False):
        ms = np.zeros(8, np.float32)
        cnt = np.zeros(8, np.uint32)
        _lib.check(self._lib.sph_stage_times(self._ctx, _lib.ptr(ms), _lib.ptr(cnt), int(reset)))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(_lib.STAGES)}
