"""Scene generators of libsph_hip.so (host code, no GPU needed)."""
import ctypes as C

import numpy as np

from . import _lib

H = np.float32(0.0457)  # include/CParticle.h:80


def grid_res(box):
    """(int) ceil(box / h), src/CBaseParticleSimulator.cpp:27-31."""
    return int(_lib.load().sph_grid_res(float(box)))


def dam_break(box):
    """The reference's dam-break lattice for a cubic box, id order
    (CBaseParticleSimulator::setupScene, src/CBaseParticleSimulator.cpp:38-65)."""
    lib = _lib.load()
    n = C.c_uint32(0)
    _lib.check(lib.sph_scene_dam_break(float(box), None, 0, C.byref(n)))
    pos = np.zeros((n.value, 3), np.float32)
    _lib.check(lib.sph_scene_dam_break(float(box), _lib.ptr(pos), n.value, C.byref(n)))
    return pos


def lattice(nx, ny, nz, box, jitter_h=0.01, seed=0):
    """Benchmark lattice: nx*ny*nz sites at spacing h/2 from the corner -box/2,
    jittered by +-jitter_h*h (SplitMix64 with the given seed)."""
    lib = _lib.load()
    b = np.ascontiguousarray(np.broadcast_to(np.asarray(box, np.float32), (3,)))
    pos = np.zeros((nx * ny * nz, 3), np.float32)
    _lib.check(lib.sph_scene_lattice(nx, ny, nz, _lib.ptr(b), float(jitter_h), int(seed), _lib.ptr(pos)))
    return pos


# Benchmark workloads: (lattice dims, box edge, steps, seed).
WORKLOADS = {
    "dam_break_4k": dict(dims=(16, 16, 16), box=0.4, steps=200, seed=0),
    "dam_break_1m": dict(dims=(128, 64, 128), box=3.2, steps=1000, seed=1),
}


def workload(name):
    w = WORKLOADS[name]
    return lattice(*w["dims"], box=w["box"], jitter_h=0.01, seed=w["seed"]), w
