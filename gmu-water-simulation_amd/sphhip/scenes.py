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


def block(origin, nx, ny, nz, count=None, jitter_h=0.01, seed=0):
    """The first `count` sites of an nx*ny*nz lattice (spacing h/2, nested y, x,
    z) whose first site is at `origin`, jittered like lattice()."""
    lib = _lib.load()
    count = nx * ny * nz if count is None else int(count)
    o = np.ascontiguousarray(np.asarray(origin, np.float32).reshape(3))
    pos = np.zeros((count, 3), np.float32)
    _lib.check(lib.sph_scene_block(_lib.ptr(o), nx, ny, nz, count, float(jitter_h), int(seed), _lib.ptr(pos)))
    return pos


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014); uniform() = (x >> 40) * 2^-24."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.state = int(seed) & self.MASK

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 40) * 2.0 ** -24


HALF = float(np.float32(H / np.float32(2.0)))


def batch_scene(rng, nmin=2048, nmax=65536, scene_seed=0, kind=None):
    """One scene of the batched workload: N log-uniform in [nmin, nmax], a dam
    break or a dropped cuboid (drawn, or `kind`), per-axis box aspect U[0.5, 2]
    with the volume scaled to N so that the particles fill about 1/4 of the box
    (the fill of the reference dam break). Returns dict(box, pos, kind, n)."""
    u = [rng.uniform() for _ in range(6)]
    n = int(round(np.exp(np.log(nmin) + u[0] * (np.log(nmax) - np.log(nmin)))))
    if kind is None:
        kind = "dam_break" if u[1] < 0.5 else "cuboid_drop"
    aspect = np.array([0.5 + 1.5 * u[2], 0.5 + 1.5 * u[3], 0.5 + 1.5 * u[4]])
    volume = 4.0 * n * HALF ** 3
    box = (aspect * (volume / aspect.prod()) ** (1.0 / 3.0)).astype(np.float32)
    if kind == "dam_break":
        # x in the first quarter of the box, all of z, as many layers in y as needed
        nx = max(1, int(np.ceil(box[0] / 4.0 / HALF)))
        nz = max(1, int(np.floor(box[2] / HALF)))
        ny_max = max(1, int(np.floor(box[1] / HALF)))
        if n > nx * ny_max * nz:      # too tall for the box: widen the column instead
            nx = int(np.ceil(n / (ny_max * nz)))
        ny = int(np.ceil(n / (nx * nz)))
        origin = -box / np.float32(2.0)
    else:
        # a block of ~0.63 of the box per axis (1/4 of its volume), centred in x and
        # z, its bottom at a random height of the free space above it
        nx = max(1, int(np.floor(0.63 * box[0] / HALF)))
        nz = max(1, int(np.floor(0.63 * box[2] / HALF)))
        ny = int(np.ceil(n / (nx * nz)))
        free = max(0.0, float(box[1]) - ny * HALF - HALF)
        origin = np.array([-box[0] / 2 + (box[0] - nx * HALF) / 2, -box[1] / 2 + HALF / 2 + u[5] * free,
                           -box[2] / 2 + (box[2] - nz * HALF) / 2], np.float32)
    pos = block(origin, nx, ny, nz, count=n, jitter_h=0.01, seed=scene_seed)
    return dict(box=box, pos=pos, kind=kind, n=n)


def batch_workload(nscenes=512, seed=3, nmin=2048, nmax=65536):
    """The batched workload: nscenes independent scenes drawn with SplitMix64(seed)."""
    rng = SplitMix64(seed)
    return [batch_scene(rng, nmin, nmax, scene_seed=(seed << 32) + i) for i in range(nscenes)]


# The fountain workload: a disc emitter of radius 8h at the floor centre adding
# 1,024 particles per step with v_y = 2.0 m/s +-10% until 524,288 particles.
FOUNTAIN = dict(box=1.6, rate=1024, limit=524288, radius=float(np.float32(8) * H), speed=2.0, spread=0.1,
                steps=2000, seed=2)

# Benchmark workloads: (lattice dims, box edge, steps, seed).
WORKLOADS = {
    "dam_break_4k": dict(dims=(16, 16, 16), box=0.4, steps=200, seed=0),
    "dam_break_1m": dict(dims=(128, 64, 128), box=3.2, steps=1000, seed=1),
}


def workload(name):
    w = WORKLOADS[name]
    return lattice(*w["dims"], box=w["box"], jitter_h=0.01, seed=w["seed"]), w
