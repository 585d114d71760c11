"""Server: continuous batching of "advance this scene by k steps" jobs
(sph_server_* in include/sph_hip.h). The scheduler runs natively inside
libsph_hip.so; this module submits jobs, waits for them, and generates the
serving workload (Poisson arrivals of dam-break, fountain and cuboid-drop jobs).
"""
import ctypes as C
import time

import numpy as np

from . import _lib
from . import scenes
from .batch import scene_desc

JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = range(4)


class sph_server_params(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_jobs", C.c_uint32), ("chunk_steps", C.c_uint32),
                ("max_neighbours", C.c_uint32), ("flags", C.c_uint32), ("variant", C.c_int32),
                ("graphs", C.c_uint32), ("max_bypass", C.c_uint32)]


class sph_job_desc(C.Structure):
    _fields_ = [("scene", _lib.sph_scene_desc), ("pos3", C.c_void_p), ("vel3", C.c_void_p), ("n", C.c_uint32),
                ("steps", C.c_uint32), ("emitter", C.c_void_p)]


class sph_job_info(C.Structure):
    _fields_ = [("state", C.c_int), ("count", C.c_uint32), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double)]


class sph_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("queued", C.c_uint64), ("running", C.c_uint64), ("particle_steps", C.c_uint64),
                ("batch_steps", C.c_uint64), ("busy_s", C.c_double), ("elapsed_s", C.c_double),
                ("pool", C.c_uint32), ("free_slots", C.c_uint32), ("admin_s", C.c_double)]


def _load():
    return _lib.load()


def slot_bytes(max_neighbours=0):
    return int(_load().sph_slot_bytes(int(max_neighbours)))


class Server:
    """A serving loop over one batch context of `hbm_budget` device bytes."""

    def __init__(self, hbm_budget, max_jobs=1024, chunk_steps=8, max_neighbours=0, variant=-1, graphs=False,
                 flags=0, max_bypass=0):
        self._lib = _load()
        p = sph_server_params(int(hbm_budget), int(max_jobs), int(chunk_steps), int(max_neighbours), int(flags),
                              int(variant), int(bool(graphs)), int(max_bypass))
        srv = C.c_void_p()
        _lib.check(self._lib.sph_server_create(C.byref(p), C.byref(srv)))
        self._srv = srv

    def close(self):
        if self._srv is not None:
            self._lib.sph_server_destroy(self._srv)
            self._srv = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, box, pos, steps, vel=None, capacity=None, emitter=None, gravity=(0.0, -9.80665, 0.0),
               dt=0.01):
        """Queue a job; returns its id. emitter: None or a dict of SceneBatch.set_emitter arguments."""
        pos = _lib.f32(pos, (-1, 3)) if pos is not None and len(pos) else np.zeros((0, 3), np.float32)
        vel = None if vel is None else _lib.f32(vel, (-1, 3))
        n = len(pos)
        cap = int(capacity if capacity is not None else max(n, 1))
        d = sph_job_desc()
        d.scene = scene_desc(box, cap, gravity, dt)
        d.pos3 = pos.ctypes.data if n else None
        d.vel3 = vel.ctypes.data if vel is not None else None
        d.n = n
        d.steps = int(steps)
        e = None
        if emitter is not None:
            e = _lib.sph_emitter()
            e.kind = int(emitter.get("kind", _lib.EMIT_DISC))
            e.rate = int(emitter.get("rate", 0))
            e.limit = int(emitter.get("limit", cap))
            c = emitter.get("centre", (0.0, 0.0))
            e.centre[0], e.centre[1] = float(c[0]), float(c[1])
            e.radius = float(emitter.get("radius", 0.0))
            e.speed = float(emitter.get("speed", 2.0))
            e.spread = float(emitter.get("spread", 0.1))
            e.seed = int(emitter.get("seed", 0))
            d.emitter = C.cast(C.pointer(e), C.c_void_p)
        jid = C.c_uint64(0)
        _lib.check(self._lib.sph_server_submit(self._srv, C.byref(d), C.byref(jid)))
        return jid.value

    def info(self, job):
        i = sph_job_info()
        _lib.check(self._lib.sph_server_job(self._srv, int(job), C.byref(i)))
        return i

    def wait(self, job, timeout=-1.0):
        i = sph_job_info()
        _lib.check(self._lib.sph_server_wait(self._srv, int(job), float(timeout), C.byref(i)))
        return i

    def result(self, job):
        """Positions and velocities of a done job, by particle id (frees the job)."""
        n = self.info(job).count
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        _lib.check(self._lib.sph_server_result(self._srv, int(job), _lib.ptr(pos), _lib.ptr(vel), n))
        return pos, vel

    def stats(self):
        s = sph_server_stats()
        _lib.check(self._lib.sph_server_stats_get(self._srv, C.byref(s)))
        return {name: getattr(s, name) for name, _ in sph_server_stats._fields_}


# ------------------------------------------------------------------ workload

def serving_job(rng, nmin=4096, nmax=1048576, smin=100, smax=2000, seed=0):
    """One serving job: a dam break, a fountain or a dropped cuboid with N
    log-uniform in [nmin, nmax] and a step count uniform in [smin, smax]. Boxes
    hold N at about 1/4 fill (as the batched workload); a fountain starts empty
    and a floor-disc emitter adds max(64, N/256) particles per step up to N."""
    kind = ("dam_break", "fountain", "cuboid_drop")[min(2, int(rng.uniform() * 3))]
    steps = smin + min(smax - smin, int(rng.uniform() * (smax - smin + 1)))
    if kind == "fountain":
        n = int(round(np.exp(np.log(nmin) + rng.uniform() * (np.log(nmax) - np.log(nmin)))))
        edge = np.float32((4.0 * n * scenes.HALF ** 3) ** (1.0 / 3.0))
        box = np.array([edge, edge, edge], np.float32)
        em = dict(kind=_lib.EMIT_DISC, rate=max(64, n // 256), limit=n, centre=(0.0, 0.0),
                  radius=float(min(np.float32(8) * scenes.H, edge / 4)), speed=2.0, spread=0.1, seed=seed)
        return dict(kind=kind, box=box, pos=np.zeros((0, 3), np.float32), n=n, steps=steps, emitter=em, capacity=n)
    sc = scenes.batch_scene(rng, nmin, nmax, scene_seed=seed, kind=kind)
    return dict(kind=kind, box=sc["box"], pos=sc["pos"], n=sc["n"], steps=steps, emitter=None, capacity=sc["n"])


def poisson_run(server, rate_hz, duration_s, seed=4, nmin=4096, nmax=1048576, smin=100, smax=2000, log=None,
                max_jobs=None):
    """Generate the jobs, submit them at Poisson arrival times (rate_hz) for
    duration_s seconds, wait for all of them, and return per-job records and the
    server's statistics. Latency = completion - submission on the server's clock."""
    rng = scenes.SplitMix64(seed)
    jobs = []
    submitted = []
    pending = []
    t0 = time.perf_counter()
    t_next = 0.0
    k = 0
    while t_next < duration_s and (max_jobs is None or k < max_jobs):
        jobs.append((t_next, serving_job(rng, nmin, nmax, smin, smax, seed=(seed << 32) + k)))
        t_next += -np.log(max(1e-12, 1.0 - rng.uniform())) / rate_hz
        k += 1
    for t_arr, job in jobs:
        delay = t_arr - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        jid = server.submit(job["box"], job["pos"], job["steps"], capacity=job["capacity"], emitter=job["emitter"])
        submitted.append((jid, job))
        if log:
            log("t=%.2fs submit job %d: %s N=%d steps=%d" % (time.perf_counter() - t0, jid, job["kind"], job["n"],
                                                           job["steps"]))
    records = []
    for jid, job in submitted:
        info = server.wait(jid)
        records.append(dict(id=jid, kind=job["kind"], n=job["n"], steps=job["steps"], latency=info.done_s - info.submit_s,
                            queue=info.admit_s - info.submit_s, count=info.count))
        pending.append(jid)
    for jid in pending:
        server.result(jid)
    return records, server.stats()
