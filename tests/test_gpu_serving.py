"""The serving loop (sph_server_* of libsph_hip.so): continuous batching of
"advance scene S by k steps" jobs. Bar: every job's result is bitwise its result
alone (a context of its own advanced by the same k steps), whatever else shared
the batch, whenever it was admitted, and however the pool forced it to queue."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def assert_bits_equal(a, b, what):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, "%s: shapes %s vs %s" % (what, a.shape, b.shape)
    bad = np.nonzero(a.view(np.uint32).ravel() != b.view(np.uint32).ravel())[0]
    assert len(bad) == 0, "%s: %d values differ, first %r vs %r" % (what, len(bad), a.ravel()[bad[0]],
                                                                     b.ravel()[bad[0]])


def alone(hip, job):
    with hip.SceneBatch([(job["box"], job["capacity"])]) as b:
        if len(job["pos"]):
            b.upload(0, job["pos"])
        if job["emitter"] is not None:
            b.set_emitter(0, **job["emitter"])
        b.step(job["steps"])
        p, v, _ = b.download(0)
    return p, v


def small_jobs(hip, count, seed):
    rng = hip.scenes.SplitMix64(seed)
    return [hip.serving.serving_job(rng, nmin=300, nmax=2500, smin=3, smax=30, seed=100 + i) for i in range(count)]


def test_jobs_in_the_mix_equal_alone(hip):
    jobs = small_jobs(hip, 10, seed=7)
    print("jobs", [(j["kind"], j["n"], j["steps"]) for j in jobs], flush=True)
    budget = hip.serving.slot_bytes(0) * 9000     # room for a few jobs at a time: the rest queue
    with hip.Server(budget, max_jobs=4, chunk_steps=4) as srv:
        ids = []
        for k, job in enumerate(jobs):
            ids.append(srv.submit(job["box"], job["pos"], job["steps"], capacity=job["capacity"],
                                  emitter=job["emitter"]))
            if k % 3 == 2:
                time.sleep(0.01)                    # arrivals while others run
        for jid, job in zip(ids, jobs):
            info = srv.wait(jid, timeout=120.0)
            assert info.state == hip.serving.JOB_DONE
            p, v = srv.result(jid)
            pa, va = alone(hip, job)
            assert_bits_equal(p, pa, "job %d (%s) positions" % (jid, job["kind"]))
            assert_bits_equal(v, va, "job %d (%s) velocities" % (jid, job["kind"]))
        st = srv.stats()
        print("stats", st, flush=True)
        assert st["completed"] == len(jobs) and st["failed"] == 0
        assert st["batch_steps"] > 0 and st["particle_steps"] > 0


def test_admission_against_the_budget(hip):
    per = hip.serving.slot_bytes(0)
    with hip.Server(per * 1500, max_jobs=8, chunk_steps=2) as srv:
        pos = hip.scenes.lattice(10, 10, 10, 0.3, seed=1)
        with pytest.raises(hip.SphError):
            srv.submit(0.6, np.zeros((3000, 3), np.float32), 5)      # can never fit the pool
        a = srv.submit(0.3, pos, 6)
        b = srv.submit(0.3, pos, 6)                                   # waits for a's slots
        ia, ib = srv.wait(a, 60.0), srv.wait(b, 60.0)
        assert ib.admit_s >= ia.done_s - 1e-6
        pa, _ = srv.result(a)
        pb, _ = srv.result(b)
        assert_bits_equal(pa, pb, "same job twice")
        with pytest.raises(hip.SphError):
            srv.result(a)                                             # already collected
        with pytest.raises(hip.SphError):
            srv.wait(12345, 1.0)                                      # unknown job
