"""The serving loop (sph_server_* of libsph_hip.so): continuous batching of
"advance scene S by k steps" jobs, which replaces the reference's QTimer/doWork
driver of one scene (src/CBaseParticleSimulator.cpp:35, 146-152). Bar: every
job's result is bitwise its result alone (a context of its own advanced by the
same k steps), whatever else shared the batch, whenever it was admitted, and
however the pool forced it to queue; and bitwise the canonical-order oracle
(the restated CCPUParticleSimulator) where the oracle finishes in seconds."""
import os
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


def oracle_job(orc, job):
    """The job on the oracle in the canonical order: the scene uploaded, or for a
    fountain job the disc emitter's particles (the oracle's restatement of it)
    appended before every step as the device emitter adds them."""
    o = orc.OracleSim(job["box"], job["capacity"])
    o.set_order_mode(orc.ORDER_CANONICAL)
    em = job["emitter"]
    if em is None:
        o.add_particles(job["pos"])
        o.step(job["steps"])
    else:
        floor_y = np.float32(-np.float32(job["box"][1])) / np.float32(2.0)
        for _ in range(job["steps"]):
            k = min(em["rate"], em["limit"] - o.count)
            if k > 0:
                pos, vel = orc.disc_emission(em["seed"], o.count, k, em["centre"], em["radius"], em["speed"],
                                             em["spread"], floor_y)
                o.add_particles(pos, vel)
            o.step(1)
    st = o.state()
    return st["pos"], st["vel"]


def small_jobs(hip, count, seed):
    rng = hip.scenes.SplitMix64(seed)
    return [hip.serving.serving_job(rng, nmin=300, nmax=2500, smin=3, smax=30, seed=100 + i) for i in range(count)]


def test_jobs_in_the_mix_equal_alone(hip, orc):
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
            po, vo = oracle_job(orc, job)
            assert_bits_equal(p, po, "job %d (%s) positions against the oracle" % (jid, job["kind"]))
            assert_bits_equal(v, vo, "job %d (%s) velocities against the oracle" % (jid, job["kind"]))
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


def test_served_workload_jobs_at_full_size(hip, orc):
    """The first 9 jobs of the serving workload's arrival stream (seed 4): dam
    breaks, fountains and dropped cuboids of 8,631-929,547 particles and 201-1,796
    steps, through a pool of 1.4 M slots, so that the largest jobs queue while
    smaller ones behind them are admitted past them. Every job is bitwise its
    solo run; the two whose oracle run takes seconds (n x steps <= 6e6) are
    bitwise the oracle too."""
    rng = hip.scenes.SplitMix64(4)
    jobs = []
    for k in range(9):
        jobs.append(hip.serving.serving_job(rng, 4096, 1048576, 100, 2000, seed=(4 << 32) + k))
        rng.uniform()   # the stream's arrival-time draw
    print("jobs", [(j["kind"], j["n"], j["steps"]) for j in jobs], flush=True)
    assert max(j["n"] for j in jobs) >= 262144 and min(j["n"] for j in jobs) >= 4096
    assert min(j["steps"] for j in jobs) >= 100
    budget = hip.serving.slot_bytes(0) * 1400000
    with hip.Server(budget, max_jobs=16, chunk_steps=8) as srv:
        ids = [srv.submit(j["box"], j["pos"], j["steps"], capacity=j["capacity"], emitter=j["emitter"]) for j in jobs]
        infos = [srv.wait(i, timeout=110.0) for i in ids]
        results = [srv.result(i) for i in ids]
        st = srv.stats()
    print("stats", st, flush=True)
    assert st["completed"] == len(jobs) and st["failed"] == 0
    # the queue was exercised: some job was admitted after a later-submitted one
    admits = [i.admit_s for i in infos]
    assert any(admits[a] > admits[b] for a in range(len(jobs)) for b in range(a + 1, len(jobs)))
    for jid, job, (p, v) in zip(ids, jobs, results):
        print("check job %d: %s N=%d steps=%d" % (jid, job["kind"], job["n"], job["steps"]), flush=True)
        pa, va = alone(hip, job)
        assert_bits_equal(p, pa, "job %d (%s) positions" % (jid, job["kind"]))
        assert_bits_equal(v, va, "job %d (%s) velocities" % (jid, job["kind"]))
        if job["n"] * job["steps"] <= 6e6:
            po, vo = oracle_job(orc, job)
            assert_bits_equal(p, po, "job %d (%s) positions against the oracle" % (jid, job["kind"]))
            assert_bits_equal(v, vo, "job %d (%s) velocities against the oracle" % (jid, job["kind"]))


def test_backfill_is_bounded(hip):
    """A job that does not fit yet is passed by smaller later jobs only max_bypass
    times; then admission waits for it."""
    per = hip.serving.slot_bytes(0)
    small = hip.scenes.lattice(6, 6, 6, 0.3, seed=3)            # 216 particles: one 256-slot group
    big = hip.scenes.lattice(13, 12, 12, 0.6, seed=3)           # 1,872 particles
    with hip.Server(per * 2048, max_jobs=16, chunk_steps=2, max_bypass=2) as srv:
        first = srv.submit(0.3, small, 40)                        # holds 256 of the 2,048 slots for 40 steps
        blocked = srv.submit(0.6, big, 4)                         # needs 1,872 of the 1,792 left: waits
        later = [srv.submit(0.3, small, 4) for _ in range(6)]
        infos = {j: srv.wait(j, 60.0) for j in [first, blocked] + later}
        for j in [first, blocked] + later:
            srv.result(j)
    # exactly max_bypass later jobs went before the blocked one; the rest after it
    before = [j for j in later if infos[j].admit_s < infos[blocked].admit_s]
    assert len(before) == 2, [(infos[j].admit_s, infos[blocked].admit_s) for j in later]


def test_failed_jobs_free_their_slots_and_are_collected(hip, monkeypatch):
    """A chunk that fails (SPH_FAULT_AT_STEP injects a failure at the context's 3rd
    step, from the host) fails its jobs; their slots go back to the pool and the
    failed jobs are collected, with the error, by sph_server_result."""
    monkeypatch.setenv("SPH_FAULT_AT_STEP", "3")
    per = hip.serving.slot_bytes(0)
    pos = hip.scenes.lattice(8, 8, 8, 0.3, seed=1)
    with hip.Server(per * 4096, max_jobs=4, chunk_steps=2) as srv:
        monkeypatch.delenv("SPH_FAULT_AT_STEP")
        a = srv.submit(0.3, pos, 10)
        b = srv.submit(0.3, pos, 10)
        for j in (a, b):
            with pytest.raises(hip.SphError) as e:
                srv.wait(j, 60.0)
            assert e.value.status == hip._lib.SPH_ERR_FAULT
        st = srv.stats()
        assert st["failed"] == 2 and st["free_slots"] == st["pool"]
        for j in (a, b):
            with pytest.raises(hip.SphError) as e:
                srv.result(j)
            assert e.value.status == hip._lib.SPH_ERR_FAULT and b"injected" in str(e.value).encode()
            with pytest.raises(hip.SphError):
                srv.info(j)                                       # forgotten
