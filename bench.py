"""Benchmark of the MI355X grid-SPH step (libsph_hip.so).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dam_break_1m]
                    [--batched-steps 500] [--serve-seconds 30]

Runs W untimed steps, then times exactly K fused steps (grid + density + forces
+ walls + integrate) of the workload on one GPU with the state resident in HBM,
and prints one JSON line: particle-steps/s, ms per step, the roofline of the
dominant kernel (timed live with HIP events on the library's stream) and the
CPU oracle timed on this host on a bounded sample of the same workload.

The same line carries the batched and served forms of the step:
  * "batched": 512 independent scenes (2,048-65,536 particles each, log-uniform,
    seed 3) advanced together for 500 steps (autotuned kernel shape, HIP graph
    replay), against the same scenes stepped by plain per-scene launches (one
    context and stream per scene); both time steps 4-503 from the same start;
  * "fountain": a disc emitter adding 1,024 particles per step in the 1.6 m box
    until 524,288, 2,000 steps with the on-device emitter, and its growth phase
    against host-side generation plus a tail upload before every step;
  * "serving": Poisson arrivals of dam-break, fountain and cuboid-drop jobs
    (4,096-1,048,576 particles, 100-2,000 steps) into the native continuous-
    batching server for a bounded time; p50/p99 job latency and aggregate
    particle-steps/s.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gmu-water-simulation_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per particle per stage (compulsory reads + writes of the
# step's data flow; per-cell terms separately).
ALG_BYTES = {
    "hash": 16,       # pos r 12 + cell w 4
    "sort": 8,        # cell r 4 + permutation w 4
    "reorder": 56,    # id, pos, vel r 28 + w 28
    "density": 16,    # pos r 12 + density w 4
    "forces": 52,     # pos, vel, density r 28 + pos, vel w 24
}
ALG_BYTES_PER_CELL = 24


def host_cpu():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(pos, box, steps):
    """The CPU oracle (restated CCPUParticleSimulator, reference summation order,
    single thread) on the first `steps` steps of the same workload."""
    import oracle
    s = oracle.OracleSim(box, len(pos))
    s.add_particles(pos)
    t0 = time.perf_counter()
    s.step(steps)
    dt = time.perf_counter() - t0
    model, ncpu = host_cpu()
    return {"value": len(pos) * steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "steps 0-%d of the %d-particle workload, oracle in the reference's bucket order, "
                      "1 thread (gcc -O2 -ffp-contract=off)" % (steps - 1, len(pos)),
            "seconds": dt, "host": model, "host_threads": ncpu}


def copy_calibration(nbytes):
    """One plain device-to-device copy of a known size, so that the FETCH_SIZE /
    WRITE_SIZE counters of a profiled run can be scaled to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)


def load_traffic(workload):
    """HBM bytes per launch of each kernel from the committed counter summary
    (profiles/traffic_<workload>.json, FETCH_SIZE + WRITE_SIZE scaled by a copy
    of known size, over the steps its "steps" field names), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def bench_batched(sphhip, steps, per_scene_steps=None, nscenes=512, seed=3, variant=None):
    """The batched workload: one context for all scenes vs one context per scene.
    Both time steps 4 onwards of the same scenes (the autotuner restores the
    state it stepped); per_scene_steps None times the per-scene run over the
    same window as the batched one. variant: a kernel shape (row masks) to use
    instead of the autotuner's choice, for profiling runs that must compare
    like with like."""
    from sphhip import _lib
    lib = _lib.load()
    w = sphhip.scenes.batch_workload(nscenes=nscenes, seed=seed)
    total = sum(s["n"] for s in w)
    b = sphhip.SceneBatch([(s["box"], s["n"]) for s in w])
    for i, s in enumerate(w):
        b.upload(i, s["pos"])
    b.step(2)
    if variant is None:
        tune = b.autotune(steps=3)
    else:   # the autotuner leaves the state at step 2 as well, so the same window is timed
        b.set_variant(variant, 2)
        tune = {}
    variant, mode = b.variant()
    b.set_graphs(True)
    b.step(2)   # captures the graphs of both buffer parities (untimed)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    wall = time.perf_counter() - t0
    b.close()
    out = {"workload": "%d scenes, 2,048-65,536 particles each (log-uniform, seed %d), dam break or cuboid drop, "
                       "boxes at 1/4 fill with per-axis aspect U[0.5, 2]" % (nscenes, seed),
           "scenes": nscenes, "particles": total, "steps": steps, "value": total * steps / wall,
           "unit": "particle-steps/s", "ms_per_step": wall * 1e3 / steps, "graphs": True,
           "timed_steps": "4-%d" % (steps + 3),
           "variant": {"index": variant, "threads_unroll_pairs_stall_queue": sphhip.variant_shape(variant),
                       "neighbours": ("walk", "lists", "masks")[mode]},
           "autotune_ms_per_step": {"%d%s" % (v, "wlm"[m]): round(ms, 4) for (v, m), ms in tune.items()}}
    if per_scene_steps is None:
        per_scene_steps = steps
    if per_scene_steps:
        sims = []
        for s in w:
            sim = sphhip.HIPGridParticleSimulator(s["box"], capacity=s["n"])
            sim.upload(s["pos"])
            sims.append(sim)
        for sim in sims:
            _lib.check(lib.sph_step(sim._ctx, 4, None))
        for sim in sims:
            sim.sync()
        t0 = time.perf_counter()
        for _ in range(per_scene_steps):
            for sim in sims:
                _lib.check(lib.sph_step(sim._ctx, 1, None))
        for sim in sims:
            sim.sync()
        pw = time.perf_counter() - t0
        for sim in sims:
            sim.close()
        ps = total * per_scene_steps / pw
        out["per_scene"] = {"value": ps, "ms_per_step": pw * 1e3 / per_scene_steps, "steps": per_scene_steps,
                            "how": "one context and stream per scene, sph_step(ctx, 1) per scene per step from one "
                                   "host thread, steps 4-%d" % (per_scene_steps + 3)}
        out["speedup_vs_per_scene"] = out["value"] / ps
    return out


def bench_fountain(sphhip, steps, host_steps):
    """The fountain workload: a disc emitter at the floor centre of the 1.6 m box
    adds 1,024 particles per step until 524,288 (step 512), run for `steps`
    steps with the on-device emitter and graph replay. Beside it, the growth
    phase with the particles generated on the host and appended before every
    step (the reference's pattern, with a tail upload instead of its full
    re-upload), for the first `host_steps` steps."""
    from sphhip import _lib
    f = sphhip.scenes.FOUNTAIN
    box, rate, limit = f["box"], f["rate"], f["limit"]
    b = sphhip.SceneBatch([(box, limit)])
    b.set_emitter(0, kind=_lib.EMIT_DISC, rate=rate, limit=limit, centre=(0.0, 0.0), radius=f["radius"],
                  speed=f["speed"], spread=f["spread"], seed=f["seed"])
    b.set_graphs(True)
    b.sync()
    t0 = time.perf_counter()
    b.step(host_steps)
    b.sync()
    grow = time.perf_counter() - t0
    b.step(steps - host_steps)
    b.sync()
    wall = time.perf_counter() - t0
    count = b.count(0)
    b.close()
    live = np.minimum(np.arange(1, steps + 1, dtype=np.int64) * rate, limit)
    out = {"workload": "fountain: disc of radius 8h at the floor centre of the %.1f m box, %d particles per step with "
                       "v_y = %.1f m/s +-%d%% (seed %d) until %d, %d steps" %
                       (box, rate, f["speed"], int(f["spread"] * 100), f["seed"], limit, steps),
           "steps": steps, "final_particles": count, "value": float(live.sum()) / wall,
           "unit": "particle-steps/s", "ms_per_step": wall * 1e3 / steps, "graphs": True,
           "growth": {"steps": host_steps, "device_emitter_ms_per_step": grow * 1e3 / host_steps}}

    # Per-stage device times from HIP events on the library's stream, in a
    # separate profiled context (eager launches), over the steps after the pool
    # is full, where the particle count is constant.
    full = -(-limit // rate)
    if steps > full:
        pb = sphhip.SceneBatch([(box, limit)], flags=_lib.FLAG_PROFILE)
        pb.set_emitter(0, kind=_lib.EMIT_DISC, rate=rate, limit=limit, centre=(0.0, 0.0), radius=f["radius"],
                       speed=f["speed"], spread=f["spread"], seed=f["seed"])
        pb.step(full)
        pb.sync()
        pb.stage_times(reset=True)
        pb.step(steps - full)
        pb.sync()
        st = pb.stage_times()
        pb.close()
        stages = {k: ms / cnt for k, (ms, cnt) in st.items() if cnt}
        dominant = max(("density", "forces"), key=lambda k: stages.get(k, 0.0))
        out["stage_ms"] = stages
        out["stage_how"] = ("HIP events per stage in a second, profiled context with eager launches, steps %d-%d "
                            "(%d particles)" % (full, steps - 1, limit))
        out["dominant"] = {"kernel": dominant, "ms": stages[dominant],
                           "alg_GBs": ALG_BYTES[dominant] * limit / (stages[dominant] * 1e-3) / 1e9}

    # Host-side injection: the same number of particles per step drawn on the
    # host (uniform on the disc), appended before each step.
    rng = np.random.default_rng(f["seed"])
    floor_y = np.float32(-np.float32(box)) / np.float32(2.0)
    sim = sphhip.HIPGridParticleSimulator(box, capacity=limit)

    def batch_of(k):
        r = f["radius"] * np.sqrt(rng.random(k))
        a = 2 * np.pi * rng.random(k)
        pos = np.stack([r * np.cos(a), np.full(k, floor_y), r * np.sin(a)], 1).astype(np.float32)
        vel = np.zeros((k, 3), np.float32)
        vel[:, 1] = f["speed"] * (1 + f["spread"] * (2 * rng.random(k) - 1))
        return pos, vel

    batches = [batch_of(rate) for _ in range(host_steps)]
    sim.ensure_context(limit)
    sim.sync()
    t0 = time.perf_counter()
    for i, (pos, vel) in enumerate(batches):
        if i == 0:
            sim.upload(pos, vel)
        else:
            sim.append(pos, vel)
        sim.advance(1)
    sim.sync()
    host = time.perf_counter() - t0
    sim.close()
    out["growth"]["host_append_ms_per_step"] = host * 1e3 / host_steps
    out["growth"]["how"] = ("steps 0-%d: device emitter under graph replay vs particles drawn on the host beforehand "
                            "and appended (sph_append, %d particles) before each sph_step(ctx, 1)" %
                            (host_steps - 1, rate))
    return out


def bench_serving(sphhip, seconds, rate, budget_gb, seed=4, graphs=True):
    budget = int(budget_gb * 2 ** 30)
    submits = [0]

    def log(msg):   # every 20th submission, so that a long run reports its progress
        submits[0] += 1
        if submits[0] % 20 == 0:
            print("bench: serving %s" % msg, file=sys.stderr, flush=True)

    with sphhip.Server(budget, max_jobs=1024, chunk_steps=8, graphs=graphs) as srv:
        recs, st = sphhip.serving.poisson_run(srv, rate, seconds, seed=seed, log=log)
    lat = np.array([r["latency"] for r in recs]) if recs else np.zeros(1)
    return {"workload": "Poisson arrivals at %.2f jobs/s for %.0f s (seed %d): dam break, fountain or cuboid drop, "
                        "4,096-1,048,576 particles (log-uniform), 100-2,000 steps (uniform)" % (rate, seconds, seed),
            "jobs": len(recs), "completed": st["completed"], "failed": st["failed"],
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "max_latency_s": float(lat.max()),
            "value": st["particle_steps"] / st["elapsed_s"], "unit": "particle-steps/s",
            "busy_fraction": st["busy_s"] / st["elapsed_s"], "admin_fraction": st["admin_s"] / st["elapsed_s"],
            "graphs": bool(graphs), "hbm_budget_bytes": budget, "pool_slots": st["pool"],
            "elapsed_s": st["elapsed_s"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="dam_break_1m")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--steady-steps", type=int, default=100,
                    help="steps timed from step 200 on as well (0 skips; only when the main window ends by 200)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=500, help="0 skips the batched measurement")
    ap.add_argument("--batched-variant", type=int, default=None,
                    help="kernel shape of the batched step (row masks) instead of the autotuner's choice")
    ap.add_argument("--per-scene-steps", type=int, default=None,
                    help="steps of the per-scene comparison (default: --batched-steps; 0 skips it)")
    ap.add_argument("--fountain-steps", type=int, default=2000, help="0 skips the fountain measurement")
    ap.add_argument("--serve-seconds", type=float, default=30.0, help="0 skips the serving measurement")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="job arrivals per second")
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-graphs", type=int, default=1, help="1: the server replays its steps from HIP graphs")
    ap.add_argument("--only", choices=("batched", "fountain", "serving"),
                    help="run only this part and print its object (for profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (scales HBM counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (a scene would be "
              "sharded by spatial slabs with halo exchange, which is not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import sphhip
    from sphhip import _lib

    if args.only:
        print("bench: only the %s part" % args.only, file=sys.stderr, flush=True)
        if args.only == "batched":
            part = bench_batched(sphhip, args.batched_steps, args.per_scene_steps, variant=args.batched_variant)
        elif args.only == "fountain":
            part = bench_fountain(sphhip, args.fountain_steps, min(512, args.fountain_steps))
        else:
            part = bench_serving(sphhip, args.serve_seconds, args.serve_rate, args.serve_budget_gb,
                                 graphs=bool(args.serve_graphs))
        print(json.dumps({args.only: part}), flush=True)
        return

    pos, w = sphhip.scenes.workload(args.workload)
    n = len(pos)
    box = w["box"]
    print("bench: workload %s, %d particles, warmup %d, steps %d" % (args.workload, n, args.warmup, args.steps),
          file=sys.stderr, flush=True)

    sim = sphhip.HIPGridParticleSimulator(box, capacity=n, flags=0)
    sim.upload(pos)
    (rx, ry, rz), ncells = sim.grid_dims()
    sim.advance(args.warmup) if args.warmup else None
    sim.sync()
    t0 = time.perf_counter()
    ev_ms = sim.advance(args.steps)
    sim.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = n * args.steps / wall

    # Per-stage device times from HIP events on the library's stream, in a
    # separate profiled context over the same steps as the timed window.
    prof = sphhip.HIPGridParticleSimulator(box, capacity=n, flags=_lib.FLAG_PROFILE)
    prof.upload(pos)
    prof.advance(args.warmup) if args.warmup else None
    prof.stage_times(reset=True)
    prof.advance(args.steps)
    st = prof.stage_times()
    stages = {k: (ms / cnt if cnt else 0.0) for k, (ms, cnt) in st.items()}
    dominant = max(("density", "forces"), key=lambda k: stages[k])
    dom_ms = stages[dominant]
    achieved = ALG_BYTES[dominant] * n / (dom_ms * 1e-3) / 1e9
    step_bytes = sum(ALG_BYTES.values()) * n + ALG_BYTES_PER_CELL * ncells
    # the streaming stages against the HBM peak (algorithmic bytes / HIP-event stage time); the
    # cell table reads the sorted keys and writes one start per cell
    stream_bytes = {"hash": ALG_BYTES["hash"] * n, "sort": ALG_BYTES["sort"] * n,
                    "cells": 4 * n + 4 * (ncells + 1), "reorder": ALG_BYTES["reorder"] * n}
    stage_roofline = {}
    for name, nbytes in stream_bytes.items():
        ms = stages.get(name, 0.0)
        if ms > 0:
            gbs = nbytes / (ms * 1e-3) / 1e9
            stage_roofline[name] = {"ms": ms, "alg_bytes": nbytes, "GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    # the counter summary counts only if it covers the steps timed here
    timed = [args.warmup, args.warmup + args.steps - 1]
    traffic = load_traffic(args.workload)
    traffic_ok = bool(traffic) and traffic.get("steps") == timed
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic["kernels"].get(dominant, {}).get("bytes_per_launch") if traffic_ok else None,
                "traffic_steps": traffic.get("steps") if traffic else None,
                "avg_launch_ms": dom_ms, "timed_steps": "%d-%d" % tuple(timed),
                "alg_bytes_per_particle": ALG_BYTES[dominant],
                "step_alg_GBs": step_bytes / (ms_per_step * 1e-3) / 1e9}

    # The early window (the lattice has not collapsed yet) next to a later one:
    # the same context continued to step 200, then steps 200-299 timed.
    steady = None
    if args.steady_steps and args.warmup + args.steps <= 200:
        prof.advance(200 - (args.warmup + args.steps))
        prof.stage_times(reset=True)
        sim.advance(200 - (args.warmup + args.steps))
        sim.sync()
        t0 = time.perf_counter()
        sim.advance(args.steady_steps)
        sim.sync()
        sw = time.perf_counter() - t0
        prof.advance(args.steady_steps)
        sst = prof.stage_times()
        sst = {k: (ms / cnt if cnt else 0.0) for k, (ms, cnt) in sst.items()}
        sdom = max(("density", "forces"), key=lambda k: sst[k])
        steady = {"timed_steps": "200-%d" % (199 + args.steady_steps), "ms_per_step": sw * 1e3 / args.steady_steps,
                  "value": n * args.steady_steps / sw, "stage_ms": sst,
                  "dominant": {"kernel": sdom, "avg_launch_ms": sst[sdom],
                               "achieved": ALG_BYTES[sdom] * n / (sst[sdom] * 1e-3) / 1e9}}
    prof.close()

    line = {
        "metric": "particle-steps/s on 1M dam break; batched-serve agg throughput + p99 latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "timed_steps": "%d-%d" % tuple(timed),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": n, "lattice": list(w["dims"]), "box_m": box,
                   "grid": [rx, ry, rz], "cells": ncells, "seed": w["seed"], "parallelism": "single GPU"},
        "device": sim.getSelectedDevice(),
        "event_ms_per_step": ev_ms / args.steps,
        "stage_ms": stages,
        "stage_roofline": stage_roofline,
        "roofline": roofline,
        "steady_state": steady,
    }
    sim.close()
    if args.batched_steps:
        print("bench: batched workload, %d steps" % args.batched_steps, file=sys.stderr, flush=True)
        line["batched"] = bench_batched(sphhip, args.batched_steps, args.per_scene_steps,
                                        variant=args.batched_variant)
    if args.fountain_steps:
        print("bench: fountain workload, %d steps" % args.fountain_steps, file=sys.stderr, flush=True)
        line["fountain"] = bench_fountain(sphhip, args.fountain_steps, min(512, args.fountain_steps))
    if args.serve_seconds:
        print("bench: serving for %.0f s" % args.serve_seconds, file=sys.stderr, flush=True)
        line["serving"] = bench_serving(sphhip, args.serve_seconds, args.serve_rate, args.serve_budget_gb,
                                        graphs=bool(args.serve_graphs))
    if not args.no_cpu_baseline:
        print("bench: timing the CPU oracle on %d steps" % args.cpu_steps, file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(pos, box, args.cpu_steps)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
