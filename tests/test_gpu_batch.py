"""Batched scenes, HIP graphs, kernel variants, the autotuner and device
emitters of libsph_hip.so, through its C ABI.

Bars:
  * every scene of a batch is bitwise what it is alone in a context of its own
    (same positions, velocities, cells, permutation and cell table), and alone
    contexts are bitwise the canonical-order oracle (test_gpu_parity.py);
  * graph replays, every kernel variant, every neighbour mode of the force pass
    (walk, lists, row masks), and a state that went through the autotuner give
    bitwise the same trajectories;
  * the device fountain is bitwise the oracle's fountain (the reference's
    generateParticles, src/CBaseParticleSimulator.cpp:187-210), and the device
    disc emitter is bitwise a host append of the oracle's restatement of it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = np.float32(0.0457)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bits_equal(a, b, what):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    assert a.shape == b.shape, "%s: shapes %s vs %s" % (what, a.shape, b.shape)
    diff = np.nonzero(bits(a).ravel() != bits(b).ravel())[0]
    if len(diff):
        k = diff[0]
        raise AssertionError("%s: %d of %d values differ, first at flat %d: %r vs %r" %
                             (what, len(diff), a.size, k, a.ravel()[k], b.ravel()[k]))


def small_scenes(hip, count=6, seed=11, nmin=300, nmax=3000):
    rng = hip.scenes.SplitMix64(seed)
    return [hip.scenes.batch_scene(rng, nmin, nmax, scene_seed=1000 + i) for i in range(count)]


def run_alone(hip, sc, steps):
    sim = hip.HIPGridParticleSimulator(sc["box"], capacity=sc["n"])
    sim.upload(sc["pos"])
    if steps:
        sim.advance(steps)
    out = sim.download()
    sim.close()
    return out


def make_batch(hip, scs, slack=37, **kw):
    b = hip.SceneBatch([(sc["box"], sc["n"] + slack * i) for i, sc in enumerate(scs)], **kw)
    for i, sc in enumerate(scs):
        b.upload(i, sc["pos"])
    return b


def test_batch_equals_alone(hip):
    scs = small_scenes(hip)
    print("scene sizes", [sc["n"] for sc in scs], flush=True)
    with make_batch(hip, scs) as b:
        b.step(25)
        for i, sc in enumerate(scs):
            p, v, ids = b.download(i)
            pa, va, _ = run_alone(hip, sc, 25)
            assert_bits_equal(p, pa, "scene %d positions" % i)
            assert_bits_equal(v, va, "scene %d velocities" % i)
            assert np.array_equal(ids, np.arange(sc["n"]))


def test_batch_grid_tables_equal_alone(hip):
    """Cells, permutation and cell table of every scene, scene-relative."""
    lib = hip._lib
    scs = small_scenes(hip, count=4, seed=5)
    with make_batch(hip, scs) as b:
        b.phase(lib.PHASE_GRID)
        for i, sc in enumerate(scs):
            sim = hip.HIPGridParticleSimulator(sc["box"], capacity=sc["n"])
            sim.upload(sc["pos"])
            sim.updateGrid()
            for kind in (lib.DEBUG_CELL_PRESORT, lib.DEBUG_PERM, lib.DEBUG_CELL, lib.DEBUG_CELL_START,
                         lib.DEBUG_IDS):
                assert np.array_equal(b.debug(i, kind), sim.debug(kind)), "scene %d debug kind %d" % (i, kind)
            sim.close()


def test_sort_passes_across_segment_sizes(hip):
    """One scene of 41^3 = 68,921 cells (17-bit local keys: three 8-bit passes) with
    two small scenes (one or two passes on their own): the batch sorts every
    segment with the passes of the largest, by segment-local digits. Grid tables
    and 8 steps equal each scene alone."""
    lib = hip._lib
    H = hip.scenes.H
    edge = np.float32(41) * H
    half = edge / np.float32(2)
    big = dict(box=np.array([edge, edge, edge], np.float32),
               pos=hip.scenes.block((-half + H, -half + H, -half + H), 16, 16, 16, seed=21))
    big["n"] = len(big["pos"])
    scs = [big] + small_scenes(hip, count=2, seed=23)
    with make_batch(hip, scs) as b:
        b.phase(lib.PHASE_GRID)
        for i, sc in enumerate(scs):
            sim = hip.HIPGridParticleSimulator(sc["box"], capacity=sc["n"])
            sim.upload(sc["pos"])
            sim.updateGrid()
            for kind in (lib.DEBUG_PERM, lib.DEBUG_CELL, lib.DEBUG_CELL_START):
                assert np.array_equal(b.debug(i, kind), sim.debug(kind)), "scene %d debug kind %d" % (i, kind)
            sim.close()
    with make_batch(hip, scs) as b:
        b.step(8)
        for i, sc in enumerate(scs):
            p, v, _ = b.download(i)
            pa, va, _ = run_alone(hip, sc, 8)
            assert_bits_equal(p, pa, "scene %d positions" % i)
            assert_bits_equal(v, va, "scene %d velocities" % i)


def test_batch_vs_oracle(hip, orc):
    scs = small_scenes(hip, count=3, seed=21, nmin=500, nmax=1500)
    with make_batch(hip, scs) as b:
        b.step(10)
        for i, sc in enumerate(scs):
            o = orc.OracleSim(sc["box"], sc["n"])
            o.add_particles(sc["pos"])
            o.set_order_mode(orc.ORDER_CANONICAL)
            o.step(10)
            p, v, _ = b.download(i)
            assert_bits_equal(p, o.state()["pos"], "scene %d positions vs oracle" % i)
            assert_bits_equal(v, o.state()["vel"], "scene %d velocities vs oracle" % i)
            o.close()


def test_tiny_and_empty_scenes(hip, orc):
    """Scenes of 0, 1, 2 and 3 particles beside a 4,096-particle one, stepped
    together eagerly and from graphs: every live scene bitwise the oracle after
    12 steps, the empty one still empty (its segment sorts to dead keys only)."""
    H = np.float32(0.0457)
    lattice = hip.scenes.lattice(16, 16, 16, 0.4, jitter_h=0.01, seed=9)
    tiny = {1: np.array([[0.01, -0.15, 0.02]], np.float32),
            2: np.array([[0.0, 0.0, 0.0], [0.3 * H, 0.2 * H, -0.1 * H]], np.float32),
            3: np.array([[0.05, 0.1, -0.05], [0.05 + 0.5 * H, 0.1, -0.05], [0.05, 0.1 - 0.7 * H, -0.05 + 0.4 * H]],
                        np.float32)}
    scenes = [(0.4, 4), (0.3, 4), (0.2, 4), (0.25, 4), (0.4, len(lattice))]
    content = [None, tiny[1], tiny[2], tiny[3], lattice]
    for graphs in (False, True):
        with hip.SceneBatch(scenes) as b:
            for i, pos in enumerate(content):
                if pos is not None:
                    b.upload(i, pos)
            b.set_graphs(graphs)
            b.step(12)
            assert b.count(0) == 0
            for i, pos in enumerate(content):
                if pos is None:
                    continue
                o = orc.OracleSim(scenes[i][0], len(pos))
                o.add_particles(pos)
                o.set_order_mode(orc.ORDER_CANONICAL)
                o.step(12)
                p, v, _ = b.download(i)
                assert_bits_equal(p, o.state()["pos"], "scene %d (graphs %d) positions vs oracle" % (i, graphs))
                assert_bits_equal(v, o.state()["vel"], "scene %d (graphs %d) velocities vs oracle" % (i, graphs))
                o.close()


def test_many_tiny_scenes(hip, orc):
    """1,024 scenes of 1-40 particles (boxes 0.1-0.3 m), one sort tile each: the
    histograms of every pass span 1,024 segments (a multi-tile scan), and every
    scene after 6 steps is bitwise the oracle's."""
    rng = np.random.default_rng(77)
    scenes, content = [], []
    for k in range(1024):
        box = float(np.float32(rng.uniform(0.1, 0.3)))
        n = int(rng.integers(1, 41))
        pos = ((rng.random((n, 3)) - 0.5) * np.float32(box * 0.8)).astype(np.float32)
        scenes.append((box, n))
        content.append(pos)
    with hip.SceneBatch(scenes) as b:
        for i, pos in enumerate(content):
            b.upload(i, pos)
        b.step(6)
        for i in range(0, 1024, 7):
            o = orc.OracleSim(scenes[i][0], scenes[i][1])
            o.add_particles(content[i])
            o.set_order_mode(orc.ORDER_CANONICAL)
            o.step(6)
            p, v, _ = b.download(i)
            assert_bits_equal(p, o.state()["pos"], "scene %d positions vs oracle" % i)
            assert_bits_equal(v, o.state()["vel"], "scene %d velocities vs oracle" % i)
            o.close()


def test_graph_replays_equal_eager(hip):
    """Graph replays with eager work (downloads, syncs) between them, every
    replay checked against the eager trajectory."""
    scs = small_scenes(hip, count=5, seed=3)
    with make_batch(hip, scs) as eager, make_batch(hip, scs) as graph:
        graph.set_graphs(True)
        for chunk in (1, 1, 2, 5, 1, 8):
            eager.step(chunk)
            graph.step(chunk)
            for i in range(len(scs)):
                pe, ve, _ = eager.download(i)
                pg, vg, _ = graph.download(i)
                assert_bits_equal(pg, pe, "graph scene %d positions" % i)
                assert_bits_equal(vg, ve, "graph scene %d velocities" % i)


def test_single_scene_graphs_equal_eager(hip):
    pos = hip.scenes.lattice(16, 16, 16, 0.4, jitter_h=0.01, seed=0)
    a = hip.HIPGridParticleSimulator(0.4, capacity=len(pos))
    b = hip.HIPGridParticleSimulator(0.4, capacity=len(pos))
    a.upload(pos)
    b.upload(pos)
    hip._lib.check(b._lib.sph_set_graphs(b._ctx, 1))
    for chunk in (3, 1, 6):
        a.advance(chunk)
        b.advance(chunk)
        assert_bits_equal(b.download()[0], a.download()[0], "graph positions")
    a.close()
    b.close()


def test_variants_and_modes_identical(hip):
    scs = small_scenes(hip, count=3, seed=8)
    ref = None
    for mode in (hip._lib.NEIGHBOURS_WALK, hip._lib.NEIGHBOURS_LISTS, hip._lib.NEIGHBOURS_MASKS):
        for v in range(hip._lib.NUM_VARIANTS):
            with make_batch(hip, scs, max_neighbours=64) as b:
                b.set_variant(v, mode)
                assert b.variant() == (v, mode)
                b.step(8)
                got = [b.download(i)[:2] for i in range(len(scs))]
            if ref is None:
                ref = got
                continue
            for i in range(len(scs)):
                assert_bits_equal(got[i][0], ref[i][0], "variant %d mode %d scene %d pos" % (v, mode, i))
                assert_bits_equal(got[i][1], ref[i][1], "variant %d mode %d scene %d vel" % (v, mode, i))


def test_masks_with_overflowing_runs(hip, orc):
    """Row masks hold 512 candidates per row; a crowded scene (about 375 particles
    per cell, runs of 750 and more) walks those particles instead: still bitwise
    the oracle."""
    rng = np.random.default_rng(5)
    pos = (rng.random((3000, 3), dtype=np.float32) * np.float32(0.09) - np.float32(0.2)).astype(np.float32)
    with hip.SceneBatch([(0.4, len(pos))]) as b:
        b.upload(0, pos)
        b.set_variant(2, hip._lib.NEIGHBOURS_MASKS)
        b.step(3)
        p, v, _ = b.download(0)
    o = orc.OracleSim(0.4, len(pos))
    o.add_particles(pos)
    o.set_order_mode(orc.ORDER_CANONICAL)
    o.step(3)
    assert_bits_equal(p, o.state()["pos"], "crowded scene with masks")
    assert_bits_equal(v, o.state()["vel"], "crowded scene with masks (velocities)")


def test_masks_with_long_runs(hip, orc):
    """A lattice at spacing h/5 (about 125 particles per cell): runs of about
    375 candidates, 12 mask words per row, held by the masks (no walk). Every
    variant equals the oracle bitwise."""
    H = float(hip.scenes.H)
    rng = np.random.default_rng(21)
    g = np.arange(15, dtype=np.float64) * (H / 5) - 0.15
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = (pos + rng.uniform(-0.01 * H, 0.01 * H, pos.shape)).astype(np.float32)
    o = orc.OracleSim(0.4, len(pos))
    o.add_particles(pos)
    o.set_order_mode(orc.ORDER_CANONICAL)
    o.step(2)
    want = o.state()
    for v in range(hip._lib.NUM_VARIANTS):
        with hip.SceneBatch([(0.4, len(pos))]) as b:
            b.upload(0, pos)
            b.set_variant(v, hip._lib.NEIGHBOURS_MASKS)
            b.step(2)
            p, vel, _ = b.download(0)
        assert_bits_equal(p, want["pos"], "variant %d positions" % v)
        assert_bits_equal(vel, want["vel"], "variant %d velocities" % v)


def test_word_queues_with_many_words(hip, orc):
    """A lattice at spacing h/4 (about 64 particles per cell): runs of ~190
    candidates, six 32-bit mask words per row, more words per particle than any
    queue variant holds, so the queues drain several times per wave. Every queue
    variant equals the oracle bitwise."""
    H = float(hip.scenes.H)
    rng = np.random.default_rng(12)
    g = np.arange(16, dtype=np.float64) * (H / 4) - 0.15
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = (pos + rng.uniform(-0.01 * H, 0.01 * H, pos.shape)).astype(np.float32)
    o = orc.OracleSim(0.4, len(pos))
    o.add_particles(pos)
    o.set_order_mode(orc.ORDER_CANONICAL)
    o.step(3)
    want = o.state()
    for v in range(hip._lib.NUM_VARIANTS):
        if not hip.variant_shape(v)[4]:
            continue
        with hip.SceneBatch([(0.4, len(pos))]) as b:
            b.upload(0, pos)
            b.set_variant(v, hip._lib.NEIGHBOURS_MASKS)
            b.step(3)
            p, vel, _ = b.download(0)
        assert_bits_equal(p, want["pos"], "queue variant %d positions" % v)
        assert_bits_equal(vel, want["vel"], "queue variant %d velocities" % v)


def test_autotune_changes_only_the_choice(hip):
    scs = small_scenes(hip, count=4, seed=9)
    with make_batch(hip, scs, max_neighbours=64) as a, make_batch(hip, scs, max_neighbours=64) as b:
        a.step(3)
        b.step(3)
        times = a.autotune(steps=2)
        print("autotune ms/step:", {k: round(t, 4) for k, t in times.items()}, flush=True)
        assert len(times) == 3 * hip._lib.NUM_VARIANTS
        v, mode = a.variant()
        assert times[(v, mode)] == min(times.values())
        a.step(10)
        b.step(10)
        for i in range(len(scs)):
            assert_bits_equal(a.download(i)[0], b.download(i)[0], "scene %d after autotune" % i)


def test_add_and_remove_scenes(hip):
    scs = small_scenes(hip, count=4, seed=13, nmin=400, nmax=1200)
    pool = 256 * 20
    with hip.SceneBatch(pool=pool, max_scenes=3) as b:
        ids = [b.add_scene(sc["box"], sc["n"]) for sc in scs[:3]]
        assert ids == [0, 1, 2]
        for i in range(3):
            b.upload(ids[i], scs[i]["pos"])
        with pytest.raises(hip.SphError):
            b.add_scene(scs[3]["box"], scs[3]["n"])          # no free id
        b.step(5)
        b.remove_scene(1)
        d = b.add_scene(scs[3]["box"], scs[3]["n"])         # reuses id 1 and the first fitting range
        assert d == 1
        b.upload(d, scs[3]["pos"])
        info = b.info()
        assert info["scenes"] == 3 and info["pool"] == pool
        b.step(7)
        for sid, sc, steps in ((0, scs[0], 12), (2, scs[2], 12), (1, scs[3], 7)):
            p, v, _ = b.download(sid)
            pa, va, _ = run_alone(hip, sc, steps)
            assert_bits_equal(p, pa, "scene %d positions" % sid)
            assert_bits_equal(v, va, "scene %d velocities" % sid)
        with pytest.raises(hip.SphError):
            b.upload(0, np.zeros((scs[0]["n"] + 1, 3), np.float32))   # beyond its capacity


def test_device_fountain_matches_oracle(hip, orc):
    """SPH_EMIT_FOUNTAIN: the reference's 7 particles per step added on the device."""
    for fused, steps in ((True, 40), (False, 12)):
        sim = hip.HIPGridParticleSimulator(0.4, scenario=hip.FOUNTAIN, device_emitter=True, fused=fused)
        sim.setupScene()
        o = orc.OracleSim(0.4, scenario=orc.FOUNTAIN)
        o.setup_scene()
        o.set_order_mode(orc.ORDER_CANONICAL)
        for _ in range(steps):
            sim.step()
        o.step(steps)
        assert sim.getParticlesCount() == o.count == 7 * steps
        p, v, ids = sim.download()
        assert_bits_equal(p, o.state()["pos"], "device fountain positions (fused=%s)" % fused)
        assert_bits_equal(v, o.state()["vel"], "device fountain velocities (fused=%s)" % fused)
        sim.close()
        o.close()


def test_device_fountain_stops_at_limit(hip):
    sim = hip.HIPGridParticleSimulator(0.4, scenario=hip.FOUNTAIN, device_emitter=True, capacity=2000)
    sim.setupScene()
    e = hip._lib.sph_emitter()
    e.kind = hip._lib.EMIT_FOUNTAIN
    e.limit = 30          # emits while count < 23: 0, 7, 14, 21 -> 28
    import ctypes as C
    hip._lib.check(sim._lib.sph_scene_set_emitter(sim._ctx, 0, C.byref(e)))
    sim.advance(10)
    assert sim.getParticlesCount() == 28


def disc_host_reference(hip, orc, box, em, steps):
    """The same run with the emitter's particles computed by the oracle's
    restatement and appended from the host before every step."""
    sim = hip.HIPGridParticleSimulator(box, capacity=em["limit"])
    n = 0
    floor_y = np.float32(-np.float32(box)) / np.float32(2.0)
    for _ in range(steps):
        k = min(em["rate"], em["limit"] - n)
        if k > 0:
            pos, vel = orc.disc_emission(em["seed"], n, k, em["centre"], em["radius"], em["speed"], em["spread"],
                                         floor_y)
            if n == 0:
                sim.upload(pos, vel)
            else:
                sim.append(pos, vel)
            n += k
        sim.advance(1)
    return sim.download(), n


def test_disc_emitter_matches_host_append(hip, orc):
    box = 0.6
    em = dict(rate=53, limit=450, centre=(0.01, -0.02), radius=float(4 * H), speed=2.0, spread=0.1, seed=77)
    steps = 14
    for graphs in (False, True):
        with hip.SceneBatch([(box, em["limit"])]) as b:
            b.set_graphs(graphs)
            b.set_emitter(0, kind=hip._lib.EMIT_DISC, **em)
            for _ in range(steps):
                b.step(1)
            p, v, ids = b.download(0)
        (pr, vr, _), n = disc_host_reference(hip, orc, box, em, steps)
        assert len(p) == n == 450
        assert_bits_equal(p, pr, "disc emitter positions (graphs=%s)" % graphs)
        assert_bits_equal(v, vr, "disc emitter velocities (graphs=%s)" % graphs)


def test_emitter_scene_in_batch_equals_alone(hip):
    """One emitting scene among others: each scene as alone."""
    scs = small_scenes(hip, count=3, seed=17)
    em = dict(rate=40, limit=300, centre=(0.0, 0.0), radius=float(3 * H), speed=1.5, spread=0.1, seed=3)
    with hip.SceneBatch([(sc["box"], sc["n"]) for sc in scs] + [(0.5, em["limit"])]) as b:
        for i, sc in enumerate(scs):
            b.upload(i, sc["pos"])
        b.set_emitter(3, kind=hip._lib.EMIT_DISC, **em)
        b.set_graphs(True)
        b.step(9)
        for i, sc in enumerate(scs):
            assert_bits_equal(b.download(i)[0], run_alone(hip, sc, 9)[0], "scene %d" % i)
        with hip.SceneBatch([(0.5, em["limit"])]) as alone:
            alone.set_emitter(0, kind=hip._lib.EMIT_DISC, **em)
            alone.step(9)
            assert_bits_equal(b.download(3)[0], alone.download(0)[0], "emitter scene")
            assert b.count(3) == alone.count(0) == 300


def test_batch_errors(hip):
    with hip.SceneBatch([(0.4, 100)]) as b:
        with pytest.raises(hip.SphError):
            b.step(1)                       # nothing uploaded
        with pytest.raises(hip.SphError):
            b.set_variant(hip._lib.NUM_VARIANTS)
        with pytest.raises(hip.SphError):
            b.set_variant(0, hip._lib.NEIGHBOURS_LISTS)    # created without lists
        with pytest.raises(hip.SphError):
            b.download(1)                   # no scene 1
        with pytest.raises(hip.SphError):
            b.set_emitter(0, kind=7)


def test_readback_by_id_equals_download(hip):
    scs = small_scenes(hip, count=3, seed=31)
    with make_batch(hip, scs) as b:
        b.step(6)
        for i in range(len(scs)):
            p, v, _ = b.download(i)
            rp, rv = b.readback(i)
            assert_bits_equal(rp, p, "readback positions scene %d" % i)
            assert_bits_equal(rv, v, "readback velocities scene %d" % i)
    pos = hip.scenes.lattice(8, 8, 8, 0.3, seed=2)
    sim = hip.HIPGridParticleSimulator(0.3, capacity=len(pos))
    sim.upload(pos, ids=np.arange(len(pos))[::-1].copy())     # ids in reverse order
    sim.advance(3)
    p, v, _ = sim.download()
    rp, rv = sim.readback()
    assert_bits_equal(rp, p, "single-scene readback")
    sim.close()


def test_readback_by_id_equals_oracle(hip, orc):
    """sph_scene_readback of each scene of a batch against the canonical-order
    oracle's state by id (the render refresh of CGPUBaseParticleSimulator::integrate,
    src/CGPUBaseParticleSimulator.cpp:83-91)."""
    scs = small_scenes(hip, count=3, seed=41)
    with make_batch(hip, scs) as b:
        b.step(7)
        for i, sc in enumerate(scs):
            rp, rv = b.readback(i)
            o = orc.OracleSim(sc["box"], sc["n"])
            o.add_particles(sc["pos"])
            o.set_order_mode(orc.ORDER_CANONICAL)
            o.step(7)
            st = o.state()
            assert_bits_equal(rp, st["pos"], "readback positions scene %d" % i)
            assert_bits_equal(rv, st["vel"], "readback velocities scene %d" % i)


def test_readback_rejects_bad_ids(hip):
    pos = hip.scenes.lattice(4, 4, 4, 0.3, seed=2)
    sim = hip.HIPGridParticleSimulator(0.3, capacity=len(pos))
    ids = np.arange(len(pos), dtype=np.uint32)
    ids[5] = 1000                                            # outside 0..n-1: not written, reported
    sim.upload(pos, ids=ids)
    with pytest.raises(hip.SphError):
        sim.readback()
    sim.close()


def test_profile_events_and_export(hip, tmp_path):
    lib = hip._lib
    pos = hip.scenes.lattice(12, 12, 12, 0.4, seed=3)
    for fused in (True, False):
        sim = hip.HIPGridParticleSimulator(0.4, capacity=len(pos), flags=lib.FLAG_PROFILE, fused=fused)
        sim.upload(pos)
        for _ in range(25):
            sim.step()
        ev = sim.profile_events()
        assert [e["iteration"] for e in ev] == [0, 10, 20]
        for e in ev:
            assert e["update_grid"] > 0 and e["update_density_pressure"] > 0 and e["update_forces"] > 0
            if fused:
                assert e["update_collisions"] == 0 and e["integrate"] == 0
            else:
                assert e["update_collisions"] > 0 and e["integrate"] > 0
        stem = tmp_path / ("fused" if fused else "phases")
        sim.exportLogs(stem, "HIP Grid")
        rows = (tmp_path / (stem.name + "_detail.csv")).read_text().splitlines()
        assert rows[0] == "HIP Grid" and rows[1].startswith("Grid;") and rows[5].startswith("Integrate;")
        assert len(rows[2].rstrip(";").split(";")) == 4                   # label + 3 events
        sim.close()


def test_batch_stage_times(hip):
    """Per-stage HIP-event times of a batch context: one launch of each step
    stage per step for all scenes together, cleared by a reset; a context
    without FLAG_PROFILE refuses."""
    lib = hip._lib
    scenes = [hip.scenes.lattice(n, n, n, 0.4, seed=s) for n, s in ((8, 1), (11, 2))]
    with hip.SceneBatch([(0.4, len(p)) for p in scenes], flags=lib.FLAG_PROFILE) as b:
        for i, p in enumerate(scenes):
            b.upload(i, p)
        b.step(3)
        b.stage_times(reset=True)
        b.step(5)
        st = b.stage_times()
        for name in ("sort", "cells", "reorder", "density", "forces"):
            ms, cnt = st[name]
            assert cnt == 5 and ms > 0, (name, st[name])
        assert b.stage_times(reset=True) == st
        assert all(v == (0.0, 0) for v in b.stage_times().values())
    with hip.SceneBatch([(0.4, len(scenes[0]))]) as b:
        with pytest.raises(lib.SphError):
            b.stage_times()


def test_full_rate_fountain_emitter_matches_host_append(hip, orc):
    """The fountain workload's emitter at its full rate (1,024 particles per step
    on a disc of radius 8h in the 1.6 m box), with graph replay: bitwise a host
    append of the oracle's restatement, for 12 steps (12,288 particles)."""
    f = hip.scenes.FOUNTAIN
    em = dict(rate=f["rate"], limit=f["limit"], centre=(0.0, 0.0), radius=f["radius"], speed=f["speed"],
              spread=f["spread"], seed=f["seed"])
    steps = 12
    with hip.SceneBatch([(f["box"], f["limit"])]) as b:
        b.set_emitter(0, kind=hip._lib.EMIT_DISC, **em)
        b.set_graphs(True)
        b.step(steps)
        p, v, _ = b.download(0)
    (pr, vr, _), n = disc_host_reference(hip, orc, f["box"], em, steps)
    assert len(p) == n == steps * f["rate"]
    assert_bits_equal(p, pr, "fountain positions")
    assert_bits_equal(v, vr, "fountain velocities")


def test_full_batch_scenes_equal_alone(hip):
    """The 512-scene batched workload (9.65 M particles), 3 steps: every scene
    keeps its particles (ids a permutation of 0..n-1) and sampled scenes are
    bitwise their solo runs."""
    ws = hip.scenes.batch_workload()
    with hip.SceneBatch([(s["box"], s["n"]) for s in ws]) as b:
        for i, s in enumerate(ws):
            b.upload(i, s["pos"])
        b.step(3)
        for i, s in enumerate(ws):
            ids = b.debug(i, hip._lib.DEBUG_IDS)
            assert np.array_equal(np.sort(ids), np.arange(s["n"], dtype=np.uint32)), "scene %d ids" % i
        for i in (0, 1, 100, 257, 511):
            p, v, _ = b.download(i)
            pa, va, _ = run_alone(hip, ws[i], 3)
            assert_bits_equal(p, pa, "scene %d positions" % i)
            assert_bits_equal(v, va, "scene %d velocities" % i)

