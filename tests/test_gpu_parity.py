"""Parity of the HIP grid-SPH path (libsph_hip.so through its C ABI) with the
CPU oracle (the restated CCPUParticleSimulator, src/CCPUParticleSimulator.cpp).

Bars (stated per test):
  * cell ids, sort permutation, cell table, neighbour counts: bit-exact;
  * density, acceleration, positions and velocities: bit-exact against the
    oracle in its canonical summation order (the order the HIP path is
    specified in), for single phases and over 100+ steps;
  * the canonical order against the oracle in the reference's own summation
    order: density within 1e-5 relative, acceleration within 1e-5 * S_i per
    component (S_i = the sum of absolute force terms, emitted by the oracle),
    positions within 1e-3*h for the first 10 steps (the reference-order mode is
    bit-exact against it: test_gpu_reference_order.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = np.float32(0.0457)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bits_equal(a, b, what):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    assert a.shape == b.shape, what
    diff = np.nonzero(bits(a).ravel() != bits(b).ravel())[0]
    if len(diff):
        k = diff[0]
        raise AssertionError("%s: %d of %d values differ, first at flat %d: %r vs %r" %
                             (what, len(diff), a.size, k, a.ravel()[k], b.ravel()[k]))


def gpu_sim(hip, box, pos, vel=None, diag=True, capacity=None, max_neighbours=0):
    flags = hip._lib.FLAG_DIAGNOSTICS if diag else 0
    sim = hip.HIPGridParticleSimulator(box, flags=flags, capacity=capacity or len(pos), max_neighbours=max_neighbours)
    sim.upload(pos, vel)
    return sim


def oracle_sim(orc, box, pos, vel=None, canonical=True, capacity=0):
    s = orc.OracleSim(box, capacity or len(pos))
    s.add_particles(pos, vel)
    s.set_order_mode(orc.ORDER_CANONICAL if canonical else orc.ORDER_REFERENCE)
    s.set_diagnostics(True)
    return s


def lattice_4k(hip):
    return hip.scenes.lattice(16, 16, 16, 0.4, jitter_h=0.01, seed=0)


def test_device_is_gfx950(hip):
    sim = hip.HIPGridParticleSimulator(0.4)
    name = sim.getSelectedDevice()
    print("device:", name)
    assert "gfx950" in name


def test_phases_bit_exact_dam_break_4k(hip, orc):
    """Every phase of one step on the 4,096-particle lattice (box 0.4)."""
    pos = lattice_4k(hip)
    n = len(pos)
    sim = gpu_sim(hip, 0.4, pos)
    o = oracle_sim(orc, 0.4, pos)
    (rx, ry, rz), ncells = sim.grid_dims()
    assert (rx, ry, rz) == (9, 9, 9) and ncells == 729

    sim.updateGrid()
    o.update_grid()
    pre = sim.debug(hip._lib.DEBUG_CELL_PRESORT)
    np.testing.assert_array_equal(pre, orc.hash_positions(pos, 0.4, 9))
    np.testing.assert_array_equal(pre, o.cells())
    perm = sim.debug(hip._lib.DEBUG_PERM)
    np.testing.assert_array_equal(perm, np.argsort(pre, kind="stable"))
    ids = sim.debug(hip._lib.DEBUG_IDS)
    np.testing.assert_array_equal(ids, perm)
    np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_CELL), pre[perm])
    start = sim.debug(hip._lib.DEBUG_CELL_START)
    np.testing.assert_array_equal(start, np.searchsorted(pre[perm], np.arange(ncells + 1), side="left"))
    ostart, oids = o.buckets()
    np.testing.assert_array_equal(oids, ids)
    np.testing.assert_array_equal(ostart, start)

    sim.updateDensityPressure()
    o.update_density_pressure()
    st = o.state()
    np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_NEIGHBOURS), o.neighbour_counts()[ids])
    assert_bits_equal(sim.debug(hip._lib.DEBUG_DENSITY), st["density"][ids], "density")
    assert_bits_equal(sim.debug(hip._lib.DEBUG_PRESSURE), st["pressure"][ids], "pressure")

    sim.updateForces()
    sim.updateCollisions()
    o.update_forces()
    st = o.state()
    assert_bits_equal(sim.debug(hip._lib.DEBUG_ACCEL), st["acc"][ids], "acceleration")

    sim.integrate()
    o.integrate()
    st = o.state()
    p, v, _ = sim.download(by_id=True)
    assert_bits_equal(p, st["pos"], "position")
    assert_bits_equal(v, st["vel"], "velocity")
    assert n == 4096


def test_single_step_vs_reference_order(hip, orc):
    """Against the oracle in the reference's own bucket order: density within 1e-5
    relative, acceleration within 1e-5 * S_i per component."""
    pos = lattice_4k(hip)
    sim = gpu_sim(hip, 0.4, pos)
    o = oracle_sim(orc, 0.4, pos, canonical=False)
    o.update_grid()
    o.update_density_pressure()
    o.update_forces()
    st = o.state()
    scale = o.force_scale()
    sim.updateGrid()
    sim.updateDensityPressure()
    sim.updateForces()
    sim.updateCollisions()
    ids = sim.debug(hip._lib.DEBUG_IDS)
    rho = sim.debug(hip._lib.DEBUG_DENSITY)
    acc = sim.debug(hip._lib.DEBUG_ACCEL)
    np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_NEIGHBOURS), o.neighbour_counts()[ids])
    rel = np.abs(rho.astype(np.float64) - st["density"][ids]) / np.abs(st["density"][ids])
    assert rel.max() <= 1e-5, rel.max()
    err = np.abs(acc.astype(np.float64) - st["acc"][ids]) / scale[ids]
    assert err.max() <= 1e-5, err.max()


def test_fused_step_equals_phases(hip, orc):
    """sph_step (forces + walls + integrate fused) is bitwise the five phases."""
    pos = lattice_4k(hip)
    a = gpu_sim(hip, 0.4, pos)
    b = gpu_sim(hip, 0.4, pos)
    for _ in range(5):
        a.step()
        b.updateGrid(); b.updateDensityPressure(); b.updateForces(); b.updateCollisions(); b.integrate()
    pa, va, _ = a.download()
    pb, vb, _ = b.download()
    assert_bits_equal(pa, pb, "position")
    assert_bits_equal(va, vb, "velocity")


def test_100_steps_bit_exact_canonical(hip, orc):
    """100 fused steps of the 4,096-particle dam break in the canonical order:
    bitwise equal to the canonical-order oracle at steps 10 and 100. Against the
    reference-order oracle the canonical order stays within 1e-3*h only for the
    first 10 steps; the 1e-3*h bound at step 100 is met by the reference-order
    mode (test_gpu_reference_order.py), whose distance there is 0."""
    pos = lattice_4k(hip)
    sim = gpu_sim(hip, 0.4, pos, diag=False)
    can = oracle_sim(orc, 0.4, pos)
    ref = oracle_sim(orc, 0.4, pos, canonical=False)
    sim.advance(10)
    can.step(10)
    ref.step(10)
    p10, _, _ = sim.download()
    assert_bits_equal(p10, can.state()["pos"], "positions at step 10")
    d10 = np.abs(p10.astype(np.float64) - ref.state()["pos"]).max()
    assert d10 <= 1e-3 * float(H), d10
    sim.advance(90)
    can.step(90)
    ref.step(90)
    p100, v100, _ = sim.download()
    cs = can.state()
    assert_bits_equal(p100, cs["pos"], "positions at step 100")
    assert_bits_equal(v100, cs["vel"], "velocities at step 100")
    d100 = np.abs(p100.astype(np.float64) - ref.state()["pos"]).max()
    print("step 100: |gpu(canonical order) - oracle(reference order)| = %.3e m = %.2f h (summation-order chaos)"
          % (d100, d100 / float(H)))


def test_reference_dam_break_scene(hip, orc):
    """The reference's own scene (setupScene, box 0.4 m, 1,620 particles), 30 steps."""
    sim = hip.HIPGridParticleSimulator(0.4, flags=0)
    sim.setupScene()
    assert sim.getParticlesCount() == 1620
    o = orc.OracleSim(0.4)
    assert o.setup_scene() == 1620
    o.set_order_mode(orc.ORDER_CANONICAL)
    p0, _, _ = sim.download()
    assert_bits_equal(p0, o.state()["pos"], "setupScene positions")
    for _ in range(30):
        sim.step()
    o.step(30)
    p, v, _ = sim.download()
    assert_bits_equal(p, o.state()["pos"], "positions after 30 steps")
    assert_bits_equal(v, o.state()["vel"], "velocities after 30 steps")


def test_reference_fountain_appends(hip, orc):
    """FOUNTAIN: 7 particles appended per step (generateParticles), 40 steps."""
    sim = hip.HIPGridParticleSimulator(0.4, scenario=hip.FOUNTAIN, flags=0)
    sim.setupScene()
    o = orc.OracleSim(0.4, scenario=orc.FOUNTAIN)
    o.setup_scene()
    o.set_order_mode(orc.ORDER_CANONICAL)
    assert sim.getMaxParticlesCount() == o.max_count
    for _ in range(40):
        sim.step()
    o.step(40)
    assert sim.getParticlesCount() == o.count == 280
    p, v, _ = sim.download()
    assert_bits_equal(p, o.state()["pos"], "fountain positions")
    assert_bits_equal(v, o.state()["vel"], "fountain velocities")


def test_neighbour_list_overflow_is_exact(hip, orc):
    """Particles whose neighbour list overflows (capacity 4 here, about 30 needed)
    re-walk their cells in the force pass: still bitwise the oracle."""
    pos = lattice_4k(hip)
    sim = gpu_sim(hip, 0.4, pos, diag=False, max_neighbours=4)
    o = oracle_sim(orc, 0.4, pos)
    sim.advance(8)
    o.step(8)
    p, v, _ = sim.download()
    assert_bits_equal(p, o.state()["pos"], "positions with overflowing lists")
    assert_bits_equal(v, o.state()["vel"], "velocities with overflowing lists")


def test_neighbour_list_mode_is_exact(hip, orc):
    """With neighbour lists kept by the density pass (capacity 256, no overflow)
    the force pass reads them instead of walking the cells: bitwise the oracle."""
    pos = lattice_4k(hip)
    sim = gpu_sim(hip, 0.4, pos, diag=False, max_neighbours=256)
    o = oracle_sim(orc, 0.4, pos)
    sim.advance(8)
    o.step(8)
    p, v, _ = sim.download()
    assert_bits_equal(p, o.state()["pos"], "positions with neighbour lists")
    assert_bits_equal(v, o.state()["vel"], "velocities with neighbour lists")


def test_determinism_repeat_runs(hip):
    """Two runs from the same state are bitwise equal (no atomics in sums)."""
    pos = lattice_4k(hip)
    out = []
    for _ in range(2):
        sim = gpu_sim(hip, 0.4, pos, diag=False)
        sim.advance(25)
        out.append(sim.download()[0])
    assert_bits_equal(out[0], out[1], "repeat run")


def test_edge_cases_bit_exact(hip, orc):
    """Isolated particle, a crowded cell, particles outside the box (clamped cells)
    and on cell boundaries: cells and 3 steps bit-exact."""
    rng = np.random.default_rng(5)
    box = np.float32(0.4)
    half = box / np.float32(2)
    crowd = (rng.random((300, 3), dtype=np.float32) * H * np.float32(0.999)).astype(np.float32) + np.float32(-0.1)
    outside = np.array([[half + 0.03, 0, 0], [-half - 0.05, 0.01, 0.02], [0, -half - 0.2, 0],
                        [0.1, 0.1, half + 0.01]], np.float32)
    edges = np.array([[-half + k * H, -half + H, -half + 2 * H] for k in range(1, 5)], np.float32)
    lone = np.array([[0.15, 0.15, 0.15]], np.float32)
    pos = np.concatenate([crowd, outside, edges, lone]).astype(np.float32)
    sim = gpu_sim(hip, 0.4, pos)
    o = oracle_sim(orc, 0.4, pos)
    sim.updateGrid()
    o.update_grid()
    np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_CELL_PRESORT), o.cells())
    sim.updateDensityPressure()
    sim.updateForces()
    sim.updateCollisions()
    sim.integrate()
    o.update_density_pressure()
    o.update_forces()
    o.integrate()
    for _ in range(2):
        sim.step()
        o.step(1)
    p, v, _ = sim.download()
    st = o.state()
    assert_bits_equal(p, st["pos"], "edge-case positions")
    assert_bits_equal(v, st["vel"], "edge-case velocities")


def test_non_finite_positions_hash_like_reference(hip, orc):
    """NaN, infinities and huge coordinates hash to the reference's cells (x86
    (int) cast semantics, then the clamp)."""
    pos = np.array([[np.nan, 0, 0], [np.inf, 0, 0], [-np.inf, 0.1, 0], [1e30, 0, 0], [-1e30, 0, 0.1],
                    [0.05, 3e9, -3e9], [0, 0, 0]], np.float32)
    sim = gpu_sim(hip, 0.4, pos)
    sim.updateGrid()
    np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_CELL_PRESORT), orc.hash_positions(pos, 0.4, 9))


def test_error_paths(hip):
    sim = hip.HIPGridParticleSimulator(0.4, capacity=16)
    sim.ensure_context(16)
    with pytest.raises(hip.SphError) as e:
        sim.updateGrid()
    assert e.value.status == hip._lib.SPH_ERR_STATE
    sim.upload(np.zeros((4, 3), np.float32))
    with pytest.raises(hip.SphError) as e:
        sim.updateForces()
    assert e.value.status == hip._lib.SPH_ERR_STATE
    with pytest.raises(hip.SphError) as e:
        sim.upload(np.zeros((17, 3), np.float32))
    assert e.value.status == hip._lib.SPH_ERR_CAPACITY
    with pytest.raises(hip.SphError) as e:
        sim.append(np.zeros((13, 3), np.float32))
    assert e.value.status == hip._lib.SPH_ERR_CAPACITY


def test_injected_fault_is_reported(hip, monkeypatch):
    """SPH_FAULT_AT_STEP=k makes step k fail on the host with SPH_ERR_FAULT before
    any launch (the failure path of the caller, without a device fault)."""
    monkeypatch.setenv("SPH_FAULT_AT_STEP", "3")
    sim = gpu_sim(hip, 0.4, lattice_4k(hip), diag=False)
    sim.advance(3)
    with pytest.raises(hip.SphError) as e:
        sim.advance(1)
    assert e.value.status == hip._lib.SPH_ERR_FAULT
    assert "step 3" in str(e.value)


def test_full_size_one_step(hip, orc):
    """The 1,048,576-particle lattice (box 3.2 m, 71^3 cells): one step through
    the phases, bit-exact against the canonical oracle; sort invariants."""
    pos = hip.scenes.lattice(128, 64, 128, 3.2, jitter_h=0.01, seed=1)
    n = len(pos)
    sim = gpu_sim(hip, 3.2, pos)
    (rx, ry, rz), ncells = sim.grid_dims()
    assert (rx, ry, rz) == (71, 71, 71)
    sim.updateGrid()
    pre = sim.debug(hip._lib.DEBUG_CELL_PRESORT)
    perm = sim.debug(hip._lib.DEBUG_PERM)
    keys = sim.debug(hip._lib.DEBUG_CELL)
    start = sim.debug(hip._lib.DEBUG_CELL_START)
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint32))
    same = keys[1:] == keys[:-1]
    assert np.all(perm[1:][same] > perm[:-1][same])  # stable within a cell
    np.testing.assert_array_equal(np.diff(start.astype(np.int64)), np.bincount(pre, minlength=ncells))
    np.testing.assert_array_equal(pre, orc.hash_positions(pos, 3.2, 71))
    o = oracle_sim(orc, 3.2, pos)
    o.update_grid()
    o.update_density_pressure()
    o.update_forces()
    o.integrate()
    sim.updateDensityPressure()
    sim.updateForces()
    sim.updateCollisions()
    ids = sim.debug(hip._lib.DEBUG_IDS)
    st = o.state()
    np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_NEIGHBOURS), o.neighbour_counts()[ids])
    assert_bits_equal(sim.debug(hip._lib.DEBUG_DENSITY), st["density"][ids], "density 1M")
    assert_bits_equal(sim.debug(hip._lib.DEBUG_ACCEL), st["acc"][ids], "acceleration 1M")
    sim.integrate()
    p, v, _ = sim.download()
    assert_bits_equal(p, st["pos"], "positions 1M")


@pytest.mark.parametrize("res,n", [(30, 200_000), (102, 200_000), (257, 200_000), (257, 1_200_000),
                                   (257, 8_000_000)])
def test_sort_across_key_widths(hip, orc, res, n):
    """The grid phase's stable sort of (cell, slot) against numpy's stable argsort
    of the oracle's cells, for grids whose keys take two 8-bit passes (30^3 cells,
    15-bit keys), three 8-bit passes (102^3, 21 bits) and three 10-bit passes
    (257^3, 25 bits); n particles uniform over the box, in random order. With
    1,200,000 particles the histograms of a 10-bit pass fill 74 scan tiles, more
    than the 64 predecessors one step of the scan's look-back reads; with
    8,000,000 they fill 489."""
    H = hip.scenes.H
    box = float(np.float32((res - 0.5) * float(H)))
    rng = np.random.default_rng(res + n)
    pos = ((rng.random((n, 3)) - 0.5) * box).astype(np.float32)
    print("grid %d^3, box %.4f m" % (res, box), flush=True)
    sim = gpu_sim(hip, box, pos, diag=False)
    (rx, ry, rz), ncells = sim.grid_dims()
    assert (rx, ry, rz) == (res, res, res)
    sim.updateGrid()
    pre = sim.debug(hip._lib.DEBUG_CELL_PRESORT)
    np.testing.assert_array_equal(pre, orc.hash_positions(pos, box, res))
    want = np.argsort(pre, kind="stable").astype(np.uint32)
    np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_PERM), want)
    np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_CELL), pre[want])
    start = sim.debug(hip._lib.DEBUG_CELL_START)
    np.testing.assert_array_equal(start, np.searchsorted(pre[want], np.arange(ncells + 1), side="left").astype(np.uint32))
    sim.close()


@pytest.mark.parametrize("limit", ["0", "1"])
def test_scan_fallback_is_exact(hip, orc, monkeypatch, limit):
    """The histogram scan's bounded wait: with SPH_SCAN_POLL_LIMIT=0 every scan
    tile adds up its predecessors' raw counts itself, with 1 a tile does so as
    soon as a predecessor has not published; the sort (three 10-bit passes, 13
    scan tiles each) is still a stable argsort, and 5 steps of a 110,592-particle
    dam break in the 3.2 m box (two 10-bit passes, 7 scan tiles each) are bitwise
    those of a context with the default limit."""
    monkeypatch.setenv("SPH_SCAN_POLL_LIMIT", limit)
    H = hip.scenes.H
    res, n = 257, 200_000
    box = float(np.float32((res - 0.5) * float(H)))
    rng = np.random.default_rng(5)
    pos = ((rng.random((n, 3)) - 0.5) * box).astype(np.float32)
    sim = gpu_sim(hip, box, pos, diag=False)
    sim.updateGrid()
    pre = sim.debug(hip._lib.DEBUG_CELL_PRESORT)
    np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_PERM), np.argsort(pre, kind="stable").astype(np.uint32))
    sim.close()
    lat = hip.scenes.lattice(48, 48, 48, 3.2, jitter_h=0.01, seed=3)   # 71^3 cells: 19-bit keys
    a = gpu_sim(hip, 3.2, lat, diag=False)
    monkeypatch.delenv("SPH_SCAN_POLL_LIMIT")
    b = gpu_sim(hip, 3.2, lat, diag=False)
    a.advance(5)
    b.advance(5)
    pa, va, _ = a.download()
    pb, vb, _ = b.download()
    assert_bits_equal(pa, pb, "positions, limit %s vs default" % limit)
    assert_bits_equal(va, vb, "velocities, limit %s vs default" % limit)


def test_full_workloads_stay_whole_finite_and_in_the_box(hip):
    """The benchmark workloads at full size and length: the 1M dam break for 1,000
    steps and the fountain to 2,000 steps (524,288 particles, graph replay). Every
    id is present once, every coordinate is finite, no particle is more than h
    outside the box (the wall spring acts within 0.01 m of a wall; measured: 0.31h
    and 0.37h), and the dam break's mass centre has fallen."""
    H = float(hip.scenes.H)
    pos, w = hip.scenes.workload("dam_break_1m")
    with hip.SceneBatch([(w["box"], len(pos))]) as b:
        b.upload(0, pos)
        b.step(1000)
        p, v, ids = b.download(0, by_id=False)
    assert np.array_equal(np.sort(ids), np.arange(len(pos), dtype=np.uint32))
    assert np.isfinite(p).all() and np.isfinite(v).all()
    assert np.abs(p).max() - np.float32(w["box"]) / 2 < H
    assert p[:, 1].mean() < pos[:, 1].mean()
    f = hip.scenes.FOUNTAIN
    with hip.SceneBatch([(f["box"], f["limit"])]) as b:
        b.set_emitter(0, kind=hip._lib.EMIT_DISC, rate=f["rate"], limit=f["limit"], centre=(0.0, 0.0),
                      radius=f["radius"], speed=f["speed"], spread=f["spread"], seed=f["seed"])
        b.set_graphs(True)
        b.step(2000)
        assert b.count(0) == f["limit"]
        p, v, ids = b.download(0, by_id=False)
    assert np.array_equal(np.sort(ids), np.arange(f["limit"], dtype=np.uint32))
    assert np.isfinite(p).all() and np.isfinite(v).all()
    assert np.abs(p).max() - np.float32(f["box"]) / 2 < H

