"""The reference's own summation order on the GPU (SPH_FLAG_REFERENCE_ORDER)
against the oracle in ORDER_REFERENCE: the restated CCPUParticleSimulator with
its swap-remove buckets (src/CCPUParticleSimulator.cpp:37-88), its x-outer,
z-inner neighbour loops (:108-127) and new particles pushed to the back of cell
(0,0,0)'s bucket (src/CBaseParticleSimulator.cpp:67-74).

Bar: bit-exact. The array order after every grid phase is the oracle's buckets
concatenated in the reference's traversal order, and the cells, neighbour counts,
density, acceleration, positions and velocities are bitwise the oracle's. Over
100 steps the positions therefore stay within 0 <= 1e-3*h of the reference
order, the bound that the canonical order cannot meet (its summation-order chaos
reaches about h by step 100).
"""
import numpy as np
import pytest

from test_gpu_parity import assert_bits_equal

pytestmark = pytest.mark.gpu

H = np.float32(0.0457)


def traversal_to_linear(t, dims):
    """Cell number z + y*Z + x*Z*Y (the reference's visiting order) -> x + y*X + z*X*Y (CGrid)."""
    X, Y, Z = dims
    t = np.asarray(t, np.int64)
    z, y, x = t % Z, (t // Z) % Y, t // (Z * Y)
    return (x + y * X + z * X * Y).astype(np.uint32)


def oracle_array_order(o):
    """The oracle's buckets concatenated in its traversal order (x outer, z inner), and
    the traversal number of each bucket's first slot (a cell table in that order)."""
    X, Y, Z = o.grid_dims
    start, ids = o.buckets()
    order, table = [], [0]
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                c = x + y * X + z * X * Y
                order.append(ids[start[c]:start[c + 1]])
                table.append(table[-1] + int(start[c + 1] - start[c]))
    return np.concatenate(order).astype(np.uint32), np.array(table, np.uint32)


def ref_sim(hip, box, pos=None, vel=None, diag=True, capacity=None, scenario=0, device_emitter=False, **kw):
    flags = hip._lib.FLAG_REFERENCE_ORDER | (hip._lib.FLAG_DIAGNOSTICS if diag else 0)
    sim = hip.HIPGridParticleSimulator(box, flags=flags, capacity=capacity or (len(pos) if pos is not None else None),
                                       scenario=scenario, device_emitter=device_emitter, **kw)
    if pos is not None:
        sim.upload(pos, vel)
    return sim


def ref_oracle(orc, box, pos=None, vel=None, scenario=0):
    o = orc.OracleSim(box, len(pos) if pos is not None else 0, scenario=scenario)
    if pos is not None:
        o.add_particles(pos, vel)
    o.set_order_mode(orc.ORDER_REFERENCE)
    o.set_diagnostics(True)
    return o


def check_grid(hip, sim, o, what):
    """After both grid phases: array order, cells and the cell table."""
    dims = o.grid_dims
    ids = sim.debug(hip._lib.DEBUG_IDS)
    want, table = oracle_array_order(o)
    np.testing.assert_array_equal(ids, want, err_msg="%s: bucket order" % what)
    cells = traversal_to_linear(sim.debug(hip._lib.DEBUG_CELL), dims)
    np.testing.assert_array_equal(cells, o.cells()[ids], err_msg="%s: cells" % what)
    np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_CELL_START), table, err_msg="%s: cell table" % what)
    return ids


def test_phases_bitwise_reference_order_4k(hip, orc, scene_4k):
    """Three steps phase by phase on the 4,096-particle lattice (all particles start
    in cell (0,0,0) in id order, as setupScene leaves them)."""
    sim = ref_sim(hip, 0.4, scene_4k)
    o = ref_oracle(orc, 0.4, scene_4k)
    for step in range(3):
        sim.updateGrid()
        o.update_grid()
        ids = check_grid(hip, sim, o, "step %d" % step)
        sim.updateDensityPressure()
        o.update_density_pressure()
        st = o.state()
        np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_NEIGHBOURS), o.neighbour_counts()[ids])
        assert_bits_equal(sim.debug(hip._lib.DEBUG_DENSITY), st["density"][ids], "density, step %d" % step)
        sim.updateForces()
        sim.updateCollisions()
        o.update_forces()
        assert_bits_equal(sim.debug(hip._lib.DEBUG_ACCEL), o.state()["acc"][ids], "acceleration, step %d" % step)
        sim.integrate()
        o.integrate()
        p, v, _ = sim.download()
        st = o.state()
        assert_bits_equal(p, st["pos"], "positions, step %d" % step)
        assert_bits_equal(v, st["vel"], "velocities, step %d" % step)


def test_100_steps_within_1e3_h_of_the_reference_order(hip, orc, scene_4k):
    """100 fused steps of the 4,096-particle dam break in the reference order:
    bitwise the reference-order oracle at steps 10, 50 and 100 (positions,
    velocities and bucket order), so the distance at step 100 is 0 <= 1e-3*h."""
    sim = ref_sim(hip, 0.4, scene_4k, diag=False)
    o = ref_oracle(orc, 0.4, scene_4k)
    done = 0
    for upto in (10, 50, 100):
        sim.advance(upto - done)
        o.step(upto - done)
        done = upto
        p, v, _ = sim.download()
        st = o.state()
        assert_bits_equal(p, st["pos"], "positions at step %d" % upto)
        assert_bits_equal(v, st["vel"], "velocities at step %d" % upto)
        np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_IDS), oracle_array_order(o)[0])
    d100 = float(np.abs(p.astype(np.float64) - st["pos"]).max())
    print("step 100: |gpu(reference order) - oracle(reference order)| = %.3e m (bound 1e-3 h = %.3e m)"
          % (d100, 1e-3 * float(H)))
    assert d100 <= 1e-3 * float(H)


def test_reference_scene_100_steps(hip, orc):
    """The reference's setupScene (box 0.4 m, 1,620 particles in cell (0,0,0) in id
    order), 100 steps in the reference order, with the steps replayed from graphs."""
    sim = hip.HIPGridParticleSimulator(0.4, flags=hip._lib.FLAG_REFERENCE_ORDER)
    sim.setupScene()
    assert sim.getParticlesCount() == 1620
    hip._lib.check(sim._lib.sph_set_graphs(sim._ctx, 1))
    o = orc.OracleSim(0.4)
    assert o.setup_scene() == 1620
    o.set_order_mode(orc.ORDER_REFERENCE)
    done = 0
    for upto in (1, 30, 100):
        k, done = upto - done, upto
        sim.advance(k)
        o.step(k)
        p, v, _ = sim.download()
        assert_bits_equal(p, o.state()["pos"], "positions after %d steps" % upto)
        assert_bits_equal(v, o.state()["vel"], "velocities after %d steps" % upto)
        np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_IDS), oracle_array_order(o)[0])


@pytest.mark.parametrize("device_emitter", [False, True])
def test_reference_fountain_40_steps(hip, orc, device_emitter):
    """FOUNTAIN in the reference order: 7 particles per step pushed to the back of
    cell (0,0,0)'s bucket before the grid phase, appended from the host
    (sph_append) or by the device emitter; 40 steps bitwise, buckets included."""
    sim = hip.HIPGridParticleSimulator(0.4, scenario=hip.FOUNTAIN, flags=hip._lib.FLAG_REFERENCE_ORDER,
                                       device_emitter=device_emitter)
    sim.setupScene()
    o = orc.OracleSim(0.4, scenario=orc.FOUNTAIN)
    o.setup_scene()
    o.set_order_mode(orc.ORDER_REFERENCE)
    done = 0
    for upto in (1, 12, 40):
        k, done = upto - done, upto
        for _ in range(k):
            sim.step()
        o.step(k)
        assert sim.getParticlesCount() == o.count
        p, v, _ = sim.download()
        assert_bits_equal(p, o.state()["pos"], "fountain positions after %d steps" % upto)
        assert_bits_equal(v, o.state()["vel"], "fountain velocities after %d steps" % upto)
        np.testing.assert_array_equal(sim.debug(hip._lib.DEBUG_IDS), oracle_array_order(o)[0])
    assert o.count == 280


def test_modes_and_variants_identical_in_reference_order(hip, orc, scene_4k):
    """Every neighbour mode and a spread of kernel shapes read the reference order's
    rows (x outer, runs along z) identically: bitwise the oracle after 12 steps."""
    o = ref_oracle(orc, 0.4, scene_4k)
    o.step(12)
    st = o.state()
    for variant, mode in ((12, hip._lib.NEIGHBOURS_MASKS), (0, hip._lib.NEIGHBOURS_WALK), (18, hip._lib.NEIGHBOURS_MASKS),
                          (5, hip._lib.NEIGHBOURS_MASKS), (3, hip._lib.NEIGHBOURS_LISTS)):
        sim = ref_sim(hip, 0.4, scene_4k, diag=False, max_neighbours=64 if mode == hip._lib.NEIGHBOURS_LISTS else 0)
        hip._lib.check(sim._lib.sph_set_variant(sim._ctx, variant, mode))
        sim.advance(12)
        p, v, _ = sim.download()
        assert_bits_equal(p, st["pos"], "positions, variant %d mode %d" % (variant, mode))
        assert_bits_equal(v, st["vel"], "velocities, variant %d mode %d" % (variant, mode))
