"""Host-side pieces of the batched path that need no GPU: the scene generators
of libsph_hip.so (lattice blocks, the batched workload) and the oracle's
restatement of the device disc emitter."""
import numpy as np

from conftest import lattice_numpy

H = np.float32(0.0457)


def test_block_at_corner_is_the_lattice(hip):
    box = np.array([0.5, 0.4, 0.6], np.float32)
    full = hip.scenes.lattice(7, 5, 9, box, jitter_h=0.01, seed=42)
    blk = hip.scenes.block(-box / np.float32(2.0), 7, 5, 9, jitter_h=0.01, seed=42)
    assert np.array_equal(full.view(np.uint32), blk.view(np.uint32))
    part = hip.scenes.block(-box / np.float32(2.0), 7, 5, 9, count=100, jitter_h=0.01, seed=42)
    assert np.array_equal(part.view(np.uint32), full[:100].view(np.uint32))
    assert np.array_equal(full.view(np.uint32), lattice_numpy(7, 5, 9, box, 0.01, 42).view(np.uint32))


def test_block_origin_shift(hip):
    a = hip.scenes.block([0.1, 0.2, -0.3], 3, 4, 5, jitter_h=0.0)
    half = np.float32(H / np.float32(2.0))
    assert a.shape == (60, 3)
    assert a[0].tolist() == [np.float32(0.1), np.float32(0.2), np.float32(-0.3)]
    assert np.isclose(a[1, 2] - a[0, 2], half)       # z runs fastest, then x, then y
    assert np.isclose(a[5, 0] - a[0, 0], half)
    assert np.isclose(a[20, 1] - a[0, 1], half)


def test_batch_workload_shape(hip):
    w = hip.scenes.batch_workload(nscenes=64, seed=3)
    ns = np.array([s["n"] for s in w])
    assert len(w) == 64 and ns.min() >= 2048 and ns.max() <= 65536
    kinds = {s["kind"] for s in w}
    assert kinds == {"dam_break", "cuboid_drop"}
    half = float(np.float32(H / np.float32(2.0)))
    for s in w:
        assert len(s["pos"]) == s["n"]
        assert np.all(np.abs(s["pos"]) <= s["box"] / 2 + 0.001)            # inside the walls
        fill = s["n"] * half ** 3 / float(np.prod(s["box"].astype(np.float64)))
        assert 0.2 < fill < 0.3                                             # about 1/4 of the box
        aspect = s["box"] / s["box"].min()
        assert aspect.max() <= 4.0 + 1e-6
    again = hip.scenes.batch_workload(nscenes=64, seed=3)
    assert all(np.array_equal(a["pos"], b["pos"]) for a, b in zip(w, again))


def test_splitmix64_reference_values(hip):
    # published SplitMix64 outputs for seed 1234567 (Vigna's reference implementation)
    g = hip.scenes.SplitMix64(1234567)
    assert [g.next() for _ in range(3)] == [6457827717110365317, 3203168211198807973, 9817491932198370423]


def test_disc_emission_statistics(orc):
    pos, vel = orc.disc_emission(5, 0, 4000, (0.1, -0.1), 0.3, 2.0, 0.1, np.float32(-0.8))
    r = np.hypot(pos[:, 0].astype(np.float64) - np.float32(0.1), pos[:, 2].astype(np.float64) + np.float32(0.1))
    assert r.max() <= 0.3 * (1 + 1e-6)
    assert np.all(pos[:, 1] == np.float32(-0.8))
    assert np.all(np.abs(vel[:, 1] - 2.0) <= 0.2 * (1 + 1e-6))
    assert np.all(vel[:, 0] == 0) and np.all(vel[:, 2] == 0)
    assert abs(np.mean(r ** 2) / 0.3 ** 2 - 0.5) < 0.03        # uniform on the disc: E[r^2] = R^2/2
    assert abs(np.mean(vel[:, 1]) - 2.0) < 0.01


def test_disc_emission_is_keyed_by_particle(orc):
    """Particle g's draws depend only on (seed, g): emitting 10 then 15 equals 25."""
    a, va = orc.disc_emission(9, 0, 25, (0, 0), 0.2, 1.0, 0.1, np.float32(-0.5))
    b1, vb1 = orc.disc_emission(9, 0, 10, (0, 0), 0.2, 1.0, 0.1, np.float32(-0.5))
    b2, vb2 = orc.disc_emission(9, 10, 15, (0, 0), 0.2, 1.0, 0.1, np.float32(-0.5))
    assert np.array_equal(a, np.vstack([b1, b2])) and np.array_equal(va, np.vstack([vb1, vb2]))


def test_write_logs_matches_export_format(hip, tmp_path):
    """sph_write_logs appends the two files of MainWindow::exportLogs
    (src/mainwindow.cpp:310-368): summary 'label;together;...' and the detail
    block (label, one line per phase, two empty lines), numbers as %g."""
    ev = [dict(iteration=0, update_grid=0.125, update_density_pressure=1.5, update_forces=2.25,
               update_collisions=0.0, integrate=0.0),
          dict(iteration=10, update_grid=0.1, update_density_pressure=1234567.0, update_forces=3e-7,
               update_collisions=0.5, integrate=0.25)]
    stem = tmp_path / "Dam break_0.4"
    hip._lib.write_logs(ev, stem, "HIP Grid")
    hip._lib.write_logs(ev[:1], stem, "HIP Grid")          # appends, as the reference does
    summary = (tmp_path / "Dam break_0.4.csv").read_text()
    detail = (tmp_path / "Dam break_0.4_detail.csv").read_text()
    assert summary == "HIP Grid;3.875;1.23457e+06;\nHIP Grid;3.875;\n"
    assert detail == ("HIP Grid\nGrid;0.125;0.1;\nDensity + pressure;1.5;1.23457e+06;\nForces;2.25;3e-07;\n"
                      "Collisions;0;0.5;\nIntegrate;0;0.25;\n\n\n"
                      "HIP Grid\nGrid;0.125;\nDensity + pressure;1.5;\nForces;2.25;\nCollisions;0;\nIntegrate;0;\n\n\n")
