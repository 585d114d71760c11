"""bench.py's host-side contract (no GPU): one GPU only, and the byte model."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def run_bench(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                          timeout=120, env=e, cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = run_bench(["--gpus", "2"])
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""


def test_world_size_above_one_exits_with_a_message():
    r = run_bench([], env={"WORLD_SIZE": "4"})
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_bytes_are_the_step_model():
    import bench
    # 148 B per particle-step plus 24 B per cell (hash 16, sort 8, reorder 56, density 16, forces 52)
    assert sum(bench.ALG_BYTES.values()) == 148
    assert bench.ALG_BYTES == {"hash": 16, "sort": 8, "reorder": 56, "density": 16, "forces": 52}
    assert bench.ALG_BYTES_PER_CELL == 24
    assert bench.HBM_PEAK_GBS == 8000.0


def test_fountain_particle_steps_closed_form():
    # 1,024 particles per step until 524,288 (step 512), 2,000 steps: the count bench_fountain divides by time
    steps, rate, limit = 2000, 1024, 524288
    live = np.minimum(np.arange(1, steps + 1, dtype=np.int64) * rate, limit)
    assert live.sum() == rate * 512 * 513 // 2 + limit * (steps - 512)
