"""CPU oracle for the grid-SPH step (test infrastructure only).

A plain-C restatement of the reference's ``CCPUParticleSimulator``
(src/CCPUParticleSimulator.cpp) loaded through ctypes. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
package, and only as the checker: the HIP product path never calls it.
"""
from .oracle import (  # noqa: F401
    OracleSim, constants, grid_res, dam_break_count, hash_positions,
    stable_perm, cell_start, wall_accel, lib_path, build, disc_emission,
    DAM_BREAK, FOUNTAIN, ORDER_REFERENCE, ORDER_CANONICAL,
)
