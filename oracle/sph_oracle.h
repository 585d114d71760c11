/*
 * sph_oracle.h -- CPU restatement of the reference's CCPUParticleSimulator.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP grid-SPH
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it. The product (libsph_hip.so) never links, calls or falls back to it.
 *
 * It restates, in plain C with the reference's mixed float/double arithmetic:
 *   - constants                 include/CParticle.h:80-84, include/CBaseParticleSimulator.h:19,
 *                               src/CBaseParticleSimulator.cpp:7, include/CCollisionGeometry.h:20-21
 *   - grid resolution           src/CBaseParticleSimulator.cpp:27-31, include/CGrid.h:25-28
 *   - dam-break scene           src/CBaseParticleSimulator.cpp:38-65, 67-74
 *   - fountain generation       src/CBaseParticleSimulator.cpp:187-210
 *   - updateGrid (rebucketing)  src/CCPUParticleSimulator.cpp:32-91
 *   - updateDensityPressure     src/CCPUParticleSimulator.cpp:93-141 (+ Wpoly6 :9-15)
 *   - updateForces              src/CCPUParticleSimulator.cpp:143-203 (+ kernels :17-30)
 *   - wall response             src/CCollisionGeometry.cpp:117-132, include/CCollisionGeometry.h:79-120
 *   - integrate                 src/CCPUParticleSimulator.cpp:211-229
 *   - step order                src/CBaseParticleSimulator.cpp:116-144
 * plus the canonical order the GPU path is specified against:
 *   - stable permutation by cell over the current array order, cell start table.
 *
 * Every entry point is pinned by tests/test_oracle_*.py against closed forms,
 * brute force or textbook routines (see the docstrings there).
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* Scenario, as SimulationScenario (include/CBaseParticleSimulator.h:21-25). */
enum { ORC_DAM_BREAK = 0, ORC_FOUNTAIN = 1 };

/* Summation order. The arithmetic of every term is the same in both modes;
 * only the order in which neighbour terms are added differs.
 * ORC_ORDER_REFERENCE: buckets by swap-remove + push_back exactly as
 *   src/CCPUParticleSimulator.cpp:72-84, neighbour cells visited x-outer,
 *   z-inner as src/CCPUParticleSimulator.cpp:108-121 (the reference's order).
 * ORC_ORDER_CANONICAL: the order the HIP path is specified in. The particles
 *   keep an array order (initially id order, new particles appended); every
 *   updateGrid replaces it by its stable sort by cell, and a cell's bucket is
 *   its run in that order. Neighbour cells are visited z-outer, y, x-inner, so
 *   that the three x-adjacent cells of a row form one contiguous run. */
enum { ORC_ORDER_REFERENCE = 0, ORC_ORDER_CANONICAL = 1 };

/* Constants as the reference computes them. out[0..6] =
 * {poly6 coef, spiky coef, viscosity coef, hSquared (double of float h*h),
 *  h, mass, rest density}. */
void orc_constants(double *out);

/* Grid resolution ceil(box/h) as int, src/CBaseParticleSimulator.cpp:27-31. */
int orc_grid_res(float box);

/* Reference dam-break particle count for a cubic box (setupScene loop). */
uint32_t orc_dam_break_count(float box);

orc_sim *orc_create(float box_x, float box_y, float box_z, uint32_t capacity, int scenario);
void orc_destroy(orc_sim *s);
uint32_t orc_count(const orc_sim *s);
uint32_t orc_max_count(const orc_sim *s);
void orc_grid_dims(const orc_sim *s, int *xyz);

/* CBaseParticleSimulator::addParticle: append with id = count, put into cell (0,0,0). */
int orc_add_particle(orc_sim *s, float x, float y, float z, float vx, float vy, float vz);
/* orc_add_particle for n particles (vel3 may be NULL: zero velocity). */
void orc_add_particles(orc_sim *s, const float *pos3, const float *vel3, uint32_t n);
/* CBaseParticleSimulator::setupScene. Returns the particle count (dam break) or 0. */
int orc_setup_scene(orc_sim *s);
void orc_set_gravity(orc_sim *s, float gx, float gy, float gz);
void orc_set_order_mode(orc_sim *s, int mode);
/* When set, density/forces also record per-particle neighbour counts and the
 * absolute force scale S_i used by the force tolerance. Off for timing. */
void orc_set_diagnostics(orc_sim *s, int on);

void orc_update_grid(orc_sim *s);
void orc_update_density_pressure(orc_sim *s);
void orc_update_forces(orc_sim *s);
void orc_integrate(orc_sim *s);
/* generateParticles + the five phases, nsteps times (CBaseParticleSimulator::step). */
void orc_step(orc_sim *s, uint32_t nsteps);

/* Per-particle state indexed by id (NULL pointers are skipped). */
void orc_get_state(const orc_sim *s, float *pos3, float *vel3, float *acc3,
                   float *density, float *pressure);
/* Overwrite position/velocity of particle id (state injection for tests). */
void orc_set_state(orc_sim *s, const float *pos3, const float *vel3);
/* Cell of each particle (by id) as last assigned by orc_update_grid. */
void orc_get_cells(const orc_sim *s, uint32_t *cell_id);
/* Bucket contents: cell_start[C+1] offsets into ids[count] in grid index order.
 * In canonical mode ids[] is the current array order. */
void orc_get_buckets(const orc_sim *s, uint32_t *cell_start, uint32_t *ids);
/* Diagnostics recorded by the last density / forces pass (by id). */
void orc_get_neighbour_counts(const orc_sim *s, uint32_t *counts);
void orc_get_force_scale(const orc_sim *s, float *scale3);

/* Stateless helpers. */
/* Cell hash of one position, src/CCPUParticleSimulator.cpp:46-69 + include/CGrid.h:25-28. */
uint32_t orc_hash(float px, float py, float pz, float box_x, float box_y, float box_z,
                  int rx, int ry, int rz);
void orc_hash_many(const float *pos3, uint32_t n, float box_x, float box_y, float box_z,
                   int rx, int ry, int rz, uint32_t *cell);
/* Stable permutation of [0,n) by key (counting sort), keys < ncells. */
void orc_stable_perm(const uint32_t *keys, uint32_t n, uint32_t ncells, uint32_t *perm);
/* start[c] = first sorted index with key >= c, for c in [0, ncells]. */
void orc_cell_start(const uint32_t *sorted_keys, uint32_t n, uint32_t ncells, uint32_t *start);
/* Wall acceleration for one particle, CCollisionGeometry::inverseBoundingBoxBounce. */
void orc_wall_accel(const float *pos3, const float *vel3, float box_x, float box_y, float box_z,
                    float *acc3);

#ifdef __cplusplus
}
#endif
#endif
