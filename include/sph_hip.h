/*
 * sph_hip.h -- C ABI of libsph_hip.so, the MI355X grid-SPH step.
 *
 * The library replaces the OpenCL grid path of the reference (CGPUParticleSimulator
 * + CLWrapper + resources/kernels/sph_grid.cl, sph_common.cl) with hand-written HIP
 * kernels for gfx950. A host class deriving from CBaseParticleSimulator
 * (include/CBaseParticleSimulator.h:29-110) calls these entry points; its five
 * phase virtuals map 1:1 onto sph_phase().
 *
 * Conventions
 *  - Every entry point returns sph_status (0 = SPH_OK) and never throws. The text of
 *    the last failure on the calling thread is sph_last_error(), as CLException::what()
 *    is for the reference (include/CLWrapper.h:18-22).
 *  - Host buffers are caller-owned; the library owns all device memory.
 *  - Positions/velocities cross the ABI as packed xyz float triples (AoS, 12 B), the
 *    order of CParticle::Physics' position/velocity (include/CParticle.h:25-26).
 *  - "Array order" is the device's current particle order: stable-sorted by cell as of
 *    the last grid phase (with SPH_FLAG_REFERENCE_ORDER: the reference's buckets, cell by
 *    cell). `ids` give each slot's particle id.
 *  - Physics constants are the reference's compile-time values (include/CParticle.h:80-84,
 *    include/CCollisionGeometry.h:20-21); the arithmetic follows CCPUParticleSimulator
 *    (src/CCPUParticleSimulator.cpp) term by term in a fixed canonical summation order.
 */
#ifndef SPH_HIP_H
#define SPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int sph_status;
enum {
    SPH_OK = 0,
    SPH_ERR_INVALID = 1,    /* bad argument */
    SPH_ERR_CAPACITY = 2,   /* particle count beyond the context's capacity */
    SPH_ERR_HIP = 3,        /* a HIP runtime call failed */
    SPH_ERR_OOM = 4,        /* device (or host) allocation failed */
    SPH_ERR_NO_DEVICE = 5,  /* no usable gfx950 device */
    SPH_ERR_STATE = 6,      /* phase called out of order / nothing uploaded */
    SPH_ERR_FAULT = 7       /* injected failure (SPH_FAULT_AT_STEP env hook) */
};

/* Phases, in CBaseParticleSimulator::step order (src/CBaseParticleSimulator.cpp:116-144). */
enum {
    SPH_PHASE_GRID = 0,       /* updateGrid: hash, stable sort by cell, cell table, reorder */
    SPH_PHASE_DENSITY = 1,    /* updateDensityPressure */
    SPH_PHASE_FORCES = 2,     /* updateForces: pressure + viscosity + gravity */
    SPH_PHASE_COLLISIONS = 3, /* updateCollisions: wall response added to the acceleration */
    SPH_PHASE_INTEGRATE = 4   /* integrate */
};

/* Context flags. */
enum {
    SPH_FLAG_DIAGNOSTICS = 1u << 0, /* keep neighbour counts and accelerations for sph_debug */
    SPH_FLAG_PROFILE = 1u << 1,     /* time every stage with HIP events (sph_stage_times) */
    SPH_FLAG_REFERENCE_ORDER = 1u << 2 /* sum in the reference's own order: CCPUParticleSimulator's swap-remove
                                          buckets (src/CCPUParticleSimulator.cpp:37-88), cells visited x outer,
                                          z inner (:108-121); uploaded and emitted particles join the back of
                                          cell (0,0,0)'s bucket as addParticle does (src/CBaseParticleSimulator.cpp:
                                          67-74). Cells are then numbered z + y*Z + x*Y*Z (the reference's
                                          traversal order) in sph_debug. Slower grid phase; same arithmetic;
                                          the neighbour kernels run in one shape (variant 12's) whatever
                                          sph_set_variant chooses. */
};

/* sph_debug kinds. Per-particle arrays are in the current array order unless noted. */
enum {
    SPH_DEBUG_CELL_PRESORT = 0, /* uint32[n]: cell of each particle in the order before the last sort */
    SPH_DEBUG_PERM = 1,         /* uint32[n]: pre-sort array index of the particle at sorted slot k */
    SPH_DEBUG_CELL = 2,         /* uint32[n]: cell of each particle (sorted order) */
    SPH_DEBUG_CELL_START = 3,   /* uint32[C+1]: first sorted slot of each cell */
    SPH_DEBUG_NEIGHBOURS = 4,   /* uint32[n]: accepted neighbours incl. self (needs DIAGNOSTICS) */
    SPH_DEBUG_DENSITY = 5,      /* float[n] */
    SPH_DEBUG_PRESSURE = 6,     /* float[n]: gas_stiffness * (density - rest_density) */
    SPH_DEBUG_ACCEL = 7,        /* float[3n] xyz (needs DIAGNOSTICS in the fused step) */
    SPH_DEBUG_IDS = 8           /* uint32[n] */
};

/* Per-stage timer slots of sph_stage_times. */
enum {
    SPH_STAGE_HASH = 0, SPH_STAGE_SORT = 1, SPH_STAGE_CELLS = 2, SPH_STAGE_REORDER = 3,
    SPH_STAGE_DENSITY = 4, SPH_STAGE_FORCES = 5, SPH_STAGE_COLLISIONS = 6, SPH_STAGE_INTEGRATE = 7,
    SPH_STAGE_COUNT = 8
};

typedef struct sph_ctx sph_ctx;

typedef struct {
    float box[3];       /* m_boxSize (include/CBaseParticleSimulator.h:87); walls at +-box/2 */
    float gravity[3];   /* initial gravity, (0, -9.80665, 0) in the reference */
    float dt;           /* 0.01f in the reference (src/CBaseParticleSimulator.cpp:7) */
    uint32_t capacity;  /* maximum particle count (m_maxParticlesCount) */
    uint32_t flags;     /* SPH_FLAG_* */
    uint32_t max_neighbours; /* 0 (default) or SPH_NEIGHBOUR_LISTS_OFF: no lists (the force pass reads
                                the density pass's row masks, see sph_set_variant). N > 0: the density
                                pass keeps lists of up to N neighbours per particle for the force pass
                                (N*4 bytes per particle of HBM); particles with more neighbours are
                                still exact, only slower. */
} sph_params;

#define SPH_NEIGHBOUR_LISTS_OFF 0xFFFFFFFFu

/* Fill p with the reference defaults for a cubic box of edge `box` and the given
 * capacity (CBaseParticleSimulator ctor, src/CBaseParticleSimulator.cpp:3-36). */
void sph_default_params(sph_params *p, float box, uint32_t capacity);

/* Grid resolution (int) ceil(box / h), src/CBaseParticleSimulator.cpp:27-31. */
int sph_grid_res(float box);

/* Replaces: CGPUParticleSimulator ctor + setupKernels (src/CGPUParticleSimulator.cpp:3-55,
 * src/CGPUBaseParticleSimulator.cpp:4-9, 39-57). Allocates all device buffers. */
sph_status sph_create(const sph_params *params, sph_ctx **out);

/* Replaces: ~CGPUParticleSimulator (include/CGPUParticleSimulator.h:17). NULL is a no-op. */
void sph_destroy(sph_ctx *ctx);

/* Replaces: the blocking H2D of m_clParticles (src/CGPUParticleSimulator.cpp:61).
 * Sets the state to n particles (array order = the given order). ids may be NULL
 * (then id = index). */
sph_status sph_upload(sph_ctx *ctx, const float *pos3, const float *vel3, const uint32_t *ids, uint32_t n);

/* Replaces: addParticle via generateParticles (src/CBaseParticleSimulator.cpp:67-74,
 * 187-210), which the reference ships by re-uploading everything
 * (src/CGPUParticleSimulator.cpp:61). Appends n particles at the end of the array
 * order. ids may be NULL (then id = current count + index). */
sph_status sph_append(sph_ctx *ctx, const float *pos3, const float *vel3, const uint32_t *ids, uint32_t n);

/* Replaces: CGPUBaseParticleSimulator::setGravityVector (src/CGPUBaseParticleSimulator.cpp:11-15). */
sph_status sph_set_gravity(sph_ctx *ctx, float gx, float gy, float gz);

/* Replaces: updateGrid / updateDensityPressure / updateForces / updateCollisions /
 * integrate (include/CBaseParticleSimulator.h:92-96; src/CGPUParticleSimulator.cpp:57-172,
 * src/CGPUBaseParticleSimulator.cpp:59-97). Runs one phase on the device. If ms_out is
 * not NULL the call waits and stores the phase's device time in ms (the reference's
 * phase methods return ms). */
sph_status sph_phase(sph_ctx *ctx, int phase, float *ms_out);

/* Replaces: CBaseParticleSimulator::step (src/CBaseParticleSimulator.cpp:116-144), nsteps
 * times, with forces + walls + integrate fused. Bit-identical to calling the five phases.
 * If ms_out is not NULL the call waits and stores the device time of all nsteps in ms. */
sph_status sph_step(sph_ctx *ctx, uint32_t nsteps, float *ms_out);

/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the steps of sph_step on particle arrays the caller owns in device memory, calling no
 * hipMalloc/hipFree or other allocating routine. Every other buffer (the second particle
 * buffer, sort keys and permutation, radix histograms, the cell table, density records)
 * lives in one static device array of SPH_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends. `state` holds n particles in array order (structure of arrays, device
 * pointers); params gives box, gravity and dt (capacity, flags and max_neighbours are
 * not used). The call enqueues nsteps steps on `stream` (a hipStream_t; NULL = the null
 * stream) and returns; once they have run, the arrays hold the state sph_step leaves
 * after sph_upload of the same arrays, bit for bit, in its new array order (the ids
 * permuted with the particles). n above SPH_NOALLOC_MAX_PARTICLES or a grid of more than
 * SPH_NOALLOC_MAX_CELLS cells is SPH_ERR_CAPACITY. All calls on one GPU share its array:
 * calls that overlap on one GPU (other threads or streams) race and must not be made. */
#define SPH_NOALLOC_MAX_PARTICLES 65536u
#define SPH_NOALLOC_MAX_CELLS 262144u
#define SPH_NOALLOC_SCRATCH_BYTES 7357696u
typedef struct {
    float *px, *py, *pz;
    float *vx, *vy, *vz;
    uint32_t *id;
} sph_device_state;
sph_status sph_step_noalloc(const sph_params *params, const sph_device_state *state, uint32_t n, uint32_t nsteps,
                            void *stream);

/* Replaces: the readback in CGPUBaseParticleSimulator::integrate
 * (src/CGPUBaseParticleSimulator.cpp:83-91). Copies n (= count) particles; by_id != 0
 * writes particle id k to row k, else rows follow the array order. Any pointer may be NULL. */
sph_status sph_download(sph_ctx *ctx, float *pos3, float *vel3, uint32_t *ids, uint32_t n, int by_id);

/* Replaces: reading fields of CParticle::Physics back from the device
 * (src/CGPUParticleSimulator.cpp:118, include/CParticle.h:22-43): cell ids, sort
 * permutation, cell table, neighbour counts, density, pressure, acceleration. `bytes`
 * is the capacity of out; too small is SPH_ERR_INVALID. */
sph_status sph_debug(sph_ctx *ctx, int kind, void *out, size_t bytes);

/* Replaces: CBaseParticleSimulator::getParticlesCount (include/CBaseParticleSimulator.h:52). */
sph_status sph_count(const sph_ctx *ctx, uint32_t *n);

/* Grid dimensions and cell count of the context (CGrid::xRes/yRes/zRes, include/CGrid.h:30-34). */
sph_status sph_grid_dims(const sph_ctx *ctx, int *xyz, uint32_t *ncells);

/* Replaces: getSelectedDevice (include/CBaseParticleSimulator.h:42,
 * src/CGPUBaseParticleSimulator.cpp:17-20). Writes the device name. */
sph_status sph_device_name(char *buf, size_t len);

/* Waits for all work of the context. */
sph_status sph_sync(sph_ctx *ctx);

/* Replaces: CLWrapper::getEventDuration with PROFILING (src/CLWrapper.cpp:128-162).
 * Accumulated device ms per stage (SPH_STAGE_COUNT floats) and launches counted, since
 * creation or the last reset; needs SPH_FLAG_PROFILE. */
sph_status sph_stage_times(sph_ctx *ctx, float *ms, uint32_t *launches, int reset);

/* Replaces: CLException::what() (include/CLWrapper.h:18-22). Thread-local; never NULL. */
const char *sph_last_error(void);

/* Replaces: CBaseParticleSimulator::setupScene DAM_BREAK (src/CBaseParticleSimulator.cpp:38-65).
 * Writes the reference's dam-break lattice for a cubic box into pos3 (id order) and
 * returns the count in *n. With pos3 == NULL only *n is computed. */
sph_status sph_scene_dam_break(float box, float *pos3, uint32_t cap, uint32_t *n);

/* Benchmark lattice (no reference counterpart: the benchmark scene definition).
 * nx*ny*nz sites at spacing h/2 from the box corner -box/2, nested y, x, z as the
 * reference's setupScene loop, coordinates accumulated in float, each coordinate then
 * jittered by (2u-1)*jitter_h*h with u from SplitMix64(seed) (24-bit uniforms drawn in
 * id order, x then y then z). */
sph_status sph_scene_lattice(uint32_t nx, uint32_t ny, uint32_t nz, const float box[3],
                             float jitter_h, uint64_t seed, float *pos3);

/* Block of a benchmark lattice (no reference counterpart): the first `count` sites
 * (count <= nx*ny*nz) of the lattice of sph_scene_lattice placed with its first site at
 * `origin` instead of -box/2, with the same jitter draws. sph_scene_lattice(nx, ny, nz,
 * box, ...) is the block at origin -box/2 with count nx*ny*nz. */
sph_status sph_scene_block(const float origin[3], uint32_t nx, uint32_t ny, uint32_t nz, uint32_t count,
                           float jitter_h, uint64_t seed, float *pos3);

/* ------------------------------------------------------------------------------
 * Batched scenes. No counterpart in the reference, which advances one scene per
 * CBaseParticleSimulator (src/CBaseParticleSimulator.cpp:116-144): a batch context
 * advances many independent scenes with one launch per stage. Scenes share a pool of
 * particle slots (each scene a segment starting on a multiple of 256) and a global
 * cell-id space (each scene its own grid plus one cell for unused slots). A context
 * made by sph_create is a batch of one scene, id 0; the single-scene entry points
 * above act on scene 0 (sph_step and sph_phase always advance every scene).
 * Every scene computes exactly what it computes alone in a context of its own.
 * ------------------------------------------------------------------------------ */

typedef struct {
    float box[3];       /* walls at +-box/2 */
    float gravity[3];
    float dt;
    uint32_t capacity;  /* particle slots reserved for the scene */
} sph_scene_desc;

typedef struct {
    uint32_t pool;           /* slots shared by all scenes; 0 = the scenes' capacities, each rounded up to 256 */
    uint32_t max_scenes;     /* scene ids 0..max_scenes-1; 0 = nscenes */
    uint32_t flags;          /* SPH_FLAG_* */
    uint32_t max_neighbours; /* as in sph_params */
} sph_batch_params;

/* Reference defaults for a cubic box of edge `box` (as sph_default_params). */
void sph_default_scene(sph_scene_desc *d, float box, uint32_t capacity);

/* Creates a context with the given scenes (ids 0..nscenes-1, placed in order).
 * nscenes may be 0 (scenes are then added with sph_scene_add). */
sph_status sph_batch_create(const sph_batch_params *bp, const sph_scene_desc *scenes, uint32_t nscenes,
                            sph_ctx **out);

/* Adds a scene in the first free range of the pool that holds it (SPH_ERR_CAPACITY if
 * none does) and returns its id. Only between steps. */
sph_status sph_scene_add(sph_ctx *ctx, const sph_scene_desc *desc, uint32_t *scene);

/* Frees a scene's slots and id. Only between steps. */
sph_status sph_scene_remove(sph_ctx *ctx, uint32_t scene);

/* sph_upload / sph_append / sph_download / sph_set_gravity / sph_count / sph_debug for
 * one scene. sph_scene_debug returns the scene's own slots and cells: slot indices
 * (SPH_DEBUG_PERM, SPH_DEBUG_CELL_START) are relative to the scene's first slot and
 * cell ids (SPH_DEBUG_CELL_PRESORT, SPH_DEBUG_CELL) to its first cell. */
sph_status sph_scene_upload(sph_ctx *ctx, uint32_t scene, const float *pos3, const float *vel3,
                            const uint32_t *ids, uint32_t n);
sph_status sph_scene_append(sph_ctx *ctx, uint32_t scene, const float *pos3, const float *vel3,
                            const uint32_t *ids, uint32_t n);
sph_status sph_scene_download(sph_ctx *ctx, uint32_t scene, float *pos3, float *vel3, uint32_t *ids,
                              uint32_t n, int by_id);
sph_status sph_scene_set_gravity(sph_ctx *ctx, uint32_t scene, float gx, float gy, float gz);
sph_status sph_scene_count(sph_ctx *ctx, uint32_t scene, uint32_t *n);
sph_status sph_scene_debug(sph_ctx *ctx, uint32_t scene, int kind, void *out, size_t bytes);

/* Grid of one scene (as sph_grid_dims) and its segment of slots. */
sph_status sph_scene_grid(sph_ctx *ctx, uint32_t scene, int *xyz, uint32_t *ncells, uint32_t *first_slot,
                          uint32_t *slots);

/* Pool summary: scene ids in use, slots in the pool, free slots, and the slots the
 * next step covers (sorted and launched over). Any pointer may be NULL. */
sph_status sph_batch_info(sph_ctx *ctx, uint32_t *scenes, uint32_t *pool, uint32_t *free_slots,
                          uint32_t *active_slots);

/* Device particle sources. */
enum {
    SPH_EMIT_FOUNTAIN = 0, /* the reference fountain: 7 particles per step at the floor centre
                              while count < limit - 7 (src/CBaseParticleSimulator.cpp:187-210) */
    SPH_EMIT_DISC = 1      /* `rate` particles per step uniformly on a floor disc until count = limit */
};

typedef struct {
    uint32_t kind;      /* SPH_EMIT_* */
    uint32_t rate;      /* DISC: particles per step */
    uint32_t limit;     /* FOUNTAIN: m_maxParticlesCount; DISC: final count (at most the capacity) */
    float centre[2];    /* DISC: x and z of the disc centre (y is the floor, -box_y/2) */
    float radius;       /* DISC */
    float speed;        /* DISC: mean upward speed */
    float spread;       /* DISC: v_y = speed * (1 + spread * (2u - 1)) */
    uint64_t seed;      /* DISC: SplitMix64 seed; particle g uses the stream at seed + 32*g*0x9E3779B97F4A7C15 */
} sph_emitter;

/* Replaces: generateParticles + addParticle (src/CBaseParticleSimulator.cpp:67-74, 187-210)
 * and the full re-upload that follows (src/CGPUParticleSimulator.cpp:61). Attaches an
 * emitter that runs on the device at the start of every step (before the grid phase),
 * without host round trips; e == NULL detaches. Particle ids continue from the count. */
sph_status sph_scene_set_emitter(sph_ctx *ctx, uint32_t scene, const sph_emitter *e);

/* Captures each step of sph_step in a HIP graph and replays it (default off; never with
 * SPH_FLAG_PROFILE). Results are identical with and without graphs. */
sph_status sph_set_graphs(sph_ctx *ctx, int enable);

/* Shape of the neighbour kernels: variant 0..SPH_NUM_VARIANTS-1 (workgroup size and
 * candidates per trip, see sph_variant_shape), and how the
 * force pass finds the neighbours:
 *   SPH_NEIGHBOURS_WALK  tests every candidate of the 27 cells again;
 *   SPH_NEIGHBOURS_LISTS reads neighbour lists the density pass kept (only if the context
 *                        was created with max_neighbours > 0);
 *   SPH_NEIGHBOURS_MASKS reads per-row masks of the accepted candidates the density pass
 *                        kept (576 bytes per slot).
 * A new context uses variant 12 and SPH_NEIGHBOURS_MASKS, or SPH_NEIGHBOURS_LISTS when it
 * was created with max_neighbours > 0. Results are identical for every choice. */
#define SPH_NUM_VARIANTS 21
enum { SPH_NEIGHBOURS_WALK = 0, SPH_NEIGHBOURS_LISTS = 1, SPH_NEIGHBOURS_MASKS = 2 };
sph_status sph_set_variant(sph_ctx *ctx, int variant, int mode);
sph_status sph_get_variant(const sph_ctx *ctx, int *variant, int *mode);
/* Threads per workgroup, candidates tested per lane per trip, and how the mask mode's
 * force pass reads the masks, of a variant: with `pairs` 32-bit LDS words per lane, either
 * a pair list (stall 0 = drain when one list is full, P = full lanes wait until P percent
 * of the wave waits) or, queue = 1, a queue of pairs / 2 mask words drained in one pass.
 * Any pointer may be NULL. */
sph_status sph_variant_shape(int variant, int *threads, int *unroll, int *pairs, int *stall, int *queue);

/* Times `steps` steps of every (variant, neighbour mode) choice from the current state,
 * keeps the fastest and restores the state, so the call changes nothing but the choice.
 * ms_out (may be NULL) receives 3*SPH_NUM_VARIANTS ms-per-step figures, index
 * 3*variant + mode, negative where a choice was not run. */
sph_status sph_autotune(sph_ctx *ctx, uint32_t steps, float *ms_out);

/* ------------------------------------------------------------------------------
 * Render/readback boundary and profiling export.
 * ------------------------------------------------------------------------------ */

/* Replaces: the blocking readback and CPU un-permutation of
 * CGPUBaseParticleSimulator::integrate (src/CGPUBaseParticleSimulator.cpp:76-97). The
 * device writes the scene's positions (and velocities if with_velocity) by particle id,
 * xyz triples, straight into library-owned pinned host memory (no staging copies);
 * *pos_xyz / *vel_xyz point to it until the next readback of the context or its
 * destruction, *count = the scene's particle count. Ids must be 0..count-1
 * (SPH_ERR_STATE otherwise). */
sph_status sph_scene_readback(sph_ctx *ctx, uint32_t scene, int with_velocity, const float **pos_xyz,
                              const float **vel_xyz, uint32_t *count);

/* One sProfilingEvent (include/SProfilingEvent.h:8-30): device ms of each phase of
 * one step. Fused steps report forces + collisions + integrate under update_forces
 * (one kernel) and 0 for the other two. fps is not measured (the library has no frame
 * loop) and is 0. */
typedef struct {
    uint64_t iteration;
    double fps;
    double update_grid;
    double update_density_pressure;
    double update_forces;
    double update_collisions;
    double integrate;
} sph_profile_event;

/* With SPH_FLAG_PROFILE, every `stride`-th step (0 = none; default 10, the reference's
 * eventLoggerStride, include/CBaseParticleSimulator.h:55) is logged as one event. */
sph_status sph_set_event_stride(sph_ctx *ctx, uint32_t stride);

/* Copies up to cap logged events (out may be NULL) and stores the total in *n; reset
 * clears the log. Needs SPH_FLAG_PROFILE. */
sph_status sph_profile_events(sph_ctx *ctx, sph_profile_event *out, uint32_t cap, uint32_t *n, int reset);

/* Replaces: MainWindow::exportLogs (src/mainwindow.cpp:310-368). Appends the events in
 * its format to <stem>.csv ("label;together;together;...") and <stem>_detail.csv (the
 * label, then one line per phase "Grid;..", "Density + pressure;..", "Forces;..",
 * "Collisions;..", "Integrate;..", then two empty lines); numbers as QString::number
 * (printf %g). No context needed. */
sph_status sph_write_logs(const sph_profile_event *events, uint32_t n, const char *stem, const char *label);

/* sph_profile_events + sph_write_logs for a context. */
sph_status sph_export_logs(sph_ctx *ctx, const char *stem, const char *label);

/* Device bytes one particle slot of a context costs (particle state, sort buffers,
 * density fields, records, lists of max_neighbours entries, its share of the cell
 * table), for sizing a pool against an HBM budget. */
uint64_t sph_slot_bytes(uint32_t max_neighbours);

/* ------------------------------------------------------------------------------
 * Serving. No counterpart in the reference, whose QTimer/doWork loop
 * (src/CBaseParticleSimulator.cpp:35, 146-152) drives one scene: a server owns one
 * batch context and a scheduler thread that admits "advance this scene by k steps"
 * jobs as they arrive and slots free up (continuous batching), advances every running
 * job with the batched step, and retires each job after exactly its k steps, so its
 * result is bitwise the result it has alone. Submission, polling and result calls are
 * thread-safe; all device work runs on the scheduler thread.
 * ------------------------------------------------------------------------------ */

typedef struct sph_server sph_server;

typedef struct {
    uint64_t hbm_budget;       /* device bytes for the pool: slots = budget / sph_slot_bytes */
    uint32_t max_jobs;         /* jobs resident at once (scene ids) */
    uint32_t chunk_steps;      /* most steps between two admission points (0 = 8) */
    uint32_t max_neighbours;   /* as in sph_params */
    uint32_t flags;            /* SPH_FLAG_* of the context */
    int32_t variant;           /* neighbour-kernel variant, -1 = default */
    uint32_t graphs;           /* != 0: replay steps from HIP graphs */
    uint32_t max_bypass;       /* a queued job that does not fit yet may be passed by this many later jobs
                                  that do before admission waits for it (0 = 32) */
} sph_server_params;

typedef struct {
    sph_scene_desc scene;      /* box, gravity, dt, capacity (>= n and the emitter's limit) */
    const float *pos3;         /* n initial positions (copied at submission) */
    const float *vel3;         /* NULL = at rest */
    uint32_t n;
    uint32_t steps;            /* steps to advance */
    const sph_emitter *emitter; /* NULL = none */
} sph_job_desc;

enum { SPH_JOB_QUEUED = 0, SPH_JOB_RUNNING = 1, SPH_JOB_DONE = 2, SPH_JOB_FAILED = 3 };

typedef struct {
    int state;                 /* SPH_JOB_* */
    uint32_t count;            /* particles (final count once done) */
    double submit_s, admit_s, done_s;   /* server clock, seconds */
} sph_job_info;

typedef struct {
    uint64_t submitted, completed, failed, queued, running;
    uint64_t particle_steps;   /* sum over batched steps of the live particles advanced */
    uint64_t batch_steps;      /* batched steps run */
    double busy_s;             /* wall time inside step chunks */
    double elapsed_s;          /* server clock now */
    uint32_t pool, free_slots;
    double admin_s;            /* wall time admitting (uploads, layout) and retiring (readback) jobs */
} sph_server_stats;

sph_status sph_server_create(const sph_server_params *params, sph_server **out);
/* Stops the scheduler (jobs not done are dropped) and frees everything. */
void sph_server_destroy(sph_server *srv);
/* Queues a job; fails with SPH_ERR_CAPACITY if it could never fit the pool. */
sph_status sph_server_submit(sph_server *srv, const sph_job_desc *job, uint64_t *job_id);
sph_status sph_server_job(sph_server *srv, uint64_t job_id, sph_job_info *info);
/* Waits until the job is done or failed (timeout_s < 0: no limit); *info as sph_server_job. */
sph_status sph_server_wait(sph_server *srv, uint64_t job_id, double timeout_s, sph_job_info *info);
/* Copies a done job's particles by id (n = its final count) and forgets the job. A failed
 * job is forgotten too and returns SPH_ERR_FAULT with its error. */
sph_status sph_server_result(sph_server *srv, uint64_t job_id, float *pos3, float *vel3, uint32_t n);
sph_status sph_server_stats_get(sph_server *srv, sph_server_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
