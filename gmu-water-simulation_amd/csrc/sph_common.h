// sph_common.h -- definitions shared by the host side of libsph_hip.so and its
// gfx950 kernels: the physics constants as the reference computes them and the
// per-scene description the kernels read.
#pragma once

#include <stdint.h>

namespace sph {

// Physics constants, include/CParticle.h:80-84, include/CBaseParticleSimulator.h:19,
// src/CBaseParticleSimulator.cpp:7, include/CCollisionGeometry.h:20-21.
constexpr float kH = 0.0457f;
constexpr float kViscosity = 3.5f;
constexpr float kMass = 0.02f;
constexpr float kGasStiffness = 3.0f;
constexpr float kRestDensity = 998.29f;
constexpr float kGravityY = -9.80665f;
constexpr float kDt = 0.01f;
constexpr double kWallK = 10000.0;
constexpr double kWallDamping = -0.9;

// Kernel coefficients with the reference's precision: the SPH kernels are double
// (src/CCPUParticleSimulator.cpp:9-30), everything meeting a QVector3D is float.
struct Consts {
    double poly6;     // 315 / (64 pi h^9)
    double spiky;     // -45 / (pi h^6)
    double visc;      // 45 / (pi h^6)
    double hsq;       // (double) fl(h * h)
    double h;         // (double) h
    float hsq_f;      // fl(h * h), the neighbour threshold (bits 0x3B08DF0C)
    float mass;       // m
    float neg_mass;   // -m   (the `-CParticle::mass` factor)
    float mu_mass;    // fl(mu * m)
    float gas_k;      // k
    float rho0;       // rest density
    float wall_k_f;   // (float) WALL_K
};

// One scene: a box with its own uniform grid, a segment of particle slots and a
// segment of global cell ids. The segment of cells ends with one extra "dead"
// cell (coff + ncells) that unused slots sort into, so that after every sort a
// scene's live particles are the first n slots of its segment. The kernels read
// scenes from a table in device memory, so live counts can change on the device
// and a captured step stays valid. A free range of slots is a scene with no cells
// and no particles.
// Cell numbering and summation order of a scene.
enum SceneOrder : uint32_t {
    // cells x + y*rx + z*rx*ry; a cell's particles in the stable order of the
    // previous array order; neighbour rows z (outer), y, then the run of x cells
    kOrderCanonical = 0,
    // the reference's own order (src/CCPUParticleSimulator.cpp:37-88, 108-121):
    // cells numbered in its traversal order T = z + y*rz + x*rz*ry, a cell's
    // particles in its swap-remove bucket order, neighbour rows x (outer), y,
    // then the run of z cells (contiguous in T)
    kOrderReference = 1,
};

struct Scene {
    int rx, ry, rz;        // grid resolution (CGrid)
    uint32_t ncells;       // rx * ry * rz
    uint32_t coff;         // first global cell id of the scene
    uint32_t poff;         // first particle slot of the scene
    uint32_t cap;          // slots of the scene
    uint32_t n;            // live particles (slots poff .. poff+n-1 after a sort)
    float bx, by, bz;      // box extents (walls at +-b/2)
    float gx, gy, gz;      // gravity
    float dt;
    uint32_t order;        // SceneOrder
};
static_assert(sizeof(Scene) == 64, "Scene is read as 16 dwords");

// Slots are grouped in runs of kGroupSlots; a scene's segment starts on a group
// boundary, so every workgroup of at most kGroupSlots threads lies in one scene.
constexpr uint32_t kGroupSlots = 256;

// The scenes of a launch as the kernels see them: slot i belongs to scene
// group_scene[i / kGroupSlots] (group_scene == nullptr: every slot belongs to
// scene 0); slots at and above nslots are not touched.
struct SceneTable {
    const Scene *scenes;
    const uint32_t *group_scene;
    uint32_t nslots;
};

// A particle source that runs on the device at the start of every step and adds
// particles to one scene (ids n, n+1, ... as CBaseParticleSimulator::addParticle
// numbers them, src/CBaseParticleSimulator.cpp:67-74).
enum EmitterKind : uint32_t {
    kEmitFountain = 0,  // the reference fountain: 7 particles per step at the floor centre
    kEmitDisc = 1,      // `rate` particles per step uniformly on a floor disc, seeded
};

struct Emitter {
    uint32_t scene;        // table index of the scene
    uint32_t kind;         // EmitterKind
    uint32_t rate;         // particles per step (disc)
    uint32_t limit;        // particle budget (fountain: m_maxParticlesCount; disc: stop at limit)
    float cx, cz;          // disc centre on the floor
    float radius;          // disc radius
    float speed;           // mean upward speed (disc)
    float spread;          // relative speed jitter: v = speed * (1 + spread * (2u - 1))
    uint32_t seed_lo, seed_hi;
    uint32_t pad[5];
};
static_assert(sizeof(Emitter) == 64, "Emitter is 64 bytes");

// Computes the constants on the host with the reference's expressions.
Consts make_consts();

}  // namespace sph
