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
// segment of global cell ids. A single-scene context has one of these.
struct Scene {
    int rx, ry, rz;        // grid resolution (CGrid), cells x + y*rx + z*rx*ry
    uint32_t coff;         // first global cell id of the scene
    uint32_t poff;         // first particle slot of the scene
    uint32_t n;            // live particles (slots poff .. poff+n-1)
    float bx, by, bz;      // box extents (walls at +-b/2)
    float gx, gy, gz;      // gravity
    float dt;
};

// Computes the constants on the host with the reference's expressions.
Consts make_consts();

}  // namespace sph
