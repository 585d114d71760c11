// sph_kernels.h -- launch wrappers of the gfx950 grid-SPH kernels (sph_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_common.h"

namespace sph {

// Structure-of-arrays particle buffers in HBM, in the current array order.
struct ParticleSoA {
    float *px, *py, *pz;
    float *vx, *vy, *vz;
    uint32_t *id;
};

constexpr int kRadixBits = 8;
constexpr int kRadixBuckets = 1 << kRadixBits;
constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;

inline uint32_t sort_blocks(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }

// Cell key of every live particle (hash of src/CCPUParticleSimulator.cpp:46-69).
void launch_hash(const ParticleSoA &p, uint32_t *key, const Scene &s, double h, hipStream_t st);

// One stable LSD radix pass over `bits` bits starting at `shift`. vals_in == nullptr
// means the identity (val = slot). hist must hold kRadixBuckets * sort_blocks(n).
void launch_radix_pass(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t *keys_out,
                       uint32_t *vals_out, uint32_t n, int shift, uint32_t *hist, hipStream_t st);

// start[c] = first slot whose sorted key is >= c, c in [0, ncells].
void launch_cell_start(const uint32_t *sorted_keys, uint32_t n, uint32_t ncells, uint32_t *start,
                       hipStream_t st);

// dst[k] = src[perm[k]] for positions, velocities and ids.
void launch_reorder(const ParticleSoA &src, const ParticleSoA &dst, const uint32_t *perm, uint32_t n,
                    hipStream_t st);

// Density and p/rho^2 (double) of every particle; ncount may be null.
void launch_density(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start, float *rho,
                    double *prr, uint32_t *ncount, const Scene &s, const Consts &k, hipStream_t st);

// Pressure + viscosity + gravity acceleration into acc (the updateForces phase).
void launch_forces(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start, const float *rho,
                   const double *prr, float *ax, float *ay, float *az, const Scene &s, const Consts &k,
                   hipStream_t st);

// acc += wall response (the updateCollisions phase).
void launch_collide(const ParticleSoA &p, float *ax, float *ay, float *az, const Scene &s, const Consts &k,
                    hipStream_t st);

// Integrate from acc into out and hash the new positions into key_next (the integrate phase).
void launch_integrate(const ParticleSoA &p, const float *ax, const float *ay, const float *az,
                      const ParticleSoA &out, uint32_t *key_next, const Scene &s, const Consts &k,
                      hipStream_t st);

// Fused forces + walls + integrate + next hash. acc pointers may be null (then not stored).
void launch_forces_fused(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start,
                         const float *rho, const double *prr, const ParticleSoA &out, uint32_t *key_next,
                         float *ax, float *ay, float *az, const Scene &s, const Consts &k, hipStream_t st);

}  // namespace sph
