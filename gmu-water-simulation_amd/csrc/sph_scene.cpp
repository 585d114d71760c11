// sph_scene.cpp -- deterministic host-side scene generators of libsph_hip.so.

#include <cmath>
#include <cstring>

#include "../../include/sph_hip.h"
#include "sph_common.h"

namespace {

// SplitMix64 (Steele, Lea, Flood 2014): the benchmark scenes' only randomness.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // 24-bit uniform in [0, 1)
    double uniform() { return (double) (next() >> 40) * (1.0 / 16777216.0); }
};

}  // namespace

namespace sph {
sph_status host_exception() noexcept;   // the status of a caught C++ exception (sph_capi.hip)
}

extern "C" {

sph_status sph_scene_dam_break(float box, float *pos3, uint32_t cap, uint32_t *n)
{
    try {
        if (!n || !(box > 0.0f)) return SPH_ERR_INVALID;
        // src/CBaseParticleSimulator.cpp:40-56: lattice at spacing h/2 filling x in [0, box/4),
        // all of y and z, from the corner -box/2; loop variables are float stepped in double.
        const double half = (double) (sph::kH / 2.0f);
        const float off = -box / 2.0f;
        uint32_t count = 0;
        for (float y = 0; y < box; y = (float) ((double) y + half)) {
            for (float x = 0; (double) x < (double) box / 4.0; x = (float) ((double) x + half)) {
                for (float z = 0; z < box; z = (float) ((double) z + half)) {
                    if (pos3) {
                        if (count >= cap) return SPH_ERR_CAPACITY;
                        pos3[3 * (size_t) count + 0] = x + off;
                        pos3[3 * (size_t) count + 1] = y + off;
                        pos3[3 * (size_t) count + 2] = z + off;
                    }
                    ++count;
                }
            }
        }
        *n = count;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_block(const float origin[3], uint32_t nx, uint32_t ny, uint32_t nz, uint32_t count,
                           float jitter_h, uint64_t seed, float *pos3)
{
    try {
        if (!origin || (!pos3 && count > 0)) return SPH_ERR_INVALID;
        if ((uint64_t) count > (uint64_t) nx * ny * nz) return SPH_ERR_INVALID;
        const double half = (double) (sph::kH / 2.0f);
        float cx = 0, cy = 0, cz;
        size_t id = 0;
        for (uint32_t j = 0; j < ny && id < count; ++j) {
            cx = 0;
            for (uint32_t i = 0; i < nx && id < count; ++i) {
                cz = 0;
                for (uint32_t k = 0; k < nz && id < count; ++k) {
                    pos3[3 * id + 0] = cx + origin[0];
                    pos3[3 * id + 1] = cy + origin[1];
                    pos3[3 * id + 2] = cz + origin[2];
                    ++id;
                    cz = (float) ((double) cz + half);
                }
                cx = (float) ((double) cx + half);
            }
            cy = (float) ((double) cy + half);
        }
        SplitMix64 rng(seed);
        const double amp = (double) jitter_h * (double) sph::kH;
        for (size_t p = 0; p < id; ++p)
            for (int a = 0; a < 3; ++a) {
                const float j = (float) ((2.0 * rng.uniform() - 1.0) * amp);
                pos3[3 * p + a] = pos3[3 * p + a] + j;
            }
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_lattice(uint32_t nx, uint32_t ny, uint32_t nz, const float box[3], float jitter_h,
                             uint64_t seed, float *pos3)
{
    try {
        if (!box || !pos3) return SPH_ERR_INVALID;
        const float origin[3] = {-box[0] / 2.0f, -box[1] / 2.0f, -box[2] / 2.0f};
        const uint64_t count = (uint64_t) nx * ny * nz;
        if (count > 0xFFFFFFFFull) return SPH_ERR_INVALID;
        return sph_scene_block(origin, nx, ny, nz, (uint32_t) count, jitter_h, seed, pos3);
    } catch (...) {
        return sph::host_exception();
    }
}

}  // extern "C"
