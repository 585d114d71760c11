// sph_simulator_core.h -- the Qt-free part of the Qt drop-in
// (CHIPGridParticleSimulator): one libsph_hip.so context driven the way
// CBaseParticleSimulator drives its subclasses (include/CBaseParticleSimulator.h:
// 29-110). The caller keeps the host particle array in id order, as the base
// class keeps m_clParticles, and reaches it through two callbacks; the core
// appends the array's new tail to the device before each grid phase, maps the
// five phase virtuals onto sph_phase, and after integrate writes the device
// state back into the array by id, as CGPUBaseParticleSimulator::integrate does
// (src/CGPUBaseParticleSimulator.cpp:76-97).
//
// Errors: the first failing library call puts the core in a failed state (every
// later call is a no-op returning 0 / false) and keeps the library's message,
// which the Qt adapter reports as the OpenCL simulator reports a CLException
// (src/CGPUBaseParticleSimulator.cpp:28-37).
#ifndef SPH_SIMULATOR_CORE_H
#define SPH_SIMULATOR_CORE_H

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "sph_hip.h"

namespace sph {

class GridSimulatorCore
{
public:
    // Reads particle i (id order) of the caller's array.
    using Fetch = std::function<void(uint32_t i, float pos[3], float vel[3], uint32_t &id)>;
    // Writes the device state of the particle with id i into the caller's array.
    using Store = std::function<void(uint32_t i, const float pos[3], const float vel[3])>;

    GridSimulatorCore() = default;
    ~GridSimulatorCore();
    GridSimulatorCore(const GridSimulatorCore &) = delete;
    GridSimulatorCore &operator=(const GridSimulatorCore &) = delete;

    // setupScene's device half: a new context for `capacity` particles
    // (m_maxParticlesCount) in a box of edges box[0..2], with gravity and dt, and
    // flags (SPH_FLAG_*, e.g. SPH_FLAG_REFERENCE_ORDER). Drops the previous one.
    bool setup(const float box[3], uint32_t capacity, const float gravity[3], float dt, uint32_t flags = 0);

    // Appends particles [onDevice(), count) of the caller's array to the device.
    bool appendNew(uint32_t count, const Fetch &fetch);

    // The phase virtuals. updateGrid appends the array's new tail first
    // (particles generateParticles added since the last step). Each returns the
    // phase's device time in ms when `profile` is set, else 0.
    double updateGrid(uint32_t count, const Fetch &fetch, bool profile = false);
    double updateDensityPressure(bool profile = false) { return phase(SPH_PHASE_DENSITY, profile); }
    double updateForces(bool profile = false) { return phase(SPH_PHASE_FORCES, profile); }
    double updateCollisions(bool profile = false) { return phase(SPH_PHASE_COLLISIONS, profile); }
    // integrate, then the readback by id through `store` (every live particle).
    double integrate(const Store &store, bool profile = false);

    // setGravityVector's device half (src/CGPUBaseParticleSimulator.cpp:11-15).
    bool setGravity(float gx, float gy, float gz);

    // getSelectedDevice: "HIP <device name>", or why there is none.
    static std::string deviceName();

    bool failed() const { return m_failed; }
    const std::string &error() const { return m_error; }
    uint32_t onDevice() const { return m_onDevice; }
    sph_ctx *context() const { return m_ctx; }

private:
    double phase(int which, bool profile);
    bool ok(sph_status status);

    sph_ctx *m_ctx = nullptr;
    uint32_t m_onDevice = 0;
    bool m_failed = false;
    std::string m_error;
    std::vector<float> m_pos, m_vel;
    std::vector<uint32_t> m_ids;
};

}  // namespace sph

#endif
