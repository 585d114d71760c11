#include "sph_simulator_core.h"

namespace sph {

GridSimulatorCore::~GridSimulatorCore()
{
    sph_destroy(m_ctx);
}

bool GridSimulatorCore::ok(sph_status status)
{
    if (status == SPH_OK) return true;
    if (!m_failed) m_error = sph_last_error();
    m_failed = true;
    return false;
}

bool GridSimulatorCore::setup(const float box[3], uint32_t capacity, const float gravity[3], float dt, uint32_t flags)
{
    sph_destroy(m_ctx);
    m_ctx = nullptr;
    m_onDevice = 0;
    m_failed = false;
    m_error.clear();
    sph_params params;
    sph_default_params(&params, box[0], capacity ? capacity : 1u);
    params.box[1] = box[1];
    params.box[2] = box[2];
    params.gravity[0] = gravity[0];
    params.gravity[1] = gravity[1];
    params.gravity[2] = gravity[2];
    params.dt = dt;
    params.flags = flags;
    return ok(sph_create(&params, &m_ctx));
}

bool GridSimulatorCore::appendNew(uint32_t count, const Fetch &fetch)
{
    if (m_failed || !m_ctx) return false;
    if (count <= m_onDevice) return true;
    const uint32_t k = count - m_onDevice;
    m_pos.resize(3 * (size_t) k);
    m_vel.resize(3 * (size_t) k);
    m_ids.resize(k);
    for (uint32_t q = 0; q < k; ++q) fetch(m_onDevice + q, &m_pos[3 * (size_t) q], &m_vel[3 * (size_t) q], m_ids[q]);
    if (!ok(sph_append(m_ctx, m_pos.data(), m_vel.data(), m_ids.data(), k))) return false;
    m_onDevice = count;
    return true;
}

double GridSimulatorCore::phase(int which, bool profile)
{
    if (m_failed || !m_ctx) return 0.0;
    float ms = 0.0f;
    if (!ok(sph_phase(m_ctx, which, profile ? &ms : nullptr))) return 0.0;
    return (double) ms;
}

double GridSimulatorCore::updateGrid(uint32_t count, const Fetch &fetch, bool profile)
{
    if (!appendNew(count, fetch)) return 0.0;
    return phase(SPH_PHASE_GRID, profile);
}

double GridSimulatorCore::integrate(const Store &store, bool profile)
{
    const double ms = phase(SPH_PHASE_INTEGRATE, profile);
    if (m_failed) return ms;
    // the device writes positions and velocities by id into the library's pinned buffers
    const float *pos = nullptr, *vel = nullptr;
    uint32_t n = 0;
    if (!ok(sph_scene_readback(m_ctx, 0, 1, &pos, &vel, &n))) return ms;
    for (uint32_t i = 0; i < n; ++i) store(i, pos + 3 * (size_t) i, vel + 3 * (size_t) i);
    return ms;
}

bool GridSimulatorCore::setGravity(float gx, float gy, float gz)
{
    if (m_failed || !m_ctx) return false;
    return ok(sph_set_gravity(m_ctx, gx, gy, gz));
}

std::string GridSimulatorCore::deviceName()
{
    char name[256];
    if (sph_device_name(name, sizeof(name)) != SPH_OK) return std::string("HIP device unavailable: ") + sph_last_error();
    return std::string("HIP ") + name;
}

}  // namespace sph
