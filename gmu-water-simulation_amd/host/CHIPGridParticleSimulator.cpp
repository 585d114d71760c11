#include "CHIPGridParticleSimulator.h"

CHIPGridParticleSimulator::CHIPGridParticleSimulator(CScene *scene, float boxSize, SimulationScenario scenario,
                                                     QObject *parent)
    : CBaseParticleSimulator(scene, boxSize, scenario, parent)
{
}

CHIPGridParticleSimulator::~CHIPGridParticleSimulator()
{
    sph_destroy(m_ctx);
}

bool CHIPGridParticleSimulator::ok(sph_status status)
{
    if (status == SPH_OK) return true;
    m_failed = true;
    m_error = sph_last_error();
    emit errorOccured(m_error.c_str());
    stop();
    return false;
}

QString CHIPGridParticleSimulator::getSelectedDevice()
{
    char name[256];
    if (sph_device_name(name, sizeof(name)) != SPH_OK) return QString("HIP device unavailable: ") + sph_last_error();
    return QString("HIP ") + name;
}

void CHIPGridParticleSimulator::setupScene()
{
    // CBaseParticleSimulator::setupScene fills m_clParticles (dam-break lattice) or
    // sets the fountain's particle budget (src/CBaseParticleSimulator.cpp:38-65).
    CBaseParticleSimulator::setupScene();
    sph_destroy(m_ctx);
    m_ctx = nullptr;
    m_onDevice = 0;
    m_failed = false;
    sph_params params;
    sph_default_params(&params, m_boxSize.x(), m_maxParticlesCount ? m_maxParticlesCount : 1u);
    params.box[1] = m_boxSize.y();
    params.box[2] = m_boxSize.z();
    params.gravity[0] = gravity.x();
    params.gravity[1] = gravity.y();
    params.gravity[2] = gravity.z();
    params.dt = dt;
    if (!ok(sph_create(&params, &m_ctx))) return;
    appendNewParticles();
}

void CHIPGridParticleSimulator::appendNewParticles()
{
    const uint32_t n = (uint32_t) m_particlesCount;
    if (!m_ctx || n <= m_onDevice) return;
    const uint32_t k = n - m_onDevice;
    m_pos.resize(3 * (size_t) k);
    m_vel.resize(3 * (size_t) k);
    std::vector<uint32_t> ids(k);
    for (uint32_t q = 0; q < k; ++q) {
        const CParticle::Physics &p = m_clParticles[m_onDevice + q];
        m_pos[3 * q] = p.position.s[0];
        m_pos[3 * q + 1] = p.position.s[1];
        m_pos[3 * q + 2] = p.position.s[2];
        m_vel[3 * q] = p.velocity.s[0];
        m_vel[3 * q + 1] = p.velocity.s[1];
        m_vel[3 * q + 2] = p.velocity.s[2];
        ids[q] = p.id;
    }
    if (ok(sph_append(m_ctx, m_pos.data(), m_vel.data(), ids.data(), k))) m_onDevice = n;
}

void CHIPGridParticleSimulator::step()
{
    if (m_failed || !m_ctx) return;
    CBaseParticleSimulator::step();
}

void CHIPGridParticleSimulator::setGravityVector(QVector3D newGravity)
{
    CBaseParticleSimulator::setGravityVector(newGravity);
    if (m_ctx) ok(sph_set_gravity(m_ctx, gravity.x(), gravity.y(), gravity.z()));
}

double CHIPGridParticleSimulator::phase(int which)
{
    if (m_failed || !m_ctx) return 0.0;
    float ms = 0.0f;
    if (!ok(sph_phase(m_ctx, which, PROFILING ? &ms : nullptr))) return 0.0;
    return (double) ms;
}

double CHIPGridParticleSimulator::updateGrid()
{
    appendNewParticles();
    return phase(SPH_PHASE_GRID);
}

double CHIPGridParticleSimulator::updateDensityPressure()
{
    return phase(SPH_PHASE_DENSITY);
}

double CHIPGridParticleSimulator::updateForces()
{
    return phase(SPH_PHASE_FORCES);
}

double CHIPGridParticleSimulator::updateCollisions()
{
    return phase(SPH_PHASE_COLLISIONS);
}

double CHIPGridParticleSimulator::integrate()
{
    const double ms = phase(SPH_PHASE_INTEGRATE);
    if (m_failed) return ms;
    // readback by id for rendering (src/CGPUBaseParticleSimulator.cpp:83-91): the device
    // writes positions and velocities by id into the library's pinned buffers
    const float *pos = nullptr, *vel = nullptr;
    uint32_t n = 0;
    if (!ok(sph_scene_readback(m_ctx, 0, 1, &pos, &vel, &n))) return ms;
    for (uint32_t i = 0; i < n; ++i) {
        CParticle::Physics &p = m_clParticles[i];
        p.position.s[0] = pos[3 * i];
        p.position.s[1] = pos[3 * i + 1];
        p.position.s[2] = pos[3 * i + 2];
        p.velocity.s[0] = vel[3 * i];
        p.velocity.s[1] = vel[3 * i + 1];
        p.velocity.s[2] = vel[3 * i + 2];
    }
    for (auto &particle : m_grid->getData()[0]) {
        particle->updatePosition();
        particle->updateVelocity();
    }
    return ms;
}
