#include "CHIPGridParticleSimulator.h"

CHIPGridParticleSimulator::CHIPGridParticleSimulator(CScene *scene, float boxSize, SimulationScenario scenario,
                                                     QObject *parent, uint32_t flags)
    : CBaseParticleSimulator(scene, boxSize, scenario, parent), m_flags(flags)
{
}

double CHIPGridParticleSimulator::checked(double ms)
{
    if (m_core.failed() && !m_reported) {
        m_reported = true;
        emit errorOccured(m_core.error().c_str());
        stop();
    }
    return ms;
}

QString CHIPGridParticleSimulator::getSelectedDevice()
{
    return QString::fromStdString(sph::GridSimulatorCore::deviceName());
}

void CHIPGridParticleSimulator::setupScene()
{
    // CBaseParticleSimulator::setupScene fills m_clParticles (dam-break lattice) or
    // sets the fountain's particle budget (src/CBaseParticleSimulator.cpp:38-65).
    CBaseParticleSimulator::setupScene();
    const float box[3] = {m_boxSize.x(), m_boxSize.y(), m_boxSize.z()};
    const float g[3] = {gravity.x(), gravity.y(), gravity.z()};
    m_reported = false;
    m_core.setup(box, (uint32_t) m_maxParticlesCount, g, dt, m_flags);
    checked(0.0);
}

void CHIPGridParticleSimulator::step()
{
    if (m_core.failed() || !m_core.context()) return;
    CBaseParticleSimulator::step();
}

void CHIPGridParticleSimulator::setGravityVector(QVector3D newGravity)
{
    CBaseParticleSimulator::setGravityVector(newGravity);
    if (m_core.context()) m_core.setGravity(gravity.x(), gravity.y(), gravity.z());
    checked(0.0);
}

double CHIPGridParticleSimulator::updateGrid()
{
    const auto fetch = [this](uint32_t i, float pos[3], float vel[3], uint32_t &id) {
        const CParticle::Physics &p = m_clParticles[i];
        for (int c = 0; c < 3; ++c) {
            pos[c] = p.position.s[c];
            vel[c] = p.velocity.s[c];
        }
        id = p.id;
    };
    return checked(m_core.updateGrid((uint32_t) m_particlesCount, fetch, PROFILING));
}

double CHIPGridParticleSimulator::updateDensityPressure()
{
    return checked(m_core.updateDensityPressure(PROFILING));
}

double CHIPGridParticleSimulator::updateForces()
{
    return checked(m_core.updateForces(PROFILING));
}

double CHIPGridParticleSimulator::updateCollisions()
{
    return checked(m_core.updateCollisions(PROFILING));
}

double CHIPGridParticleSimulator::integrate()
{
    const auto store = [this](uint32_t i, const float pos[3], const float vel[3]) {
        CParticle::Physics &p = m_clParticles[i];
        for (int c = 0; c < 3; ++c) {
            p.position.s[c] = pos[c];
            p.velocity.s[c] = vel[c];
        }
    };
    const double ms = checked(m_core.integrate(store, PROFILING));
    if (m_core.failed()) return ms;
    for (auto &particle : m_grid->getData()[0]) {
        particle->updatePosition();
        particle->updateVelocity();
    }
    return ms;
}
