// CHIPGridParticleSimulator -- the grid simulator of the Qt application, with
// every phase running on an MI355X through libsph_hip.so (include/sph_hip.h).
//
// A drop-in for CGPUParticleSimulator: it derives directly from
// CBaseParticleSimulator (include/CBaseParticleSimulator.h:29-110) and overrides
// the same virtuals, so MainWindow::createSimulator (src/mainwindow.cpp:171-201)
// can construct it in place of `new CGPUParticleSimulator(...)` at :190. It does
// not derive from CGPUBaseParticleSimulator, which is bound to OpenCL
// (include/CGPUBaseParticleSimulator.h:22-35).
//
// The device work is sph::GridSimulatorCore (sph_simulator_core.h, no Qt); this
// class gives it the base class's state. The base class owns m_clParticles (id
// order) and the CParticle render entities, which the GPU simulators keep in
// grid cell (0,0,0). updateGrid() appends the particles generateParticles added
// (FOUNTAIN) to the device; integrate() writes the device state back into
// m_clParticles by id and refreshes the render entities, as
// CGPUBaseParticleSimulator::integrate does (src/CGPUBaseParticleSimulator.cpp:76-97).
#ifndef CHIPGRIDPARTICLESIMULATOR_H
#define CHIPGRIDPARTICLESIMULATOR_H

#include "CBaseParticleSimulator.h"
#include "sph_simulator_core.h"

class CHIPGridParticleSimulator : public CBaseParticleSimulator
{
public:
    // flags: SPH_FLAG_* of the context, e.g. SPH_FLAG_REFERENCE_ORDER for sums in
    // CCPUParticleSimulator's own order.
    explicit CHIPGridParticleSimulator(CScene *scene, float boxSize, SimulationScenario scenario = DAM_BREAK,
                                       QObject *parent = nullptr, uint32_t flags = 0);

    QString getSelectedDevice() override;
    void setupScene() override;
    void step() override;
    void setGravityVector(QVector3D newGravity) override;

protected:
    double updateGrid() override;
    double updateDensityPressure() override;
    double updateForces() override;
    double updateCollisions() override;
    double integrate() override;

private:
    // After a core call: a failure is reported once, like a CLException
    // (src/CGPUBaseParticleSimulator.cpp:28-37), and stops the timer.
    double checked(double ms);

    sph::GridSimulatorCore m_core;
    uint32_t m_flags;
    bool m_reported = false;
};

#endif
