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
// Data flow: the base class owns m_clParticles (id order) and the CParticle
// render entities, which the GPU simulators keep in grid cell (0,0,0). The
// device holds the state in cell-sorted structure-of-arrays buffers; integrate()
// downloads by id into m_clParticles and refreshes the render entities, as
// CGPUBaseParticleSimulator::integrate does (src/CGPUBaseParticleSimulator.cpp:76-97).
// Particles that generateParticles appends (FOUNTAIN) are appended on the device
// by the next updateGrid() instead of a full re-upload.
#ifndef CHIPGRIDPARTICLESIMULATOR_H
#define CHIPGRIDPARTICLESIMULATOR_H

#include <string>
#include <vector>

#include "CBaseParticleSimulator.h"
#include "sph_hip.h"

class CHIPGridParticleSimulator : public CBaseParticleSimulator
{
public:
    explicit CHIPGridParticleSimulator(CScene *scene, float boxSize, SimulationScenario scenario = DAM_BREAK,
                                       QObject *parent = nullptr);
    ~CHIPGridParticleSimulator() override;

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
    // Runs one phase; on failure reports like the OpenCL simulator does on a
    // CLException (src/CGPUBaseParticleSimulator.cpp:28-37).
    double phase(int which);
    bool ok(sph_status status);
    void appendNewParticles();

    sph_ctx *m_ctx = nullptr;
    uint32_t m_onDevice = 0;   // particles of m_clParticles already uploaded
    bool m_failed = false;
    std::string m_error;
    std::vector<float> m_pos, m_vel;
};

#endif
