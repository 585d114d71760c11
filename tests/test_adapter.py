"""The Qt drop-in CHIPGridParticleSimulator (gmu-water-simulation_amd/host)
compiles against the reference's own CBaseParticleSimulator interface
(include/CBaseParticleSimulator.h) with the Qt 5 headers of this image. The
compile reads the reference headers as text; no reference code is built or run.
Skipped where the reference checkout or the Qt headers are absent."""
import os
import subprocess

import pytest

from conftest import ROOT

REF = "/root/reference"
QT = "/opt/conda/include/qt"
HOST = os.path.join(ROOT, "gmu-water-simulation_amd", "host")
SRC = os.path.join(HOST, "CHIPGridParticleSimulator.cpp")
CORE = os.path.join(HOST, "sph_simulator_core.cpp")
DRIVER = os.path.join(ROOT, "tests", "hip", "core_driver")


@pytest.mark.skipif(not (os.path.isdir(os.path.join(REF, "include")) and os.path.isdir(QT)),
                    reason="reference headers or Qt 5 headers not present")
def test_adapter_compiles_against_reference_interface():
    mods = ["", "QtCore", "QtGui", "QtWidgets", "QtOpenGL", "Qt3DCore", "Qt3DExtras", "Qt3DRender", "Qt3DInput"]
    incs = ["-I" + os.path.join(QT, m) for m in mods]
    cmd = ["g++", "-std=gnu++11", "-fsyntax-only", "-fPIC", "-w", "-DCL_TARGET_OPENCL_VERSION=120", *incs,
           "-I" + REF, "-I" + os.path.join(REF, "include"), "-I" + os.path.join(ROOT, "include"), "-I" + HOST, SRC]
    out = subprocess.run(cmd, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr[-3000:]


def test_adapter_overrides_every_phase():
    text = open(SRC).read()
    for name in ("updateGrid", "updateDensityPressure", "updateForces", "updateCollisions", "integrate",
                 "setupScene", "getSelectedDevice", "setGravityVector", "step"):
        assert "CHIPGridParticleSimulator::%s(" % name in text


def test_core_compiles_without_qt():
    """The adapter's device half needs no Qt and no reference header."""
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), "-I" + HOST,
           CORE]
    out = subprocess.run(cmd, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr[-3000:]


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="tests/hip/core_driver not built (build())")
def test_core_driver_links_the_library():
    """The C++ driver of the core links libsph_hip.so from the tree (no GPU call:
    the usage message)."""
    out = subprocess.run(["ldd", DRIVER], capture_output=True, text=True)
    lib = [line for line in out.stdout.splitlines() if "libsph_hip.so" in line]
    assert lib and "gmu-water-simulation_amd/lib/libsph_hip.so" in lib[0], out.stdout
    r = subprocess.run([DRIVER], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
