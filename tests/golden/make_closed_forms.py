"""Writes tests/golden/closed_forms.json: exact-arithmetic values the oracle is
checked against. Nothing here runs or reads the reference; every value is a
closed form of the SPH model of the reference (Mueller, Charypar, Gross 2003
kernels as used in src/CCPUParticleSimulator.cpp:9-30, walls of
src/CCollisionGeometry.cpp:117-132, integrator of src/CCPUParticleSimulator.cpp:
220-221) evaluated with Python Fractions/Decimal at 50 digits, with the
reference's constants taken at their exact float32 values.

    python tests/golden/make_closed_forms.py
"""
import json
import os
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np

getcontext().prec = 60
PI = Decimal("3.14159265358979323846264338327950288419716939937510582097494459")

F = lambda x: Fraction(float(np.float32(x)))  # exact value of a float32 constant
H = F(0.0457)
MASS = F(0.02)
K_GAS = F(3.0)
RHO0 = F(998.29)
MU = F(3.5)
G = F(-9.80665)
DT = F(0.01)
HSQ_F = Fraction(float(np.float32(np.float32(0.0457) * np.float32(0.0457))))  # fl(h*h)


def dec(fr):
    return Decimal(fr.numerator) / Decimal(fr.denominator)


def poly6():
    return Decimal(315) / (Decimal(64) * PI * dec(H) ** 9)


def spiky():
    return Decimal(-45) / (PI * dec(H) ** 6)


def visc():
    return Decimal(45) / (PI * dec(H) ** 6)


def lattice_interior():
    """Perfect cubic lattice of spacing h/2: every site with i^2+j^2+k^2 <= 4
    (r <= h) is a neighbour; the self term included."""
    s2 = (H / 2) ** 2
    total = Decimal(0)
    count = 0
    for i in range(-2, 3):
        for j in range(-2, 3):
            for k in range(-2, 3):
                r2 = (i * i + j * j + k * k) * s2
                if r2 <= HSQ_F:
                    count += 1
                    total += dec((HSQ_F - r2) ** 3)
    return count, dec(MASS) * poly6() * total


def isolated_density():
    return dec(MASS) * poly6() * dec(HSQ_F) ** 3


def pair(r, u):
    """Two particles at x=0 and x=r (r < h), velocities 0 and (u,0,0): density,
    pressure and the x-acceleration of the particle at x=0 (no walls in reach)."""
    r = Fraction(r)
    u = Fraction(u)
    rho = dec(MASS) * poly6() * (dec(HSQ_F) ** 3 + dec(HSQ_F - r * r) ** 3)
    p = dec(K_GAS) * (rho - dec(RHO0))
    hr = dec(H) - dec(r)
    # grad W_01 = spiky (h-r)^2 (x0 - x1)/r, (x0 - x1)/r = -1
    grad_x = spiky() * hr * hr * Decimal(-1)
    f_p = (2 * p / (rho * rho)) * grad_x
    a_press = -dec(MASS) * rho * f_p / rho
    a_visc = dec(MU) * dec(MASS) * (dec(u) - 0) * visc() * hr / rho / rho
    return rho, p, a_press + a_visc


def main():
    out = {
        "h": str(dec(H)),
        "hsq_f_bits": "0x%08X" % int(np.array(np.float32(np.float32(0.0457) * np.float32(0.0457))).view(np.uint32)),
        "poly6": str(poly6()),
        "spiky": str(spiky()),
        "visc": str(visc()),
        "isolated_density": str(isolated_density()),
    }
    cnt, rho = lattice_interior()
    out["lattice_interior_count"] = cnt
    out["lattice_interior_density"] = str(rho)
    pairs = []
    for r, u in [(0.01, 0.0), (0.02, 0.5), (0.03, -1.25), (0.04, 2.0)]:
        rf = float(np.float32(r))
        uf = float(np.float32(u))
        rho_p, p_p, ax = pair(rf, uf)
        pairs.append({"r": rf, "u": uf, "density": str(rho_p), "pressure": str(p_p), "ax_sph": str(ax)})
    out["pairs"] = pairs
    # free fall of an isolated particle from rest under g, symplectic Euler:
    # v_n = n g dt, y_n = y_0 + g dt^2 n(n+1)/2
    out["free_fall"] = [{"n": n, "dy": str(dec(G * DT * DT * n * (n + 1) / 2)), "vy": str(dec(G * DT * n))}
                        for n in (1, 2, 5, 20)]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "closed_forms.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
