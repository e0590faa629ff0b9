"""Writes tests/golden/delhommeau_mpmath.json: high-precision reference values.

Independent of the oracle and of the reference code: the exact
Guevel-Delhommeau wave term is the integral (reference
src/green_functions/exact_Guevel_Delhommeau.jl:15-46)

    G(r, z)     = 4 int_0^{pi/2} [Re J + i Re e^zeta] dtheta,
    dG/dr(r, z) = 4 int_0^{pi/2} [Re(i cos t (J - 1/(pi zeta))) + i Re(i cos t e^zeta)] dtheta,
    J = e^zeta (E1(zeta)/pi + i),  zeta = z + i r cos(theta),

evaluated here with mpmath's tanh-sinh quadrature and its own E1 at 30
significant digits.  e^zeta E1(zeta) itself is also tabulated on a grid of
the closed upper-left quadrant the term samples.

Run:  python tests/golden/make_golden.py   (about a minute)
"""
import json
import os

import mpmath as mp

mp.mp.dps = 30
HERE = os.path.dirname(os.path.abspath(__file__))


def dimless(r, z):
    r, z = mp.mpf(r), mp.mpf(z)

    def parts(t):
        zeta = mp.mpc(z, r * mp.cos(t))
        ez = mp.exp(zeta)
        j = ez * (mp.e1(zeta) / mp.pi + 1j)
        return (mp.re(j), mp.re(ez), mp.re(1j * mp.cos(t) * (j - 1 / (mp.pi * zeta))),
                mp.re(1j * mp.cos(t) * ez))

    vals = [4 * mp.quad(lambda t, w=w: parts(t)[w], [0, mp.pi / 4, mp.pi / 2]) for w in range(4)]
    return [float(v) for v in vals]


def main():
    pts = []
    for r in [0.0, 1e-3, 0.05, 0.3, 1.0, 2.5, 6.0, 15.0, 40.0]:
        for z in [-1e-3, -0.02, -0.3, -1.0, -3.0, -8.0, -25.0]:
            pts.append((r, z))
    rows = []
    for r, z in pts:
        g = dimless(r, z)
        rows.append({"r": r, "z": z, "G": [g[0], g[1]], "dGdr": [g[2], g[3]]})
    e1 = []
    for x in [-60.0, -20.0, -9.0, -4.0, -1.0, -0.2, -1e-3, 0.0]:
        for y in [0.0, 1e-6, 0.3, 2.0, 7.0, 12.0, 45.0]:
            if x == 0.0 and y == 0.0:
                continue
            zc = mp.mpc(x, y)
            v = mp.exp(zc) * mp.e1(zc)
            e1.append({"z": [x, y], "expE1": [float(mp.re(v)), float(mp.im(v))]})
    with open(os.path.join(HERE, "delhommeau_mpmath.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py (mpmath, 30 digits)",
                   "dimless": rows, "exp_e1": e1}, f, indent=1)


if __name__ == "__main__":
    main()
