"""Writes jkr_roots.json: 50-digit roots of the JKR contact-radius quartic.

a^4 - 2 R d a^2 - (4 pi R^2 g / E) a + R^2 d^2 = 0 (jkrContactRadius,
ABM-3d/ABM-3d-cpp/include/jkr.hpp:67-144); the larger real root is the
contact radius. Roots come from mpmath.polyroots at 50 digits, independent of
the oracle's FP64 Newton iteration.
"""
import json
import os

import mpmath as mp

mp.mp.dps = 50


def larger_root(delta, R, E, gamma):
    d, R, E, g = (mp.mpf(str(x)) for x in (delta, R, E, gamma))
    coeffs = [1, 0, -2 * R * d, -4 * mp.pi * R * R * g / E, R * R * d * d]
    roots = mp.polyroots(coeffs, maxsteps=500, extraprec=200)
    real = [r.real for r in roots if abs(r.imag) < mp.mpf("1e-30")]
    return max(real) if real else None


def main():
    cases = []
    Req, E = 0.4, 3900.0
    for gamma in (10.0, 50.0, 100.0, 500.0, 1000.0):
        for k in range(1, 31):
            delta = 0.6 * k / 30
            cases.append(dict(delta=delta, R=Req, E=E, gamma=gamma,
                              a=mp.nstr(larger_root(delta, Req, E, gamma), 40)))
    # negative overlaps (the surrogate envelope reaches -0.2 R)
    for gamma in (10.0, 100.0, 1000.0):
        for delta in (-0.16, -0.1, -0.05, -0.01):
            r = larger_root(delta, Req, E, gamma)
            cases.append(dict(delta=delta, R=Req, E=E, gamma=gamma,
                              a=None if r is None else mp.nstr(r, 40)))
    # testContactArea.cpp:102-162 (delta = 0.1, R = 0.8, E = 3900, gamma = 100)
    cases.append(dict(delta=0.1, R=0.8, E=3900.0, gamma=100.0,
                      a=mp.nstr(larger_root(0.1, 0.8, 3900.0, 100.0), 40)))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jkr_roots.json")
    with open(out, "w") as f:
        json.dump(cases, f, indent=0)
    print(f"wrote {len(cases)} cases to {out}")


if __name__ == "__main__":
    main()
