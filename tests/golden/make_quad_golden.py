"""Writes tests/golden/quad_vectors.npz: inputs and outputs of the reference's
own quadrature routines, compiled from the reference sources into oracle/_ref
by oracle/Makefile.ref (flang).  Run in a container that has the reference
tree:

    make -C oracle ref && python tests/golden/make_quad_golden.py

Routines captured (reference file:line):
  getAdaptiveQuadPts2       Numerical_Integration.f90:213-245 with QUADPACK
                            qag / qage / qk61 / qsort (quadpack.f90), single
                            precision, integrand read from a grid
  trap_2Dc                  Numerical_Integration.f90:251-305
  getGaussLegendreQuadPts3  Numerical_Integration.f90:88-145, LegendreRuleFast.f90

The fixture holds data only (inputs and expected outputs).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as o  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "quad_vectors.npz")


def main():
    rng = np.random.default_rng(2026)
    d = {}
    # adaptive quadrature: theta columns (a0 = 0, b0 = pi) and phi vectors (0, 2 pi)
    cases = []
    for c in range(24):
        nth = [19, 37, 91, 181][c % 4]
        nph = [1, 13, 25, 37][(c // 4) % 4]
        th = np.linspace(0, np.pi, nth)
        kind = c % 3
        if kind == 0:
            f = np.outer(np.sin(th) * (1 + np.cos(3 * th)) ** 2, rng.uniform(0.5, 2.0, nph))
        elif kind == 1:
            f = rng.uniform(0.0, 1.0, (nth, nph)) * 10.0 ** rng.integers(-8, 4)
        else:
            f = np.outer(np.exp(-20 * (th - 1.0) ** 2) * np.sin(th), np.ones(nph)) * 1e3
        ph = bool(c % 5 == 4 and nph > 1)
        if ph:
            f = rng.uniform(0.0, 1.0, (nph, 1))
            jj, a0, b0 = 1, 0.0, 2 * np.pi
        else:
            jj, a0, b0 = int(rng.integers(1, nph + 1)), 0.0, np.pi
        res = o.ref_adaptive_quad(f, nth, nph, ph, jj, a0, b0)
        d[f"aq{c}_f"] = f
        d[f"aq{c}_args"] = np.array([nth, nph, int(ph), jj])
        d[f"aq{c}_ab"] = np.array([a0, b0])
        d[f"aq{c}_res"] = np.array(res)
        cases.append(c)
    d["aq_n"] = np.array(len(cases))
    # trapezoid cubature, including one phi node (0 / 0 width) and equal phi bounds
    tr = [((37, 25), 0.0, np.pi, 0.0, 2 * np.pi), ((19, 1), 0.0, np.pi, 0.0, 0.0), ((13, 7), 0.0, np.pi, 0.3, 0.3)]
    for t, (shape, a, b, cc, dd) in enumerate(tr):
        f = rng.uniform(0.0, 1.0, shape)
        d[f"tr{t}_f"] = f
        d[f"tr{t}_abcd"] = np.array([a, b, cc, dd])
        d[f"tr{t}_res"] = np.array(o.ref_trap2d(f, a, b, cc, dd))
    d["tr_n"] = np.array(len(tr))
    # Gauss-Legendre rules
    gl = [(1, 0.0, np.pi), (2, 0.0, np.pi), (7, 0.2, 2.9), (19, 0.0, np.pi), (37, 0.0, 2 * np.pi), (181, 0.0, np.pi)]
    for t, (n, a, b) in enumerate(gl):
        x, w = o.ref_gauss_legendre(n, a, b)
        d[f"gl{t}_nab"] = np.array([n, a, b])
        d[f"gl{t}_x"] = x
        d[f"gl{t}_w"] = w
    d["gl_n"] = np.array(len(gl))
    np.savez_compressed(OUT, **d)
    print("wrote", OUT, len(d), "arrays")


if __name__ == "__main__":
    main()
