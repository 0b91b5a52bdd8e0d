"""Fixture from the reference's own input deck examples/Run_Sims_p0004 (a
DDSCAT shape.dat of 11,508 dipoles and its Simulation_data.dat): the lattice
(i, j, k, m_x, m_y, m_z) parsed by whitespace splitting (independently of the
fixed-width reader under test) and the deck parameters the tests use.
Writes deck_p0004.npz.  Run from the repository root:

    python tests/golden/make_deck_fixture.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DECK = "/root/reference/examples/Run_Sims_p0004/inputs"


def main():
    path = os.path.join(DECK, "shape.dat")
    lines = open(path).read().splitlines()
    count = int(lines[1].split()[0])
    rows = [ln.split() for ln in lines[7:7 + count]]
    lat = np.array([[int(t) for t in r[1:7]] for r in rows], dtype=np.int16)
    assert lat.shape == (count, 6)
    np.savez_compressed(os.path.join(HERE, "deck_p0004.npz"), lattice=lat, count=count, header=np.array(lines[:7]),
                        # Simulation_data.dat: ap/L = 0.700283325 mm, Sc = 50 um, 13.6 GHz, CSH division with
                        # NBlock (avgNcells) = 2000, m from inputs/diel/m1.78_0.00133
                        a_eff_m=0.700283325e-3, freq_hz=13.6e9, m=1.7861 + 0.00133j, navg_cells=2000, sc_nominal_m=50e-6)
    print("wrote deck_p0004.npz", lat.shape)


if __name__ == "__main__":
    main()
