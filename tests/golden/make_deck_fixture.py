"""Fixtures from the reference's own input decks examples/Run_Sims_p0004 and
examples/Run_Sims_a0006 (DDSCAT shape.dat files of 11,508 and 140,896 dipoles
and their Simulation_data.dat): the lattice (i, j, k, m_x, m_y, m_z) parsed by
whitespace splitting (independently of the fixed-width reader under test) and
the deck parameters the tests and the deck benchmark use.  Writes
deck_p0004.npz and deck_a0006.npz.  Run from the repository root:

    python tests/golden/make_deck_fixture.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DECK = "/root/reference/examples/Run_Sims_p0004/inputs"


def lattice(deck):
    lines = open(os.path.join(deck, "shape.dat")).read().splitlines()
    count = int(lines[1].split()[0])
    rows = [ln.split() for ln in lines[7:7 + count]]
    lat = np.array([[int(t) for t in r[1:7]] for r in rows], dtype=np.int16)
    assert lat.shape == (count, 6)
    return lat, count, lines[:7]


def main():
    lat, count, header = lattice("/root/reference/examples/Run_Sims_a0006/inputs")
    np.savez_compressed(os.path.join(HERE, "deck_a0006.npz"), lattice=lat, count=count, header=np.array(header),
                        # Simulation_data.dat: ap/L = 1.614011963 mm, Sc = 50 um, 94 GHz, CSH division with
                        # NBlock (avgNcells) = 2000, IPWs sd 422, Ninc sd 482, Nscat sd 482, m from
                        # inputs/diel/m1.78_0.00133
                        a_eff_m=1.614011963e-3, freq_hz=94e9, m=1.7861 + 0.00133j, navg_cells=2000,
                        sc_nominal_m=50e-6, n_iws=422, n_inc=482, n_scat=482)
    print("wrote deck_a0006.npz", lat.shape)
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
