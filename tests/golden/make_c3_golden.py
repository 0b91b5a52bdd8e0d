"""Writes c3_dlca64.npz: the oracle's CBFM-E + ACA solution of the full C3
workload (DLCA aggregate of 64 monomers of ~300 cells, seed 3; 19,160 cells,
57,480 unknowns; SR-sparsified block solves with res_SR 0.01 %; CBFs from the
reference's uniform grid of 422 requested directions (435); ACA eps 1e-4,
kmax 50; one plane wave theta = phi = 0).  The oracle takes minutes on the
host at this size, so its result is kept as a fixture for the GPU parity test.

    python tests/golden/make_c3_golden.py        # from the repository root
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from midas_amd import geometry as g  # noqa: E402
from oracle import oracle as o  # noqa: E402


def main():
    t0 = time.time()
    s = g.dlca_aggregate(nmono=64, cells_per_mono=300, seed=3)
    th_c, ph_c = g.cbf_directions_uniform(422)
    th, ph = np.array([0.0]), np.array([0.0])
    c = o.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    fs = []
    E, K, _ = o.cbfm(c, s.blocks, s.ext, th_c, ph_c, th, ph, use_aca=True, sr=True, res_sr=0.01, fsr_out=fs)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c3_dlca64.npz")
    np.savez_compressed(out, E=E, K=np.asarray(K, np.int32), fsr=np.asarray(fs), n_cells=s.n,
                        n_cbf_dirs=len(th_c))
    print(f"{out}: {s.n} cells, K total {int(np.sum(K))}, {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
