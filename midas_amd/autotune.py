"""Autotuner of the library's schedule parameters on the GPU it runs on.

    python -m midas_amd.autotune [--quick]

Times the MoM solve (midas_compute_efields_mom; the factorisation and
refinement times the library measures on its stream) on seeded Gaussian
random spheres of several sizes while it varies, one at a time (coordinate
descent, largest size first, the values of a parameter interleaved over the
repetitions), the parameters midas_tune_set exposes:

  lu_nb_split         outer LU step of the single-term split modes (512: one
                      panel per step; 768, 1024: two panels of half the width)
  lu_two_panel_min_n  the smallest order at which that wider step pays
  panel_leaf          columns of a recursive panel's leaves
  trail_group_kib     L2 share of a row-tile group of the split GEMM
  lu_nb_fp64          panel width of the FP64 LU
  refine_graph        refinement sweeps replayed from a captured hipGraph

It writes every timing to midas_amd/tuning.json and the chosen values to
midas_amd/csrc/midas_tuned.h, the library's compiled-in defaults (rebuild to
take them).  The reference has no counterpart: its ScaLAPACK blocks are fixed
at 32 x 32 (Initialization.f90:66-67).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

from midas_amd import api, geometry as g  # noqa: E402

HEADER = os.path.join(HERE, "csrc", "midas_tuned.h")
TABLE = os.path.join(HERE, "tuning.json")

COMMENTS = {
    "lu_nb_split": "LU outer step, single-term split modes, n >= lu_two_panel_min_n",
    "lu_two_panel_min_n": "from this order on the split modes take two-panel steps",
    "lu_nb_mixed": "LU panel, the other mixed modes (and smaller n)",
    "lu_nb_fp64": "LU panel, FP64",
    "panel_leaf": "columns of a recursive panel's leaves",
    "trail_group_kib": "L2 share (KiB) of a row-tile group of the split GEMM",
    "refine_graph": "refinement sweeps replayed from one captured hipGraph",
    "zgemm_tile": "tile shape of the FP64 ZGEMM (0: 128x128 ... 4: 64x32)",
    "fp64_gemm_library": "1: rocBLAS for the FP64 LU updates and many-RHS residuals",
    "cbfm_block_mode": "CBF block solves: 0 FP64 LU, 1 c32 / 2 f16 / 4 bf16 LU + FP64 refinement, 5 by size",
    "cgemm_tile": "tile shape of the complex-FP32 GEMM without transposes (5: the 64x64 k_cgemm)",
    "cbfm_graph": "CBFM block batches (fills, excitations, FP64 LU, solve) launched as one hipGraph each",
    "lu_pivot_search": "mixed LU pivot search in each panel's diagonal block (certified): 1 MoM solves, 2 all; 0 whole columns",
    "mom_matrix_free": "MoM mixed solves (<= 2 directions): FP64 residuals from the cells, the FP64 matrix not stored",
}


def _solve_time(m, s, mode):
    """GPU time of one solve (factorisation + refinement, ms, from the
    library's stream events: the host's allocation of the matrix, which varies
    from call to call, stays out) and the solve report."""
    tth, tph = np.array([0.0]), np.array([0.0])
    rx = (np.array([0.0]), np.array([0.0]))
    _, _, _, _, info = m.compute_efields_mom(s, s.k0, (tth, tph), rx, mode=mode, want_E=False)
    if info["status"] != api.STATUS_CONVERGED:
        raise RuntimeError(f"solve did not converge while tuning: {info}")
    return info["t_factor_ms"] + info["t_refine_ms"], info


def sweep(m, name, values, s, mode, reps, log):
    """Time every value of one parameter, the values interleaved over `reps`
    rounds (so that slow drifts of the device hit them alike); keep the
    fastest (best of its rounds); return it."""
    best_ms = {v: np.inf for v in values}
    for _ in range(reps):
        for v in values:
            api.tune_set(name, v)
            t, info = _solve_time(m, s, mode)
            best_ms[v] = min(best_ms[v], t)
            log.append({"parameter": name, "value": v, "cells": s.n, "mode": api.MODE_NAMES[mode], "gpu_ms": t,
                        "factor_ms": info["t_factor_ms"], "refine_ms": info["t_refine_ms"]})
    for v in values:
        print(f"  {name} = {v:6d}  n = {3 * s.n:6d}  {api.MODE_NAMES[mode]:8s}  {best_ms[v]:9.1f} ms", flush=True)
    best = min(values, key=lambda v: best_ms[v])
    api.tune_set(name, best)
    return best, best_ms


def write_header(values, path=HEADER):
    lines = ["// Tuned defaults of the library's schedule parameters, written by",
             "// midas_amd/autotune.py from timings on one MI355X (its measurements are in",
             "// midas_amd/tuning.json).  Each can be changed at run time with",
             "// midas_tune_set (include/midas_gpu.h).",
             "#pragma once"]
    for k in api.TUNE_NAMES:
        lines.append(f"#define MIDAS_TUNED_{k.upper():<22s} {values[k]:<8d} // {COMMENTS[k]}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="smaller sizes, one repetition")
    ap.add_argument("--no-write", action="store_true", help="print only")
    ap.add_argument("--table", default=TABLE, help="where the timings go (default midas_amd/tuning.json)")
    ap.add_argument("--header", default=HEADER, help="where the chosen defaults go (default csrc/midas_tuned.h)")
    args = ap.parse_args()
    reps = 1 if args.quick else 2
    sizes = [4000, 8000] if args.quick else [4000, 8000, 12000, 20000]
    spheres = {n: g.gaussian_random_sphere(size_param=6.0 * (n / 20000) ** (1 / 3), target_cells=n, eps=1.5 + 0.01j,
                                           seed=2) for n in sizes}
    big = spheres[sizes[-1]]
    m = api.Midas(0)
    log = []
    start = {k: api.tune_get(k) for k in api.TUNE_NAMES}
    print("start", start, flush=True)
    f16 = api.SOLVE_F16_IR
    # the wider LU step applies from lu_two_panel_min_n on: open it for the sweep
    api.tune_set("lu_two_panel_min_n", 0)
    sweep(m, "lu_nb_split", [512, 768, 1024], big, f16, reps, log)
    sweep(m, "panel_leaf", [32, 64, 128], big, f16, reps, log)
    sweep(m, "trail_group_kib", [1024, 2048, 4096], big, f16, reps, log)
    sweep(m, "refine_graph", [0, 1], big, f16, reps, log)
    # the smallest size from which on the chosen wider step beats 512
    chosen = api.tune_get("lu_nb_split")
    min_n = 1 << 30
    if chosen > 512:
        for n in sizes:
            _, t = sweep(m, "lu_nb_split", [512, chosen], spheres[n], f16, reps, log)
            if t[chosen] < t[512]:
                min_n = min(min_n, 3 * n)
            else:
                min_n = 1 << 30
        api.tune_set("lu_nb_split", chosen)
    api.tune_set("lu_two_panel_min_n", min_n)
    mid = spheres[sizes[-2]]
    sweep(m, "lu_nb_fp64", [64, 128, 256], mid, api.SOLVE_FP64, reps, log)
    m.close()
    final = {k: api.tune_get(k) for k in api.TUNE_NAMES}
    print("chosen", final, flush=True)
    if not args.no_write:
        with open(args.table, "w") as f:
            json.dump({"device": "MI355X (gfx950)", "chosen": final, "start": start, "timings": log}, f, indent=1)
        write_header(final, args.header)
        print(f"wrote {args.table} and {args.header}", flush=True)


if __name__ == "__main__":
    main()
