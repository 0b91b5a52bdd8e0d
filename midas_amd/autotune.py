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
  lu_pivot_search     the MoM solve's pivot search: whole columns or each
                      panel's diagonal block (certified)
  mom_matrix_free     FP64 residuals of the MoM solve from the cells instead
                      of a stored FP64 matrix (timed over the whole call:
                      fill, solve, far field)
  u12_split           U12 = L11^-1 A12 of the split-mode LU on the split
                      GEMM (in place, A12 -= (I - L11^-1) A12) or on the
                      complex-FP32 GEMM
  l21_split           the same for L21 = A21 U11^-1 of the diagonal-block
                      pivot search (A21 -= A21 (I - U11^-1))

and, on a DLCA aggregate solved by CBFM-E with ACA (midas_compute_efields_cbfm,
wall time of the call):

  cbfm_block_mode     the CBF block systems by block size (5: the f16 LU +
                      FP64 refinement for blocks at least 16 times taller than
                      the excitation count, the FP64 LU below) or always by
                      the f16 LU + refinement (2).  (0, the FP64 LU for all, is
                      the same code as 5 on this case, whose blocks are below
                      that size; C5 measured 5 against 0: 1.85 s against
                      1.98 s for its block phase.)

(cbfm_graph, the batches' hipGraph launch, is not swept: it is the design's
launch form, and measures equal to eager issue on the C3 and a0006 benches,
the phase being GPU-bound.)

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
    "u12_split": "single-term split LUs: U12 = L11^-1 A12 on the split-precision matrix cores",
    "l21_split": "... and L21 = A21 U11^-1 of the diagonal-block pivot search (two-panel steps)",
    "lu_lookahead": "LU panels on their own stream, one step ahead (0: on the main stream)",
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
    fastest (best of its rounds) unless the current value is within 1 % of
    it; return it."""
    best_ms = {v: np.inf for v in values}
    start = api.tune_get(name)
    for _ in range(reps):
        for v in values:
            api.tune_set(name, v)
            t, info = _solve_time(m, s, mode)
            best_ms[v] = min(best_ms[v], t)
            log.append({"parameter": name, "value": v, "cells": s.n, "mode": api.MODE_NAMES[mode], "gpu_ms": t,
                        "factor_ms": info["t_factor_ms"], "refine_ms": info["t_refine_ms"]})
    for v in values:
        print(f"  {name} = {v:6d}  n = {3 * s.n:6d}  {api.MODE_NAMES[mode]:8s}  {best_ms[v]:9.1f} ms", flush=True)
    best = _pick(values, best_ms, start)
    api.tune_set(name, best)
    return best, best_ms


def _pick(values, best_ms, current):
    """The fastest value, unless the current one is within 1 % of it (timing
    noise does not move a default)."""
    best = min(values, key=lambda v: best_ms[v])
    if current in best_ms and best_ms[current] <= 1.01 * best_ms[best]:
        return current
    return best


def _wall(f, reps):
    """Best wall time (ms) of reps calls of f."""
    best = np.inf
    for _ in range(reps):
        t0 = time.perf_counter()
        f()
        best = min(best, 1e3 * (time.perf_counter() - t0))
    return best


def sweep_wall(name, values, f, reps, log, tag, keep=True):
    """As sweep, for a timing function f() of the whole call (ms of wall
    time); with keep=False the parameter is restored after the timings."""
    old = api.tune_get(name)
    best_ms = {v: np.inf for v in values}
    for _ in range(reps):
        for v in values:
            api.tune_set(name, v)
            t = _wall(f, 1)
            best_ms[v] = min(best_ms[v], t)
            log.append({"parameter": name, "value": v, "case": tag, "wall_ms": t})
    for v in values:
        print(f"  {name} = {v:6d}  {tag:24s}  {best_ms[v]:9.1f} ms", flush=True)
    best = _pick(values, best_ms, old) if keep else old
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
    # the CBFM first: a DLCA aggregate, SR blocks, ACA, the reference's uniform
    # CBF grid (timed before the MoM solves: once the process has allocated and
    # freed their 57.5 GB matrices, the runtime's later multi-GB allocations --
    # the CBFM's, made per call -- slow down)
    tth, tph = np.array([0.0]), np.array([0.0])
    ag = g.dlca_aggregate(nmono=16 if args.quick else 64, cells_per_mono=300, seed=3)
    th_c, ph_c = g.cbf_directions_uniform(422)
    cbfm_call = lambda: m.compute_efields_cbfme(ag, ag.k0, ag.blocks, ag.ext, (th_c, ph_c), (tth, tph), use_aca=True,
                                                sr=True, res_sr=0.01, mode=f16)
    cbfm_call()
    tag = f"CBFM DLCA {ag.n} cells"
    sweep_wall("cbfm_block_mode", [5, 2], cbfm_call, reps + 1, log, tag)
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
    # the MoM solve's pivot search and residual form, on the largest sphere
    sweep(m, "lu_pivot_search", [0, 1], big, f16, reps, log)
    sweep(m, "u12_split", [0, 1], big, f16, reps, log)
    sweep(m, "l21_split", [0, 1], big, f16, reps, log)
    rx = (np.array([0.0, 90.0]), np.array([0.0, 0.0]))
    mom_call = lambda: m.compute_efields_mom(big, big.k0, (tth, tph), rx, mode=f16, want_E=False)
    mom_call()
    sweep_wall("mom_matrix_free", [0, 1], mom_call, reps, log, f"MoM {big.n} cells")
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
