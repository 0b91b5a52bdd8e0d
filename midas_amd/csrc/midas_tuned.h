// Tuned defaults of the library's schedule parameters, written by
// midas_amd/autotune.py from timings on one MI355X (its measurements are in
// midas_amd/tuning.json).  Each can be changed at run time with
// midas_tune_set (include/midas_gpu.h).
#pragma once
#define MIDAS_TUNED_LU_NB_SPLIT            1024     // LU outer step, single-term split modes, n >= lu_two_panel_min_n
#define MIDAS_TUNED_LU_TWO_PANEL_MIN_N     12000    // from this order on the split modes take two-panel steps
#define MIDAS_TUNED_LU_NB_MIXED            512      // LU panel, the other mixed modes (and smaller n)
#define MIDAS_TUNED_LU_NB_FP64             256      // LU panel, FP64
#define MIDAS_TUNED_PANEL_LEAF             128      // columns of a recursive panel's leaves
#define MIDAS_TUNED_TRAIL_GROUP_KIB        2048     // L2 share (KiB) of a row-tile group of the split GEMM
#define MIDAS_TUNED_REFINE_GRAPH           0        // refinement sweeps replayed from one captured hipGraph
#define MIDAS_TUNED_ZGEMM_TILE             2        // tile shape of the FP64 ZGEMM (0: 128x128 ... 4: 64x32)
#define MIDAS_TUNED_FP64_GEMM_LIBRARY      0        // 1: rocBLAS for the FP64 LU updates and many-RHS residuals
#define MIDAS_TUNED_CBFM_BLOCK_MODE        5        // CBF block solves: 0 FP64 LU, 1 c32 / 2 f16 / 4 bf16 LU + FP64 refinement, 5 by size
#define MIDAS_TUNED_CGEMM_TILE             2        // tile shape of the complex-FP32 GEMM without transposes (5: the 64x64 k_cgemm)
#define MIDAS_TUNED_CBFM_GRAPH             1        // CBFM block batches (fills, excitations, FP64 LU, solve) launched as one hipGraph each
#define MIDAS_TUNED_LU_PIVOT_SEARCH        1        // mixed LU pivot search in each panel's diagonal block (certified): 1 MoM solves, 2 all; 0 whole columns
#define MIDAS_TUNED_MOM_MATRIX_FREE        1        // MoM mixed solves (<= 2 directions): FP64 residuals from the cells, the FP64 matrix not stored
#define MIDAS_TUNED_U12_SPLIT              1        // single-term split LUs: U12 = L11^-1 A12 on the split-precision matrix cores
#define MIDAS_TUNED_L21_SPLIT              1        // ... and L21 = A21 U11^-1 of the diagonal-block pivot search (two-panel steps)
#define MIDAS_TUNED_LU_LOOKAHEAD           1        // LU panels on their own stream, one step ahead (0: on the main stream)
