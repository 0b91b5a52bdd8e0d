// Run-time schedule parameters (midas_tune.hip).
#pragma once
#include "midas_common.h"

namespace midas {
enum {
    TUNE_LU_NB_SPLIT,
    TUNE_LU_TWO_PANEL_MIN_N,
    TUNE_LU_NB_MIXED,
    TUNE_LU_NB_FP64,
    TUNE_PANEL_LEAF,
    TUNE_TRAIL_GROUP_KIB,
    TUNE_REFINE_GRAPH,
    TUNE_ZGEMM_TILE,
    TUNE_FP64_GEMM_LIBRARY,
    TUNE_CBFM_BLOCK_MODE,
    TUNE_CGEMM_TILE,
    TUNE_CBFM_GRAPH,
    TUNE_LU_PIVOT_SEARCH,
    TUNE_MOM_MATRIX_FREE,
    TUNE_U12_SPLIT,
    TUNE_L21_SPLIT,
    TUNE_LU_LOOKAHEAD,
    TUNE_COUNT
};
int tune(int which);
}  // namespace midas
