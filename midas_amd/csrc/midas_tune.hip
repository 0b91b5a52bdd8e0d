// Run-time schedule parameters of the library (defaults in midas_tuned.h, set
// by midas_amd/autotune.py), readable and settable through the C-ABI so that
// the autotuner can time alternatives.
#include "midas_tune.h"
#include "midas_tuned.h"
#include "../../include/midas_gpu.h"

#include <cstring>

namespace midas {

struct TuneEntry {
    const char *name;
    int value, lo, hi;
};
static TuneEntry g_tune[TUNE_COUNT] = {
    {"lu_nb_split", MIDAS_TUNED_LU_NB_SPLIT, 16, 1024},
    {"lu_two_panel_min_n", MIDAS_TUNED_LU_TWO_PANEL_MIN_N, 0, 1 << 30},
    {"lu_nb_mixed", MIDAS_TUNED_LU_NB_MIXED, 16, 512},
    {"lu_nb_fp64", MIDAS_TUNED_LU_NB_FP64, 16, 512},
    {"panel_leaf", MIDAS_TUNED_PANEL_LEAF, 16, 512},
    {"trail_group_kib", MIDAS_TUNED_TRAIL_GROUP_KIB, 64, 65536},
    {"refine_graph", MIDAS_TUNED_REFINE_GRAPH, 0, 1},
    {"zgemm_tile", MIDAS_TUNED_ZGEMM_TILE, 0, 4},
    {"fp64_gemm_library", MIDAS_TUNED_FP64_GEMM_LIBRARY, 0, 1},
    {"cbfm_block_mode", MIDAS_TUNED_CBFM_BLOCK_MODE, 0, 5},
    {"cgemm_tile", MIDAS_TUNED_CGEMM_TILE, 0, 5},
    {"cbfm_graph", MIDAS_TUNED_CBFM_GRAPH, 0, 1},
    {"lu_pivot_search", MIDAS_TUNED_LU_PIVOT_SEARCH, 0, 2},
    {"mom_matrix_free", MIDAS_TUNED_MOM_MATRIX_FREE, 0, 1},
    {"u12_split", MIDAS_TUNED_U12_SPLIT, 0, 1},
    {"l21_split", MIDAS_TUNED_L21_SPLIT, 0, 1},
    {"lu_lookahead", MIDAS_TUNED_LU_LOOKAHEAD, 0, 1},
};

int tune(int which) { return g_tune[which].value; }

}  // namespace midas

using namespace midas;

extern "C" int midas_tune_set(const char *name, int value)
{
    MIDAS_CHECK(name, -1, "tune_set: no name");
    for (TuneEntry &e : g_tune)
        if (!std::strcmp(e.name, name)) {
            MIDAS_CHECK(value >= e.lo && value <= e.hi, -1, "tune_set: %s = %d outside [%d, %d]", name, value, e.lo,
                        e.hi);
            e.value = value;
            return 0;
        }
    MIDAS_CHECK(false, -1, "tune_set: unknown parameter %s", name);
}

extern "C" int midas_tune_get(const char *name, int *value)
{
    MIDAS_CHECK(name && value, -1, "tune_get: bad arguments");
    for (const TuneEntry &e : g_tune)
        if (!std::strcmp(e.name, name)) {
            *value = e.value;
            return 0;
        }
    MIDAS_CHECK(false, -1, "tune_get: unknown parameter %s", name);
}
