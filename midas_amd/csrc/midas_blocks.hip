// Block tables of the CBFM (host code): bounding box and sphere of every block
// and the extension cells of every block.
//
//   midas_block_frames   SetBlockCtrs (Division_blocks.f90:1616-1660)
//   midas_extend_blocks  Extend_blocks (Extend_blocks.f90:1-162)
//
// Blocks are contiguous cell ranges block_start[b] .. block_start[b+1]-1, in
// block order (as Division_blocks leaves the cells).  These tables are O(N)
// integer and comparison work done once per geometry, as in the reference, on
// the host; the arrays they produce are the midas_blocks_t input of the GPU
// CBFM solver.
#include "midas_common.h"
#include "midas_host.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int FACT_NBEXT_MAX = 10;   // Fact_Nbext_max (Initialization.f90:55)

// frames of block b: box min (3), box max (3), box extent (3), sphere centre (3), radius
void frame_of(const midas_cells_t *c, int b0, int b1, double *f)
{
    const double *p[3] = {c->x, c->y, c->z};
    for (int d = 0; d < 3; ++d) {
        double lo = p[d][b0], hi = p[d][b0];
        for (int i = b0 + 1; i < b1; ++i) {
            lo = std::min(lo, p[d][i]);
            hi = std::max(hi, p[d][i]);
        }
        f[d] = lo;
        f[3 + d] = hi;
        f[6 + d] = hi - lo;
        f[9 + d] = (lo + hi) / 2.0;
    }
    double r = 0.0;
    for (int i = b0; i < b1; ++i) {
        const double dx = c->x[i] - f[9], dy = c->y[i] - f[10], dz = c->z[i] - f[11];
        r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    f[12] = r;
}

int check_blocks(const midas_cells_t *cells, int nblocks, const int *bs)
{
    MIDAS_CHECK(cells && cells->x && cells->y && cells->z && bs && nblocks > 0, -1, "blocks: bad arguments");
    MIDAS_CHECK(bs[0] == 0 && bs[nblocks] == cells->n, -1, "blocks: block_start must run 0..n");
    for (int b = 0; b < nblocks; ++b) MIDAS_CHECK(bs[b + 1] > bs[b], -1, "blocks: block %d empty", b);
    return 0;
}

}  // namespace

extern "C" int midas_block_frames(const midas_cells_t *cells, int nblocks, const int *block_start, double *frames)
{
    TRY(check_blocks(cells, nblocks, block_start));
    MIDAS_CHECK(frames, -1, "block_frames: no output");
    for (int b = 0; b < nblocks; ++b) frame_of(cells, block_start[b], block_start[b + 1], frames + 13L * b);
    return 0;
}

extern "C" int midas_extend_blocks(const midas_cells_t *cells, int nblocks, const int *block_start, double sc,
                                   int nc_extended, int *ext_start, int *ext_cells, long ext_cap, int *adj_start,
                                   int *adj, long adj_cap)
{
    TRY(check_blocks(cells, nblocks, block_start));
    MIDAS_CHECK(ext_start && sc > 0 && nc_extended >= 0, -1, "extend_blocks: bad arguments");
    const int *bs = block_start;
    std::vector<double> fr(13L * nblocks);
    for (int b = 0; b < nblocks; ++b) frame_of(cells, bs[b], bs[b + 1], fr.data() + 13L * b);
    int nbc_max = 0;
    for (int b = 0; b < nblocks; ++b) nbc_max = std::max(nbc_max, bs[b + 1] - bs[b]);
    const long nbext_max = (long)FACT_NBEXT_MAX * nbc_max;

    std::vector<int> ext, adjs;
    std::vector<int> es(nblocks + 1, 0), as(nblocks + 1, 0);
    // one block, or no extension layers: no extension (Extend_blocks.f90:145-147)
    const bool extend = nblocks > 1 && nc_extended != 0;
    const double ext_len = nc_extended * sc;
    for (int B = 0; B < nblocks && extend; ++B) {
        const double *f1 = fr.data() + 13L * B;
        // adjacent blocks: bounding spheres closer than the sum of the radii and
        // the extension length; scanned in block order, B itself skipped
        std::vector<int> badj;
        for (int Bp = 0; Bp < nblocks; ++Bp) {
            if (Bp == B) continue;
            const double *f2 = fr.data() + 13L * Bp;
            const double dx = f2[9] - f1[9], dy = f2[10] - f1[10], dz = f2[11] - f1[11];
            const double dis = std::sqrt(dx * dx + dy * dy + dz * dz);
            const double thr = f1[12] + f2[12] + ext_len;
            if (dis <= thr) badj.push_back(Bp);
        }
        // extension zone: the block's box grown by the extension length and half a cell
        double lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = f1[d] - ext_len - sc / 2.0;
            hi[d] = f1[3 + d] + ext_len + sc / 2.0;
        }
        long nb_ext = 0;
        for (const int Bp : badj) {
            long in_bp = 0;
            for (int i = bs[Bp]; i < bs[Bp + 1]; ++i) {
                const double x = cells->x[i], y = cells->y[i], z = cells->z[i];
                if (lo[0] <= x && x <= hi[0] && lo[1] <= y && y <= hi[1] && lo[2] <= z && z <= hi[2]) {
                    ++nb_ext;
                    MIDAS_CHECK(nb_ext <= nbext_max, -2,
                                "extend_blocks: extension of block %d exceeds %ld cells (10 x the largest block)", B,
                                nbext_max);
                    ++in_bp;
                    ext.push_back(i);
                }
            }
            // adjacent only if it gave cells (Extend_blocks.f90:130-134)
            if (in_bp) adjs.push_back(Bp);
        }
        es[B + 1] = (int)ext.size();
        as[B + 1] = (int)adjs.size();
    }
    if (!extend)
        for (int B = 0; B < nblocks; ++B) es[B + 1] = as[B + 1] = 0;
    std::copy(es.begin(), es.end(), ext_start);
    if (ext_cells) {
        MIDAS_CHECK((long)ext.size() <= ext_cap, -1, "extend_blocks: ext_cap %ld < %zu extension cells", ext_cap,
                    ext.size());
        std::copy(ext.begin(), ext.end(), ext_cells);
    }
    if (adj_start) std::copy(as.begin(), as.end(), adj_start);
    if (adj) {
        MIDAS_CHECK((long)adjs.size() <= adj_cap, -1, "extend_blocks: adj_cap %ld < %zu", adj_cap, adjs.size());
        std::copy(adjs.begin(), adjs.end(), adj);
    }
    return 0;
}
