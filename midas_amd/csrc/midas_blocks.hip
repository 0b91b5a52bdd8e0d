// Block tables of the CBFM (host code): the division of the cells into
// blocks, the bounding box and sphere of every block and the extension cells
// of every block.
//
//   midas_partition_sph  Division_blocks, SPH (Division_blocks.f90:95-132, 1252-1537)
//   midas_partition_csh  Division_blocks, CSH (Division_blocks.f90:133-180, 190-652)
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
#include <climits>
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

// ---------------------------------------------------------------------------
// Division into blocks (Division_blocks.f90).  Both divisions reorder the
// cells so that every block is a contiguous range, as the reference leaves
// Cells; `order` receives, for each new position, the input index of the
// cell there.  Cells keep their input order inside a block.
// ---------------------------------------------------------------------------
namespace {

constexpr int NB_BLOCK_S_MAX = 10000;   // NbBlock_s_max (Initialization.f90:54)

// Fortran NINT (halves away from zero) and CEILING to a default integer; an
// out-of-range quotient (a zero slab width) converts to -2^31 as on x86.
int f_nint(double v) { return v >= 0 ? (int)std::floor(v + 0.5) : -(int)std::floor(-v + 0.5); }
int f_ceiling(double v)
{
    if (!std::isfinite(v) || std::fabs(v) >= 2147483648.0) return INT_MIN;
    return (int)std::ceil(v);
}

struct Cloud {
    const double *p[3];
    int n;
};

// largest distance of cells[0..nc) from their box centre (SetBlockCtrs, :1631-1660)
double radius_of(const Cloud &P, const int *cells, int nc)
{
    double lo[3], hi[3], c[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = hi[a] = P.p[a][cells[0]];
        for (int i = 1; i < nc; ++i) {
            lo[a] = std::min(lo[a], P.p[a][cells[i]]);
            hi[a] = std::max(hi[a], P.p[a][cells[i]]);
        }
        c[a] = (lo[a] + hi[a]) / 2.0;
    }
    double r = 0.0;
    for (int i = 0; i < nc; ++i) {
        const double dx = P.p[0][cells[i]] - c[0], dy = P.p[1][cells[i]] - c[1], dz = P.p[2][cells[i]] - c[2];
        r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    return r;
}

// HierarchicalOctree_subdivision (Division_blocks.f90:1403-1537): cells
// in[0..nc) split at fr of their centre extent per axis into octants
// 1 + ix + 2 iy + 4 iz (a cell at or below the split goes low); written to
// out[0..nc) grouped by octant, input order kept; sizes of the non-empty
// octants in octant order.
void octree_split(const Cloud &P, const int *in, int nc, const double fr[3], int *out, std::vector<int> &sizes)
{
    double lo[3], h[3];
    for (int a = 0; a < 3; ++a) {
        double mn = P.p[a][in[0]], mx = mn;
        for (int i = 1; i < nc; ++i) {
            mn = std::min(mn, P.p[a][in[i]]);
            mx = std::max(mx, P.p[a][in[i]]);
        }
        lo[a] = mn;
        h[a] = (mx - mn) * fr[a];
    }
    std::vector<unsigned char> oct(nc);
    int cnt[8] = {0};
    for (int i = 0; i < nc; ++i) {
        int o = 0;
        for (int a = 0; a < 3; ++a)
            if (!(P.p[a][in[i]] - lo[a] <= h[a])) o |= 1 << a;
        oct[i] = (unsigned char)o;
        ++cnt[o];
    }
    int off[8], s = 0;
    for (int o = 0; o < 8; ++o) {
        off[o] = s;
        s += cnt[o];
    }
    for (int i = 0; i < nc; ++i) out[off[oct[i]]++] = in[i];
    sizes.clear();
    for (int o = 0; o < 8; ++o)
        if (cnt[o]) sizes.push_back(cnt[o]);
}

// split fractions of Division_blocks_sph_H (:1298-1312): 1/2 on every axis,
// then every combination of the single-precision 1/3. and 2/3.
void sph_fractions(double fr[9][3])
{
    const double t1 = (double)(1.0f / 3.0f), t2 = (double)(2.0f / 3.0f);
    fr[0][0] = fr[0][1] = fr[0][2] = 0.5;
    int row = 1;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k, ++row) {
                fr[row][0] = i ? t2 : t1;
                fr[row][1] = j ? t2 : t1;
                fr[row][2] = k ? t2 : t1;
            }
}

// distributeCells (Division_blocks.f90:1668-1709): slab ceiling((c - dmin) /
// dstep) clamped to [1, nbl]; out receives the cells grouped by slab (input
// order kept), cnt the slab sizes.
void distribute(const double *pos, const int *in, int nc, int nbl, double dmin, double dstep, int *out,
                std::vector<int> &cnt)
{
    cnt.assign(nbl, 0);
    std::vector<int> slab(nc);
    for (int i = 0; i < nc; ++i) {
        const int s = std::min(std::max(f_ceiling((pos[in[i]] - dmin) / dstep), 1), nbl) - 1;
        slab[i] = s;
        ++cnt[s];
    }
    std::vector<int> off(nbl);
    for (int b = 0, s = 0; b < nbl; ++b) {
        off[b] = s;
        s += cnt[b];
    }
    for (int i = 0; i < nc; ++i) out[off[slab[i]]++] = in[i];
}

// Division_blocks_csh (Division_blocks.f90:190-652) on the cells in `order`
// (reordered in place); block sizes in `sizes`.  Returns false where the
// reference sets error_division.
bool csh_once(const Cloud &P, std::vector<int> &order, double sc, int navg, double h_block, const double box_min[3],
              const double dims[3], int type_s, std::vector<int> &sizes)
{
    const int n = P.n;
    // directions by decreasing box side, the first index on ties (minloc of maxval)
    int dirs[3];
    double d[3] = {dims[0], dims[1], dims[2]};
    for (int j = 0; j < 3; ++j) {
        int k = 0;
        for (int a = 1; a < 3; ++a)
            if (d[a] > d[k]) k = a;
        dirs[j] = k;
        d[k] = 0.0;
    }
    const int dir0 = dirs[0], dir1 = dirs[1], dir2 = dirs[2];
    // slabs along the longest side
    const double hbox = dims[dir0];
    const int div = (int)(hbox / h_block + 1);
    const double h0 = hbox / div;
    std::vector<int> next(n), cnt;
    if (div > 1) {
        distribute(P.p[dir0], order.data(), n, div, box_min[dir0], f_nint(h0 / sc) * sc, next.data(), cnt);
        order.swap(next);
        sizes = cnt;
    } else {
        sizes.assign(1, n);
    }
    if (type_s == 3) {   // cylinder: slabs only, empty ones dropped
        sizes.erase(std::remove(sizes.begin(), sizes.end(), 0), sizes.end());
        return true;
    }
    const int nbmax = std::min(*std::min_element(sizes.begin(), sizes.end()), navg);
    bool possible = true;
    std::vector<int> nsizes, tmp(n), c2;
    while (possible) {
        possible = false;
        nsizes.clear();
        int start = 0;
        for (const int nbc : sizes) {
            const int *blk = order.data() + start;
            if (nbc > nbmax && std::abs(nbmax - nbc) > std::abs(nbmax - nbc / 2)) {
                possible = true;
                // split in two along the longer of the other two sides of its box
                double lo[3], hi[3];
                for (int a = 0; a < 3; ++a) {
                    lo[a] = hi[a] = P.p[a][blk[0]];
                    for (int i = 1; i < nbc; ++i) {
                        lo[a] = std::min(lo[a], P.p[a][blk[i]]);
                        hi[a] = std::max(hi[a], P.p[a][blk[i]]);
                    }
                }
                const int dd = (hi[dir1] - lo[dir1]) > (hi[dir2] - lo[dir2]) ? dir1 : dir2;
                double h = (hi[dd] - lo[dd]) / 2;
                const double dmin = lo[dd] - sc / 2.0;
                const double *pos = P.p[dd];
                distribute(pos, blk, nbc, 2, dmin, f_nint(h / sc) * sc, tmp.data(), c2);
                // move the cut by halving steps while the first piece's
                // deviation from nbc/2 does not grow (at most 20 trials); the
                // deviation is a single-precision quotient
                const int ref = nbc / 2;
                float err1 = (float)std::abs(c2[0] - ref) / (float)ref;
                double prev_err = 2.0 * (double)err1, prev_h = 0.0, step = h / 2.0;
                bool moved = false;
                for (int tr = 1; (double)err1 <= prev_err && tr <= 20; ++tr) {
                    prev_err = err1;
                    prev_h = h;
                    moved = true;
                    h = c2[0] > c2[1] ? h - step : h + step;
                    distribute(pos, blk, nbc, 2, dmin, f_nint(h / sc) * sc, tmp.data(), c2);
                    err1 = (float)std::abs(c2[0] - ref) / (float)ref;
                    step = step / 2.0;
                }
                if (!moved) return false;   // a one-cell block (nbmax = 0): no valid cut
                distribute(pos, blk, nbc, 2, dmin, f_nint(prev_h / sc) * sc, next.data() + start, c2);
                nsizes.push_back(c2[0]);
                nsizes.push_back(c2[1]);
            } else {
                std::copy(blk, blk + nbc, next.data() + start);
                nsizes.push_back(nbc);
            }
            if ((int)nsizes.size() > NB_BLOCK_S_MAX) return false;
            start += nbc;
        }
        order.swap(next);
        sizes.swap(nsizes);
    }
    return true;
}

int partition_out(const Cloud &P, const std::vector<int> &order, const std::vector<int> &sizes, int *out_order,
                  int *block_size, int cap, int *nblocks, double *hblock)
{
    *nblocks = (int)sizes.size();
    MIDAS_CHECK((int)sizes.size() <= cap, -1, "partition: %zu blocks exceed the capacity %d", sizes.size(), cap);
    std::copy(order.begin(), order.end(), out_order);
    std::copy(sizes.begin(), sizes.end(), block_size);
    double r = 0.0;
    for (size_t b = 0, s = 0; b < sizes.size(); s += sizes[b], ++b) r = std::max(r, radius_of(P, &order[s], sizes[b]));
    if (hblock) *hblock = 2.0 * r;
    return 0;
}

int check_partition_args(const midas_cells_t *cells, int navg, const int *order, const int *bsz, const int *nb)
{
    MIDAS_CHECK(cells && cells->x && cells->y && cells->z && cells->n > 0, -1, "partition: bad cells");
    MIDAS_CHECK(navg > 0 && order && bsz && nb, -1, "partition: bad arguments");
    return 0;
}

}  // namespace

extern "C" int midas_partition_sph(const midas_cells_t *cells, int navg_cells, int best_combin, int *order,
                                   int *block_size, int cap, int *nblocks, double *hblock)
{
    TRY(check_partition_args(cells, navg_cells, order, block_size, nblocks));
    const Cloud P = {{cells->x, cells->y, cells->z}, cells->n};
    const int n = cells->n;
    double fr[9][3];
    sph_fractions(fr);
    std::vector<int> ord(n), work(n), sizes, sub;
    for (int i = 0; i < n; ++i) work[i] = i;
    octree_split(P, work.data(), n, fr[0], ord.data(), sizes);
    // scan the blocks; a block worth an eighth split is replaced by its pieces
    // and the scan resumes at the first piece (Division_blocks.f90:1322-1378)
    size_t ii = 0;
    int start = 0;
    while (ii < sizes.size()) {
        const int N = sizes[ii];
        if (N > navg_cells && std::abs(N - navg_cells) > std::abs(f_nint((double)((float)N / 8.0f)) - navg_cells)) {
            int pick = 0;
            if (best_combin) {
                // the fractions whose pieces deviate least from their mean size
                // (first on ties; 1/2 when no trial splits the block)
                int best = -1;
                for (int t = 0; t < 9; ++t) {
                    octree_split(P, ord.data() + start, N, fr[t], work.data(), sub);
                    if (sub.size() > 1) {
                        const int mean = N / (int)sub.size();
                        int dev = 0;
                        for (const int s : sub) dev = std::max(dev, std::abs(s - mean));
                        if (best < 0 || dev < best) {
                            best = dev;
                            pick = t;
                        }
                    }
                }
            }
            octree_split(P, ord.data() + start, N, fr[pick], work.data(), sub);
            MIDAS_CHECK(sub.size() > 1, -2, "partition_sph: block %zu of %d coincident cells cannot be split", ii, N);
            std::copy(work.data(), work.data() + N, ord.data() + start);
            sizes.erase(sizes.begin() + ii);
            sizes.insert(sizes.begin() + ii, sub.begin(), sub.end());
        } else {
            start += N;
            ++ii;
        }
    }
    return partition_out(P, ord, sizes, order, block_size, cap, nblocks, hblock);
}

extern "C" int midas_partition_csh(const midas_cells_t *cells, double sc, int navg_cells, int type_s, int *order,
                                   int *block_size, int cap, int *nblocks, double *hblock)
{
    TRY(check_partition_args(cells, navg_cells, order, block_size, nblocks));
    MIDAS_CHECK(sc > 0, -1, "partition_csh: sc must be positive");
    const Cloud P = {{cells->x, cells->y, cells->z}, cells->n};
    const int n = cells->n;
    // the scatterer's box: centre extremes -+ sc/2 (Discretization.f90:510-519)
    double box_min[3], dims[3];
    for (int a = 0; a < 3; ++a) {
        const double *p = P.p[a];
        const double mn = *std::min_element(p, p + n), mx = *std::max_element(p, p + n);
        box_min[a] = mn - sc / 2.0;
        dims[a] = (mx + sc / 2.0) - box_min[a];
    }
    const double dmax = std::max(dims[0], std::max(dims[1], dims[2]));
    // trial divisions with hBlock halved each time (each trial reorders the
    // cells the next one starts from), while the mean block size exceeds
    // navg (Division_blocks.f90:141-168)
    std::vector<int> ord(n), sizes;
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::vector<int> diff(9, n);
    std::vector<double> htest(9, 0.0);
    double h_block = dmax;
    for (int i = 0, N = n; i < 8 && N > navg_cells; ++i) {
        if (!csh_once(P, ord, sc, navg_cells, h_block, box_min, dims, type_s, sizes)) {
            MIDAS_CHECK(i > 0, -2, "partition_csh: the division into blocks failed");
            break;
        }
        htest[i] = h_block;
        N = n / (int)sizes.size();
        diff[i] = std::abs(N - navg_cells);
        h_block = h_block * 0.5;
    }
    // at most 2 navg cells: hBlock = the largest side; else the trial closest to navg
    h_block = n <= 2 * navg_cells ? dmax : htest[std::min_element(diff.begin(), diff.end()) - diff.begin()];
    MIDAS_CHECK(csh_once(P, ord, sc, navg_cells, h_block, box_min, dims, type_s, sizes), -2,
                "partition_csh: the division into blocks failed");
    return partition_out(P, ord, sizes, order, block_size, cap, nblocks, hblock);
}
