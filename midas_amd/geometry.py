"""Seeded synthetic scatterers, voxelised into MIDAS cell lists.

The reference reads its geometry from input decks (Discretization.f90); the
hot path only needs the resulting cell table (positions, cell size Sc, relative
permittivity per cell) and, for CBFM, a block partition.  These generators
build such tables for the workloads this package is measured on:

  sphere(...)             homogeneous sphere on the reference's 7-domain lattice
                          rule (Discretization.f90:246-310)
  gaussian_random_sphere  lognormal-radius Gaussian random sphere
  dlca_aggregate          cluster-cluster aggregate of spherical monomers, one
                          CBFM block per monomer
  convex_hull_particle    hull of random points on an ellipsoid
  porous_particle         thresholded Gaussian random field with categorical
                          permittivities, split into uneven blocks

Lengths are in metres, angles in degrees; `k0` is returned with the geometry so
that the requested size parameter holds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Scatterer:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    sc: np.ndarray
    eps: np.ndarray
    k0: float
    blocks: list = field(default_factory=list)   # 0-based cell indices per block (contiguous)
    ext: list = field(default_factory=list)      # 0-based extension cells per block
    name: str = ""

    @property
    def n(self):
        return len(self.x)


def sphere(size_param=2.0, nd=16, eps=1.6 + 0.02j, radius=1e-3):
    """Reference lattice rule: inner cube of whole cells, then six shells
    keeping centres with Rc < a (Discretization.f90:246-310)."""
    dm = 2.0 * radius
    sc = dm / nd
    ap = dm / 2.0
    n_dp = int(round(dm / sc))
    n_ext = int(math.ceil((ap - (0.8 * ap / math.sqrt(3.0))) / sc))
    n_int = n_dp - 2 * n_ext
    ap_int = n_int * sc / 2.0
    sx = [-ap_int, -ap, -ap, ap_int, -ap_int, -ap_int, -ap]
    nx = [n_int, n_dp, n_ext, n_ext, n_int, n_int, n_dp]
    sy = [-ap_int, -ap, -ap, -ap, -ap, ap_int, -ap]
    ny = [n_int, n_dp, n_dp, n_dp, n_ext, n_ext, n_dp]
    sz = [-ap_int, -ap, -ap_int, -ap_int, -ap_int, -ap_int, ap_int]
    nz = [n_int, n_ext, n_int, n_int, n_int, n_int, n_ext]
    pts = []
    for d in range(7):
        iz, iy, ix = np.meshgrid(np.arange(nz[d]), np.arange(ny[d]), np.arange(nx[d]), indexing="ij")
        px = sx[d] + (ix.ravel() + 0.5) * sc
        py = sy[d] + (iy.ravel() + 0.5) * sc
        pz = sz[d] + (iz.ravel() + 0.5) * sc
        if d > 0:
            keep = np.sqrt(px ** 2 + py ** 2 + pz ** 2) < ap
            px, py, pz = px[keep], py[keep], pz[keep]
        pts.append(np.stack([px, py, pz], 1))
    p = np.concatenate(pts)
    n = len(p)
    k0 = size_param / radius
    return Scatterer(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), np.full(n, sc), np.full(n, eps, np.complex128),
                     k0, name=f"sphere x={size_param} Nd={nd}")


def _fib_dirs(n):
    """n quasi-uniform unit vectors (spherical Fibonacci lattice)."""
    i = np.arange(n) + 0.5
    zc = 1.0 - 2.0 * i / n
    r = np.sqrt(1.0 - zc ** 2)
    ph = math.pi * (1.0 + math.sqrt(5.0)) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), zc], 1)


def directions(n):
    """n quasi-uniform incidence directions as (theta, phi) in degrees with theta
    measured from +x, the reference's convention (Incident_Field.f90:31-34)."""
    u = _fib_dirs(n)
    th = np.degrees(np.arccos(np.clip(u[:, 0], -1, 1)))
    ph = np.degrees(np.arctan2(u[:, 2], u[:, 1]))
    return th, ph


def cbf_directions_uniform(niws):
    """CBF excitation directions of getTransmitters_CBFM with distr_iws = 1
    (getParameters_CBFM.f90:1-110): the grid uniform in cos(theta) and phi whose
    size is closest to `niws`, phi outer, theta inner, degrees, theta from +x.
    The grid-size expression is single precision in the reference (integer
    and default-real operands) and pi is the single-precision acos(-1.)."""
    f32 = np.float32
    delta = float(f32(f32(540.0) ** f32(2.0)) + f32(4 * (niws - 1) * (360 * 180)))
    step = (540 + math.sqrt(delta)) / (2 * (niws - 1))
    nth = int(np.rint(180.0 / step)) + 1
    nph = int(np.rint(360.0 / step)) + 1
    pq = float(f32(math.pi))
    c0 = math.cos(0.0)
    step_th = (math.cos(180.0 * pq / 180.0) - c0) / (nth - 1)
    th = np.degrees(np.arccos(c0 + step_th * np.arange(nth))) * (math.pi / pq)
    return np.tile(th, nph), np.repeat(np.arange(nph) * (360.0 / (nph - 1)), nth)


def _lattice(extent, sc):
    g = np.arange(-extent, extent + 1e-12, sc)
    g = g[:-1] + sc / 2 if len(g) > 1 else g
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1)


def _grs_radius(u, rng, sigma=0.2, corr_deg=30.0, lmax=24):
    """Gaussian random sphere radius r(u)/a for unit vectors u (Muinonen's
    lognormal model with a Gaussian correlation function)."""
    from scipy.special import eval_legendre, sph_harm_y
    beta2 = math.log(1.0 + sigma ** 2)
    ell = 2.0 * math.sin(math.radians(corr_deg) / 2.0)
    # Legendre coefficients of C(g) = exp(-2 sin^2(g/2) / ell^2)
    xg, wg = np.polynomial.legendre.leggauss(200)
    cg = np.exp(-(1.0 - xg) / ell ** 2)
    cl = np.array([(2 * l + 1) / 2.0 * np.sum(wg * cg * eval_legendre(l, xg)) for l in range(lmax + 1)])
    cl = np.maximum(cl, 0.0)
    cl[:2] = 0.0
    cl = cl / cl.sum()
    th = np.arccos(np.clip(u[:, 2], -1, 1))
    ph = np.arctan2(u[:, 1], u[:, 0])
    s = np.zeros(len(u))
    for l in range(2, lmax + 1):
        var = beta2 * cl[l] * 4 * math.pi / (2 * l + 1)
        for m in range(0, l + 1):
            ylm = sph_harm_y(l, m, th, ph)
            if m == 0:
                s += rng.normal(0, math.sqrt(var)) * ylm.real
            else:
                a, b = rng.normal(0, math.sqrt(var / 2), 2)
                s += 2 * (a * ylm.real - b * ylm.imag)
    return np.exp(s - beta2 / 2.0)


def gaussian_random_sphere(size_param=6.0, target_cells=20000, eps=1.5 + 0.01j, seed=2, radius=1e-3):
    from scipy.interpolate import RegularGridInterpolator
    k0 = size_param / radius
    # r(theta, phi) tabulated once on a 1-degree grid, then interpolated
    tg = np.radians(np.linspace(0.0, 180.0, 181))
    pg = np.radians(np.linspace(-180.0, 180.0, 361))
    T, P = np.meshgrid(tg, pg, indexing="ij")
    u = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    rtab = _grs_radius(u, np.random.default_rng(seed)).reshape(T.shape)
    interp = RegularGridInterpolator((tg, pg), rtab)
    rmax = float(rtab.max()) * radius
    sc = radius * (4.0 / 3.0 * math.pi / target_cells) ** (1.0 / 3.0)
    for _ in range(12):
        p = _lattice(rmax * 1.01, sc)
        r = np.linalg.norm(p, axis=1)
        p, r = p[r < rmax], r[r < rmax]
        th = np.arccos(np.clip(p[:, 2] / np.maximum(r, 1e-30), -1, 1))
        ph = np.arctan2(p[:, 1], p[:, 0])
        keep = r < interp(np.stack([th, ph], 1)) * radius
        n = int(keep.sum())
        if abs(n - target_cells) <= 0.01 * target_cells:
            break
        sc *= (n / target_cells) ** (1.0 / 3.0)
    p = p[keep]
    n = len(p)
    return Scatterer(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), np.full(n, sc), np.full(n, eps, np.complex128),
                     k0, name=f"gaussian random sphere x={size_param} N={n}")


def _tcca(nmono, df, kf, rng):
    """Monomer centres (unit monomer radius) of a tunable cluster-cluster
    aggregate: clusters are merged pairwise at the centre distance that keeps
    N = kf (Rg/a)^Df, in a random direction free of overlaps."""
    clusters = [np.zeros((1, 3)) for _ in range(nmono)]
    while len(clusters) > 1:
        rng.shuffle(clusters)
        merged = []
        for i in range(0, len(clusters) - 1, 2):
            a, b = clusters[i], clusters[i + 1]
            n1, n2 = len(a), len(b)
            n = n1 + n2
            rg1 = math.sqrt(np.mean(np.sum((a - a.mean(0)) ** 2, 1)) + 0.6)
            rg2 = math.sqrt(np.mean(np.sum((b - b.mean(0)) ** 2, 1)) + 0.6)
            g2 = n * n / (n1 * n2) * (n / kf) ** (2.0 / df) - n * rg1 ** 2 / n2 - n * rg2 ** 2 / n1
            gam = math.sqrt(max(g2, 4.0))
            a0, b0 = a - a.mean(0), b - b.mean(0)
            for attempt in range(4000):
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                g = gam * (1.0 + 0.002 * attempt)
                bb = b0 + g * d
                dist = np.linalg.norm(a0[:, None, :] - bb[None, :, :], axis=2)
                if dist.min() >= 2.0 and dist.min() <= 2.0 + 0.6 + 0.002 * attempt * g:
                    break
            merged.append(np.concatenate([a0, bb]))
        if len(clusters) % 2:
            merged.append(clusters[-1])
        clusters = merged
    c = clusters[0]
    return c - c.mean(0)


def dlca_aggregate(nmono=64, df=1.8, kf=1.3, mono_size_param=0.8, cells_per_mono=300,
                   eps_choices=(1.5 + 0.01j, 1.8 + 0.6j), seed=3, mono_radius=1e-3):
    rng = np.random.default_rng(seed)
    centres = _tcca(nmono, df, kf, rng) * mono_radius
    k0 = mono_size_param / mono_radius
    sc = mono_radius * (4.0 / 3.0 * math.pi / cells_per_mono) ** (1.0 / 3.0)
    lo = centres.min(0) - mono_radius
    hi = centres.max(0) + mono_radius
    axes = [np.arange(lo[d] + sc / 2, hi[d], sc) for d in range(3)]
    zz, yy, xx = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    p = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1)
    best = np.full(len(p), np.inf)
    owner = np.full(len(p), -1)
    for m, c in enumerate(centres):
        d = np.linalg.norm(p - c, axis=1)
        sel = (d < mono_radius) & (d < best)
        best[sel] = d[sel]
        owner[sel] = m
    keep = owner >= 0
    p, owner = p[keep], owner[keep]
    order = np.argsort(owner, kind="stable")
    p, owner = p[order], owner[order]
    mat = rng.integers(0, len(eps_choices), nmono)
    eps = np.array([eps_choices[mat[o]] for o in owner], np.complex128)
    blocks = []
    for m in range(nmono):
        idx = np.nonzero(owner == m)[0]
        if len(idx):
            blocks.append(idx)
    s = Scatterer(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), np.full(len(p), sc), eps, k0, blocks=blocks,
                  name=f"DLCA {nmono} monomers N={len(p)}")
    s.ext = extend_blocks(s, blocks, layers=1)
    return s


def convex_hull_particle(size_param=5.0, target_cells=12000, axes=(1.0, 0.8, 0.6), npts=40, eps=1.33 + 0.001j,
                         seed=4, radius=1e-3):
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(npts, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    pts = u * np.asarray(axes)[None, :] * radius
    hull = ConvexHull(pts)
    eq = hull.equations
    k0 = size_param / radius
    vol = hull.volume
    sc = (vol / target_cells) ** (1.0 / 3.0)
    for _ in range(8):
        p = _lattice(radius * 1.05, sc)
        inside = np.all(p @ eq[:, :3].T + eq[:, 3][None, :] <= 0, axis=1)
        n = int(inside.sum())
        if abs(n - target_cells) <= 0.02 * target_cells:
            break
        sc *= (n / target_cells) ** (1.0 / 3.0)
    p = p[inside]
    n = len(p)
    return Scatterer(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), np.full(n, sc), np.full(n, eps, np.complex128),
                     k0, name=f"convex hull x={size_param} N={n}")


def _grf(shape, corr_cells, rng):
    """Gaussian random field on a grid (FFT filtering of white noise)."""
    w = rng.normal(size=shape)
    k = [np.fft.fftfreq(s) for s in shape]
    kk = np.meshgrid(*k, indexing="ij")
    k2 = sum(q ** 2 for q in kk)
    filt = np.exp(-2.0 * (math.pi * corr_cells) ** 2 * k2)
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    return (f - f.mean()) / f.std()


def porous_particle(size_param=12.0, target_cells=150000, porosity=0.3,
                    eps_choices=(1.7 + 0.02j, 3.0 + 0.1j, 2.5 + 1.0j), probs=(0.6, 0.3, 0.1), nblocks=400,
                    seed=5, radius=1e-3):
    from scipy.stats import norm
    rng = np.random.default_rng(seed)
    k0 = size_param / radius
    sc = radius * (4.0 / 3.0 * math.pi * (1 - porosity) / target_cells) ** (1.0 / 3.0)
    ng = int(math.ceil(2 * radius / sc))
    g = -radius + (np.arange(ng) + 0.5) * sc
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    inside = xx ** 2 + yy ** 2 + zz ** 2 < radius ** 2
    f1 = _grf((ng, ng, ng), 3.0, rng)
    solid = inside & (f1 > norm.ppf(porosity))
    f2 = _grf((ng, ng, ng), 4.0, rng)
    cuts = norm.ppf(np.cumsum(probs)[:-1])
    cat = np.digitize(f2, cuts)
    p = np.stack([xx[solid], yy[solid], zz[solid]], 1)
    eps = np.asarray(eps_choices, np.complex128)[cat[solid]]
    blocks, order = _kmeans_blocks(p, nblocks, rng)
    p, eps = p[order], eps[order]
    s = Scatterer(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), np.full(len(p), sc), eps, k0, blocks=blocks,
                  name=f"porous x={size_param} N={len(p)} blocks={len(blocks)}")
    s.ext = extend_blocks(s, blocks, layers=1)
    return s


def _kmeans_blocks(p, nb, rng, iters=12):
    """Uneven spatial blocks: k-means on cell centres; returns contiguous
    block index ranges after reordering cells by block."""
    c = p[rng.choice(len(p), nb, replace=False)]
    for _ in range(iters):
        lab = _nearest(p, c)
        for b in range(nb):
            sel = lab == b
            if sel.any():
                c[b] = p[sel].mean(0)
    lab = _nearest(p, c)
    order = np.argsort(lab, kind="stable")
    lab = lab[order]
    blocks = []
    start = 0
    for b in range(nb):
        cnt = int(np.sum(lab == b))
        if cnt:
            blocks.append(np.arange(start, start + cnt))
            start += cnt
    return blocks, order


def _nearest(p, c, chunk=20000):
    out = np.empty(len(p), np.int64)
    for i in range(0, len(p), chunk):
        d = ((p[i:i + chunk, None, :] - c[None, :, :]) ** 2).sum(2)
        out[i:i + chunk] = d.argmin(1)
    return out


def contiguous_blocks(n, nblocks):
    """Split cells 0..n-1 into nblocks contiguous ranges of near-equal size."""
    edges = np.linspace(0, n, nblocks + 1).round().astype(int)
    return [np.arange(edges[i], edges[i + 1]) for i in range(nblocks)]


def extend_blocks(s: Scatterer, blocks, layers=1):
    """Extension cells of each block by the reference rule (Extend_blocks.f90:
    1-162, Nc_extended = layers, SimScatterer%Sc = the lattice step), through
    the library's midas_extend_blocks."""
    from . import api
    cells = api.CellTable(s.x, s.y, s.z, s.sc, s.eps)
    ext, _ = api.extend_blocks(cells, blocks, float(s.sc[0]), layers)
    return ext


def divide_blocks(s: Scatterer, div_type="SPH", navg_cells=300, best_combin=0, type_s=2, layers=1):
    """Division_blocks (Division_blocks.f90:1-183) then Extend_blocks: a copy of
    the scatterer with its cells reordered block by block (as the reference
    reorders Cells), its blocks (contiguous ranges) and their extensions.
    div_type 'SPH' (octree) or 'CSH' (slabs and halvings); navg_cells is the
    reference's NBlock (avgNcells) input.  Returns (scatterer, hBlock)."""
    from . import api
    cells = api.CellTable(s.x, s.y, s.z, s.sc, s.eps)
    if div_type.upper() == "SPH":
        order, sizes, hblock = api.partition_sph(cells, navg_cells, best_combin)
    elif div_type.upper() == "CSH":
        order, sizes, hblock = api.partition_csh(cells, float(np.min(s.sc)), navg_cells, type_s)
    else:
        raise ValueError("div_type must be SPH or CSH")
    starts = np.concatenate([[0], np.cumsum(sizes)])
    blocks = [np.arange(starts[b], starts[b + 1]) for b in range(len(sizes))]
    t = Scatterer(s.x[order].copy(), s.y[order].copy(), s.z[order].copy(), s.sc[order].copy(), s.eps[order].copy(),
                  s.k0, blocks=blocks, name=f"{s.name} {div_type.upper()} blocks={len(blocks)}")
    t.ext = extend_blocks(t, blocks, layers=layers)
    return t, hblock
