"""Compute_Scattering_Quantities_1 (Compute_Scattering_Quantities.f90:1-589)
and its quadrature rules: host routines of the library
(midas_scattering_quantities_grid, midas_quad_*).

Pinned by (1) golden vectors of the reference's own getAdaptiveQuadPts2
(QUADPACK qag / qk61 in single precision, quadpack.f90), trap_2Dc and
getGaussLegendreQuadPts3, compiled from the reference sources into
oracle/_ref (tests/golden/make_quad_golden.py): bit-exact for the adaptive and
trapezoid rules; (2) the oracle's independent restatement, bit-exact against
the same vectors, and equal to the library on the assembled quantities for
every receiver / transmitter rule; (3) the Mie series for a discretised
sphere on 'un' and 'gl' grids; (4) the reference's quirks: the 'tr'
transmitter average passes the lower phi bound twice (0), 'sm' has no rule
(0), Q_abs is always the internal-field value."""
import os

import numpy as np
import pytest

from midas_amd import api

from test_physics import _dirs, mie_g, mie_q

GOLD = os.path.join(os.path.dirname(__file__), "golden", "quad_vectors.npz")


@pytest.fixture(scope="module")
def qv():
    return np.load(GOLD, allow_pickle=False)


def test_adaptive_quad_golden(oracle, qv):
    for c in range(int(qv["aq_n"])):
        nth, nph, ph, jj = map(int, qv[f"aq{c}_args"])
        a0, b0 = map(float, qv[f"aq{c}_ab"])
        f = qv[f"aq{c}_f"]
        want = float(qv[f"aq{c}_res"])
        assert api.quad_adaptive(f, nth, nph, ph, jj, a0, b0) == want, c
        assert oracle.adaptive_quad(f, nth, nph, ph, jj, a0, b0) == want, c


def test_trap_and_gauss_legendre_golden(oracle, qv):
    for t in range(int(qv["tr_n"])):
        f, (a, b, c, d) = qv[f"tr{t}_f"], qv[f"tr{t}_abcd"]
        want = float(qv[f"tr{t}_res"])
        for got in (api.quad_trap2d(f, a, b, c, d), oracle.trap2d(f, a, b, c, d)):
            assert (np.isnan(want) and np.isnan(got)) or got == want, t
    for t in range(int(qv["gl_n"])):
        n, a, b = qv[f"gl{t}_nab"]
        x, w = api.quad_gauss_legendre(int(n), a, b)
        assert np.allclose(x, qv[f"gl{t}_x"], rtol=0, atol=1e-14 * max(1.0, abs(b)))
        assert np.allclose(w, qv[f"gl{t}_w"], rtol=1e-11, atol=0)
        xo, wo = oracle.gauss_legendre(int(n), a, b)
        assert np.allclose(xo, qv[f"gl{t}_x"], atol=1e-13) and np.allclose(wo, qv[f"gl{t}_w"], rtol=1e-10)


def _grid(nt, tmin, tmax, npp, pmin, pmax):
    th, ph = np.linspace(tmin, tmax, nt), np.linspace(pmin, pmax, npp)
    T, P = np.meshgrid(th, ph, indexing="xy")   # theta fastest
    return T.ravel(), P.ravel()


def _receivers(tr_th, tr_ph, nrt=37, nrp=13):
    """A receiver grid, then the forward and backward directions of every
    transmitter (the reference appends them the same way)."""
    rth, rph = _grid(nrt, 0.0, 180.0, nrp, 0.0, 360.0)
    fth, fph = np.asarray(tr_th), np.asarray(tr_ph)
    return len(rth), np.r_[rth, fth, 180.0 - fth], np.r_[rph, fph, np.fmod(fph + 180.0, 360.0)]


def _cases():
    """Every receiver rule with the non-adaptive transmitter rules on a 5 x 3
    transmitter grid (3 x 2 with adaptive receivers), and the adaptive ('un')
    transmitter average on 7 x 1 and 13 x 1 grids: the oracle's
    single-precision QUADPACK is pure Python, and a coarse nearest-node
    staircase runs it to the subdivision limit."""
    out = [(rx, tr, (3, 2) if rx == api.INT_UN else (5, 3))
           for rx in (api.INT_UN, api.INT_GL, api.INT_TR, api.INT_SM, api.INT_SD, api.INT_LB)
           for tr in (api.INT_GL, api.INT_TR, api.INT_SM)]
    return out + [(api.INT_UN, api.INT_UN, (7, 1)), (api.INT_SD, api.INT_UN, (13, 1))]


@pytest.mark.parametrize("rx_type,tr_type,tgrid", _cases())
def test_scattering_quantities_grid_vs_oracle(oracle, rx_type, tr_type, tgrid):
    rng = np.random.default_rng(10 * rx_type + tr_type)
    tth, tph = _grid(tgrid[0], 0.0, 180.0, tgrid[1], 0.0, 180.0)
    ntr = len(tth)
    if rx_type in (api.INT_SD, api.INT_LB):
        nrx = 200
        rth0, rph0 = _dirs(nrx)
        rth = np.r_[rth0, tth, 180.0 - tth]
        rph = np.r_[rph0, tph, np.fmod(tph + 180.0, 360.0)]
    else:
        nrx, rth, rph = _receivers(tth, tph)
    # smooth in both directions (the nearest-node quadrature of a random
    # staircase would run to the subdivision limit)
    ang = lambda d: np.radians(np.asarray(d))   # noqa: E731
    fr = 1 + 0.5 * np.cos(ang(rth)) + 0.3 * np.sin(ang(rph))
    ft = 1 + 0.2 * np.cos(ang(tth)) + 0.1 * np.sin(ang(tph))
    coef = rng.normal(size=4) + 1j * rng.normal(size=4)
    S = 0.3 * np.einsum("r,t,k->rtk", fr, ft, coef).reshape(len(rth), 4 * ntr)
    ce = rng.normal(size=ntr) + 1j * rng.normal(size=ntr)
    ca = rng.normal(size=ntr) + 1j * rng.normal(size=ntr)
    w = rng.uniform(0.5, 1.5, nrx) / nrx if rx_type == api.INT_LB else None
    k0, dm = 2 * np.pi / 1e-3, 0.6e-3
    Q, av = api.scattering_quantities_grid(S, nrx, (rth, rph), rx_type, (tth, tph), tr_type, k0, dm, ce, ca, rx_w=w)
    Qo, avo = oracle.scattering_quantities_grid(S, nrx, rth, rph, rx_type, tth, tph, tr_type, k0, dm, ce, ca,
                                                rx_w=w)
    print(Q[0], Qo[0], av, avo)
    assert np.allclose(Q, Qo, rtol=1e-12, atol=0)
    assert np.allclose(av, avo, rtol=1e-12, atol=0)
    # the reference's rules: Q_abs is the internal-field value; no 'sm' rule;
    # the 'tr' transmitter average integrates over a zero phi width
    assert np.allclose(Q[:, 1], ca.real / (np.pi * (dm / 2) ** 2), rtol=1e-6)
    if rx_type == api.INT_SM:
        assert np.all(Q[:, 2] == 0) and np.all(Q[:, 4] == 0)
    if tr_type in (api.INT_SM, api.INT_TR):
        assert np.all(av == 0)


def test_grid_rules_vs_mie(oracle):
    """A discretised sphere (ka = 2, 1,791 cells) on a 91 x 25 'un' receiver
    grid and on a 'gl' grid: Q_sca and g against the Mie series within the
    discretisation error, and close to the equal-weight (sd) rule of
    Compute_Scattering_Quantities_2 on 1,200 directions."""
    from midas_amd import geometry as g
    s = g.sphere(size_param=2.0, nd=15)
    c = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    th, ph = np.array([0.0]), np.array([0.0])
    E = oracle.solve(oracle.mom_fill(c), oracle.mom_rhs(c, th, ph))
    ce, ca = oracle.cext_cabs(c, E, th, ph, q6=0)
    a_eq = (3 * s.n * s.sc[0] ** 3 / (4 * np.pi)) ** (1 / 3)
    dm = 2 * a_eq
    qe, qs = mie_q(s.k0 * a_eq, np.sqrt(s.eps[0]))
    gm = mie_g(s.k0 * a_eq, np.sqrt(s.eps[0]))
    nrx, rth, rph = _receivers(th, ph, nrt=91, nrp=25)
    S = oracle.scattering_matrix(c, E, rth, rph)
    Q, av = api.scattering_quantities_grid(S, nrx, (rth, rph), api.INT_UN, (th, ph), api.INT_UN, s.k0, dm, ce, ca)
    print("un", Q[0], "Mie", qe, qs, gm)
    assert Q[0, 0] == pytest.approx(qe, rel=2e-2)
    assert Q[0, 2] == pytest.approx(qs, rel=3e-2)
    assert Q[0, 4] == pytest.approx(gm, rel=3e-2)
    assert np.array_equal(av, Q[0])
    # Gauss-Legendre receivers on their own nodes, ascending as get_trans_Receiv
    # lays them out (get_trans_Receiv.f90:458-463); the reference then builds
    # the rule on [first node, last node], which the test accepts as it is
    xg = oracle.gauss_legendre(31, 0.0, np.pi)[0][::-1]
    xp = oracle.gauss_legendre(16, 0.0, 2 * np.pi)[0][::-1]
    T, P = np.meshgrid(np.degrees(xg), np.degrees(xp), indexing="xy")
    grth, grph = T.ravel(), P.ravel()
    Sg = oracle.scattering_matrix(c, E, np.r_[grth, 0.0, 180.0], np.r_[grph, 0.0, 180.0])
    Qg, _ = api.scattering_quantities_grid(Sg, len(grth), (np.r_[grth, 0.0, 180.0], np.r_[grph, 0.0, 180.0]),
                                           api.INT_GL, (th, ph), api.INT_GL, s.k0, dm, ce, ca)
    print("gl", Qg[0])
    assert Qg[0, 2] == pytest.approx(qs, rel=3e-2)
    # the equal-weight rule on quasi-uniform directions
    nsd = 1200
    sth, sph = _dirs(nsd)
    Ssd = oracle.scattering_matrix(c, E, np.r_[sth, 0.0, 180.0], np.r_[sph, 0.0, 0.0])
    Qsd, _ = oracle.scattering_quantities(Ssd, nsd, 1, sth, sph, th, ph, s.k0, dm, ce, ca)
    assert Q[0, 2] == pytest.approx(Qsd[0, 2], rel=2e-2)
    assert Q[0, 3] == pytest.approx(Qsd[0, 3], rel=1e-9)


def test_grid_sd_lb_equal_quantities_2():
    """Compute_Scattering_Quantities_1 with 'sd' / 'lb' receivers evaluates the
    same expressions as Compute_Scattering_Quantities_2 (Q_sca, g, Q_ext, Q_bks
    of one transmitter)."""
    from oracle import oracle as o
    rng = np.random.default_rng(5)
    nrx = 100
    rth, rph = _dirs(nrx)
    th, ph = np.array([30.0]), np.array([60.0])
    Rth, Rph = np.r_[rth, th, 180 - th], np.r_[rph, ph, np.fmod(ph + 180, 360)]
    S = rng.normal(size=(nrx + 2, 4)) + 1j * rng.normal(size=(nrx + 2, 4))
    ce, ca = np.array([1.0 + 0.5j]), np.array([0.3 + 0.1j])
    k0, dm = 2 * np.pi / 1e-3, 0.6e-3
    Q1, _ = api.scattering_quantities_grid(S, nrx, (Rth, Rph), api.INT_SD, (th, ph), api.INT_UN, k0, dm, ce, ca)
    Q2, _ = o.scattering_quantities(S, nrx, 1, rth, rph, th, ph, k0, dm, ce, ca)
    assert np.allclose(Q1, Q2, rtol=1e-12)


def test_grid_errors():
    th, ph = np.array([0.0, 10.0, 20.0]), np.array([0.0, 0.0, 5.0])   # not a theta x phi grid
    S = np.zeros((5, 12), complex)
    with pytest.raises(api.MidasError):
        api.scattering_quantities_grid(S, 3, (th, ph), api.INT_SD, (th, ph), api.INT_UN, 1.0, 1.0,
                                       np.zeros(3, complex), np.zeros(3, complex))
