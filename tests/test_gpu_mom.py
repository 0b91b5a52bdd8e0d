"""GPU parity of the MoM path (fill, excitation, solve, far field, Cext/Cabs)
against the oracle on identical seeded inputs.  Needs an MI355X and the built
library; every test prints its case first."""
import numpy as np
import pytest

from midas_amd import api, geometry as g

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mid():
    m = api.Midas(0)
    yield m
    m.close()


def _case(name):
    print(f"\n[case] {name}", flush=True)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


@pytest.fixture(scope="module")
def small(oracle):
    s = g.sphere(size_param=2.0, nd=6)
    return s, oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)


def test_cell_params(mid, small):
    s, oc = small
    _case(f"cell_params n={s.n}")
    a, che, zn, kap = mid.set_cells_params(s.sc, s.eps, s.k0)
    assert np.allclose(a, oc.a, rtol=1e-15, atol=0)
    assert np.allclose(che, oc.che, rtol=1e-15, atol=0)
    assert _rel(zn, oc.znnpp) < 1e-14
    assert _rel(kap, oc.kappa) < 1e-14


def test_mom_fill(mid, small, oracle):
    s, oc = small
    _case(f"mom_fill n={s.n}")
    Z = mid.mom_fill(s, s.k0)
    Zo = oracle.mom_fill(oc)
    err = np.abs(Z - Zo).max() / np.abs(Zo).max()
    print("fill max rel err", err)
    assert err < 1e-13
    # elementwise relative error away from cancellation
    big = np.abs(Zo) > 1e-6 * np.abs(Zo).max()
    assert np.max(np.abs(Z - Zo)[big] / np.abs(Zo)[big]) < 1e-11


@pytest.mark.parametrize("seed", [3])
def test_block_fill_ref32(mid, oracle, seed):
    s = g.dlca_aggregate(nmono=4, cells_per_mono=60, seed=seed)
    _case(f"block_fill ref32 dlca n={s.n}")
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    b0, b1 = s.blocks[0], s.blocks[1]
    ext0 = np.concatenate([b0, s.ext[0]])
    ct = api.CellTable.from_scatterer(s)
    for (r, c) in [(ext0, ext0), (b0, b1)]:
        Z = mid.green_s_tr_partial(ct.subset(r), ct.subset(c), s.k0, quirks=api.Q_ALL)
        Zo = oracle.block_fill(oc.subset(r), oc.subset(c), fp32=True)
        mism = np.mean(Z != Zo)
        err = np.abs(Z - Zo).max() / np.abs(Zo).max()
        print("ref32 block: fraction not bit-equal", mism, "max rel", err)
        assert err < 2e-7
        assert mism < 0.05
    Z64 = mid.green_s_tr_partial(ct.subset(b0), ct.subset(b1), s.k0, quirks=api.Q_PI | api.Q_LITERALS)
    Zo64 = oracle.block_fill(oc.subset(b0), oc.subset(b1), fp32=False)
    assert np.abs(Z64 - Zo64).max() / np.abs(Zo64).max() < 1e-13


def test_plane_wave(mid, small, oracle):
    s, oc = small
    th, ph = np.array([0.0, 33.0, 120.0]), np.array([0.0, 71.0, 300.0])
    _case("plane_wave fp64 + ref32")
    V = mid.plane_wave(s, s.k0, th, ph)
    assert _rel(V, oracle.mom_rhs(oc, th, ph)) < 1e-14
    V32 = mid.incident_field(s, s.k0, th, ph)
    Vo = oracle.incident_field_f32(oc, th, ph)
    assert np.abs(V32 - Vo).max() < 2e-7
    assert np.mean(V32 != Vo) < 0.05


def _mom_system(n_cells_nd=10, seed=0, ntr=1, eps=1.6 + 0.02j):
    s = g.sphere(size_param=2.0, nd=n_cells_nd, eps=eps)
    return s


@pytest.mark.parametrize("nd", [6, 10])
def test_solve_fp64(mid, oracle, nd):
    s = _mom_system(nd)
    _case(f"solve fp64 n={3 * s.n}")
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    Z = oracle.mom_fill(oc)
    B = oracle.mom_rhs(oc, [0.0, 40.0], [0.0, 10.0])
    X, info = mid.solve(Z, B, api.SOLVE_FP64)
    Xo = np.linalg.solve(Z, B)
    err = _rel(X, Xo)
    res = np.linalg.norm(Z @ X - B) / (np.linalg.norm(Z) * np.linalg.norm(X))
    print("fp64 solve rel err", err, "backward", res, info)
    assert info["info"] == 0
    assert err < 1e-12
    assert res < 1e-14
    assert abs(info["anorm"] - np.abs(Z).sum(1).max()) < 1e-12 * info["anorm"]


def test_solve_fp64_random_pivoting(mid):
    """Random non-symmetric complex matrix (pivoting active), n not a multiple of the panel."""
    rng = np.random.default_rng(7)
    n = 701
    _case(f"solve fp64 random n={n}")
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    B = rng.normal(size=(n, 5)) + 1j * rng.normal(size=(n, 5))
    X, info = mid.solve(A, B, api.SOLVE_FP64)
    Xo = np.linalg.solve(A, B)
    print("random rel err", _rel(X, Xo))
    assert _rel(X, Xo) < 1e-10
    assert np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X)) < 1e-14


@pytest.mark.parametrize("mode", [api.SOLVE_FP64, api.SOLVE_C32_IR])
def test_lu_panel_paths_agree(mid, mode, monkeypatch):
    """The one-launch-per-inner-block panel (grid barrier per column) and the
    one-launch-per-column panel pick the same pivots: same solution, n over
    three 512-row workgroups and not a multiple of the panel or inner block."""
    rng = np.random.default_rng(11)
    n = 1301
    _case(f"lu panel paths {api.MODE_NAMES[mode]} n={n}")
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    B = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    X1, i1 = mid.solve(A, B, mode)
    monkeypatch.setenv("MIDAS_LU_COLUMN_STEPS", "1")
    X2, i2 = mid.solve(A, B, mode)
    monkeypatch.delenv("MIDAS_LU_COLUMN_STEPS")
    Xo = np.linalg.solve(A, B)
    print("fused vs column-step", _rel(X1, X2), "vs numpy", _rel(X1, Xo), _rel(X2, Xo))
    assert i1["info"] == 0 and i2["info"] == 0
    assert _rel(X1, X2) < 1e-12
    assert _rel(X1, Xo) < 1e-10


@pytest.mark.parametrize("mode", [api.SOLVE_FP64, api.SOLVE_C32_IR])
@pytest.mark.parametrize("nrhs", [1, 2, 3, 8])
def test_getrs_few_rhs_paths_agree(mid, mode, nrhs, monkeypatch):
    """Few right-hand sides: the streaming substitution (one pass over each
    block column of L and U) and the GEMM-based substitution give the same
    solution, and both solve the system; n over two panels and not a multiple
    of the panel."""
    rng = np.random.default_rng(13 + nrhs)
    n = 1301
    _case(f"getrs few rhs {api.MODE_NAMES[mode]} n={n} nrhs={nrhs}")
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 40.0 * np.eye(n)
    B = rng.normal(size=(n, nrhs)) + 1j * rng.normal(size=(n, nrhs))
    X1, i1 = mid.solve(A, B, mode)
    monkeypatch.setenv("MIDAS_GETRS_GEMM", "1")
    X2, i2 = mid.solve(A, B, mode)
    monkeypatch.delenv("MIDAS_GETRS_GEMM")
    Xo = np.linalg.solve(A, B)
    print("streaming vs gemm", _rel(X1, X2), "vs numpy", _rel(X1, Xo), i1, i2)
    assert i1["info"] == 0 and i2["info"] == 0
    assert _rel(X1, Xo) < 1e-11 and _rel(X2, Xo) < 1e-11
    assert _rel(X1, X2) < 1e-11
    assert np.linalg.norm(A @ X1 - B) / (np.linalg.norm(A) * np.linalg.norm(X1)) < 1e-14


@pytest.mark.parametrize("column_steps", ["0", "1"])
def test_lu_exactly_singular(mid, column_steps, monkeypatch):
    """An exactly zero column makes U(j, j) exactly zero: the factorisation
    reports the first such j (1-based), as LAPACK ZGETRF INFO."""
    rng = np.random.default_rng(12)
    n = 600
    _case(f"lu exactly singular column_steps={column_steps} n={n}")
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A[:, 37] = 0.0
    monkeypatch.setenv("MIDAS_LU_COLUMN_STEPS", column_steps)
    _, info = mid.solve(A, np.ones((n, 1), complex), api.SOLVE_FP64)
    monkeypatch.delenv("MIDAS_LU_COLUMN_STEPS")
    assert info["info"] == 38


@pytest.mark.parametrize("mode", [api.SOLVE_C32_IR, api.SOLVE_F16_IR, api.SOLVE_F16X3_IR, api.SOLVE_BF16_IR,
                                  api.SOLVE_BF16X3_IR])
def test_solve_mixed(mid, oracle, mode):
    s = _mom_system(10)
    _case(f"solve {api.MODE_NAMES[mode]} n={3 * s.n}")
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    Z = oracle.mom_fill(oc)
    B = oracle.mom_rhs(oc, [0.0, 40.0], [0.0, 10.0])
    X, info = mid.solve(Z, B, mode)
    Xo = np.linalg.solve(Z, B)
    err = _rel(X, Xo)
    print(api.MODE_NAMES[mode], "rel err", err, info)
    assert info["info"] == 0
    assert err < 1e-10
    assert info["berr"] < 1e-13


def test_scattering_and_cext(mid, small, oracle):
    s, oc = small
    _case("scattering_matrix + cext")
    rng = np.random.default_rng(1)
    E = rng.normal(size=(3 * s.n, 4)) + 1j * rng.normal(size=(3 * s.n, 4))
    rth = np.arange(0.0, 181.0, 15.0)
    rph = np.where(np.arange(len(rth)) % 2 == 0, 0.0, 90.0)
    for beta in [(0.0,), (0.0, 30.0, 90.0)]:
        S = mid.scattering_matrix(s, s.k0, E, rth, rph, beta=beta)
        So = oracle.scattering_matrix(oc, E, rth, rph, beta)
        assert _rel(S, So) < 1e-13
    th, ph = np.array([0.0, 50.0]), np.array([0.0, 20.0])
    for q6 in (1, 0):
        q = api.Q_ALL if q6 else api.Q_ALL & ~api.Q_CEXT_SUM
        ce, ca = mid.cext_cabs(s, s.k0, E, th, ph, quirks=q)
        ceo, cao = oracle.cext_cabs(oc, E, th, ph, q6=q6)
        assert _rel(ce, ceo) < 1e-12 and _rel(ca, cao) < 1e-12


@pytest.mark.parametrize("mode", [api.SOLVE_FP64, api.SOLVE_F16X3_IR])
def test_compute_efields_mom_sphere(mid, oracle, mode):
    """Compute_EFields_MoM on the small sphere: internal field, S_total, Cext/Cabs."""
    s = g.sphere(size_param=2.0, nd=10)
    _case(f"compute_efields_mom {api.MODE_NAMES[mode]} n={3 * s.n}")
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    tth, tph = np.array([0.0, 60.0]), np.array([0.0, 45.0])
    rth = np.concatenate([np.arange(0.0, 181.0, 1.0)] * 2)
    rph = np.concatenate([np.zeros(181), np.full(181, 90.0)])
    S, ce, ca, E, info = mid.compute_efields_mom(s, s.k0, (tth, tph), (rth, rph), mode=mode)
    Z = oracle.mom_fill(oc)
    Eo = np.linalg.solve(Z, oracle.mom_rhs(oc, tth, tph))
    So = oracle.scattering_matrix(oc, Eo, rth, rph)
    ceo, cao = oracle.cext_cabs(oc, Eo, tth, tph)
    print("E err", _rel(E, Eo), "S err", _rel(S, So), "Cext", ce.real, ceo.real, info)
    assert _rel(E, Eo) < 1e-9
    assert _rel(S, So) < 1e-9
    assert _rel(ce, ceo) < 1e-9 and _rel(ca, cao) < 1e-9


def test_compute_efields_mom_c1_full_size(mid, oracle):
    """The reference's CPU-runnable case at its full size: sphere x = 2 at 16
    cells per diameter (2,176 cells, 6,528 unknowns), eps 1.6+0.02i, one plane
    wave along +x (MIDAS axes), FP64 fill + LU; internal field, S on the
    1-degree theta grid at phi = 0 and 90, Cext/Cabs against the oracle's FP64
    fill + LAPACK solve."""
    s = g.sphere(size_param=2.0, nd=16)
    assert s.n == 2176
    _case(f"compute_efields_mom C1 full size n={3 * s.n}")
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    tth, tph = np.array([0.0]), np.array([0.0])
    rth = np.concatenate([np.arange(0.0, 181.0, 1.0)] * 2)
    rph = np.concatenate([np.zeros(181), np.full(181, 90.0)])
    S, ce, ca, E, info = mid.compute_efields_mom(s, s.k0, (tth, tph), (rth, rph), mode=api.SOLVE_FP64)
    Eo = np.linalg.solve(oracle.mom_fill(oc), oracle.mom_rhs(oc, tth, tph))
    So = oracle.scattering_matrix(oc, Eo, rth, rph)
    ceo, cao = oracle.cext_cabs(oc, Eo, tth, tph)
    print("C1: E err", _rel(E, Eo), "S err", _rel(S, So), "Cext", ce.real, ceo.real, "Cabs", ca.real, cao.real)
    assert info["info"] == 0
    assert _rel(E, Eo) < 1e-11
    assert _rel(S, So) < 1e-11
    assert _rel(ce, ceo) < 1e-11 and _rel(ca, cao) < 1e-11


@pytest.mark.parametrize("mode", [api.SOLVE_FP64, api.SOLVE_F16_IR])
def test_compute_efields_mom_many_directions(mid, oracle, mode):
    """Multi-RHS (config C4 style): 60 incidence directions (120 right-hand
    sides) through one factorisation; far field on both scattering planes."""
    s = g.sphere(size_param=2.0, nd=10)
    th, ph = g.directions(60)
    _case(f"compute_efields_mom {api.MODE_NAMES[mode]} n={3 * s.n} ntr=60")
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    rth = np.concatenate([np.arange(0.0, 181.0, 2.0)] * 2)
    rph = np.concatenate([np.zeros(91), np.full(91, 90.0)])
    S, ce, ca, E, info = mid.compute_efields_mom(s, s.k0, (th, ph), (rth, rph), mode=mode)
    Eo = np.linalg.solve(oracle.mom_fill(oc), oracle.mom_rhs(oc, th, ph))
    So = oracle.scattering_matrix(oc, Eo, rth, rph)
    ceo, cao = oracle.cext_cabs(oc, Eo, th, ph)
    print("E err", _rel(E, Eo), "S err", _rel(S, So), info)
    assert E.shape == (3 * s.n, 120) and S.shape == (182, 240)
    assert _rel(E, Eo) < 1e-9 and _rel(S, So) < 1e-9
    assert _rel(ce, ceo) < 1e-9 and _rel(ca, cao) < 1e-9
    # a (voxelised) sphere: the dot-product Cext barely depends on the incidence
    # direction (the reference's component-sum formula, Q6, does)
    ce0, _ = mid.cext_cabs(s, s.k0, E, th, ph, quirks=api.Q_ALL & ~api.Q_CEXT_SUM)
    print("Cext spread over directions", np.ptp(ce0.real) / abs(ce0.real).max())
    assert np.ptp(ce0.real) < 2e-2 * abs(ce0.real).max()


@pytest.mark.parametrize("layout", ["separate", "shared"])
@pytest.mark.parametrize("rx_type", [api.RX_SD, api.RX_RF, api.RX_LB])
def test_scattering_quantities_vs_oracle(mid, oracle, rx_type, layout):
    """Orientation-averaged quantities (Compute_Scattering_Quantities_2) from
    the same S on the GPU and in the oracle: the receiver sums, the
    reference's constants, both forward / backward layouts (ntr == nrx with
    sd / lb receivers puts the forward direction among the receivers) and
    the Lebedev-weighted transmitter average."""
    rng = np.random.default_rng(5 + rx_type)
    nrx = 40 if layout == "shared" else 57
    ntr = 40 if layout == "shared" else 6
    nrx_tot = nrx + (ntr if layout == "shared" else 2 * ntr)
    _case(f"scattering quantities rx_type={rx_type} layout={layout} nrx={nrx} ntr={ntr}")
    S = rng.normal(size=(nrx_tot, 4 * ntr)) + 1j * rng.normal(size=(nrx_tot, 4 * ntr))
    rth, rph = rng.uniform(0, 180, nrx), rng.uniform(0, 360, nrx)
    tth, tph = rng.uniform(0, 180, ntr), rng.uniform(0, 360, ntr)
    rw = rng.uniform(0.5, 1.5, nrx) / nrx if rx_type == api.RX_LB else None
    tw = rng.uniform(0.5, 1.5, ntr) / ntr
    ce = rng.normal(size=ntr) + 1j * rng.normal(size=ntr)
    ca = rng.normal(size=ntr) + 1j * rng.normal(size=ntr)
    k0, dm = 2 * np.pi / 5e-4, 3e-4
    for tr_type, w in ((api.RX_SD, None), (api.RX_LB, tw)):
        Q, av = mid.scattering_quantities(S, nrx, ntr, (rth, rph), (tth, tph), k0, dm, ce, ca, rx_type=rx_type,
                                          rx_w=rw, tr_type=tr_type, tr_w=w)
        Qo, avo = oracle.scattering_quantities(S, nrx, ntr, rth, rph, tth, tph, k0, dm, ce, ca, rx_type=rx_type,
                                               rx_w=rw, tr_type=tr_type, tr_w=w)
        print("max rel", np.abs(Q - Qo).max() / np.abs(Qo).max(), np.abs(av - avo).max())
        assert np.allclose(Q, Qo, rtol=1e-12, atol=0)
        assert np.allclose(av, avo, rtol=1e-12, atol=0)


@pytest.mark.parametrize("rso_wavelengths", [0.6, 50.0])
def test_near_field_vs_oracle(mid, small, oracle, rso_wavelengths):
    """Scattered near field (Green_s_dt + projection, FP64) on the GPU against
    the oracle, receivers close to the particle and far from it."""
    s, oc = small
    rng = np.random.default_rng(8)
    E = rng.normal(size=(3 * s.n, 4)) + 1j * rng.normal(size=(3 * s.n, 4))
    rth = np.linspace(0.0, 180.0, 23)
    rph = np.linspace(0.0, 330.0, 23)
    rso = rso_wavelengths * 2 * np.pi / s.k0
    _case(f"near field rso={rso_wavelengths} wavelengths nrx={len(rth)}")
    Es = mid.near_field(s, s.k0, E, rth, rph, rso)
    Eo = oracle.near_field(oc, E, rso, rth, rph)
    print("near field rel err", _rel(Es, Eo))
    assert _rel(Es, Eo) < 1e-12

