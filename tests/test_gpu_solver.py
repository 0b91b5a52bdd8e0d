"""GPU checks of the solver's control paths: the outcome of iterative
refinement (converged, FP64 fallback, or reported non-convergence with the
best iterate), the context's reusable workspace across sizes, and the CBFM
batching budgets forced down to one item per batch (every batched branch then
runs many batches) against the oracle."""
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


def _berr(A, X, B):
    """max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)"""
    R = B - A @ X
    an = np.abs(A).sum(1).max()
    return max(np.abs(R[:, j]).max() / (an * np.abs(X[:, j]).max() + np.abs(B[:, j]).max())
               for j in range(B.shape[1]))


def _ill_conditioned(n, cond, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    V, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    A = (U * np.logspace(0, -np.log10(cond), n)) @ V.conj().T
    B = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    return A, B


def test_refinement_converges_well_conditioned(mid):
    rng = np.random.default_rng(3)
    n = 700
    _case(f"refinement status converged, f16, n={n}")
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 30 * np.eye(n)
    B = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    X, info = mid.solve(A, B, api.SOLVE_F16_IR, no_fallback=True)
    print(info)
    assert info["status"] == api.STATUS_CONVERGED
    assert info["berr"] <= 1e-14
    assert _berr(A, X, B) <= 1e-14


def test_refinement_falls_back_to_fp64(mid):
    """cond(A) = 1e7 with BF16 trailing updates (unit roundoff 2^-8) and 64-wide
    panels: refinement cannot converge (cond x u >> 1).  By default the solve
    then takes the FP64 LU of A, as LAPACK's ZCGESV does, and says so."""
    n = 600
    A, B = _ill_conditioned(n, 1e7, 5)
    _case(f"refinement fallback, bf16 nb=64, cond 1e7, n={n}")
    X, info = mid.solve(A, B, api.SOLVE_BF16_IR, nb=64)
    print(info)
    Xo = np.linalg.solve(A, B)
    assert info["status"] == api.STATUS_FP64_FALLBACK
    assert info["info"] == 0
    assert _berr(A, X, B) < 1e-14
    assert _rel(X, Xo) < 1e-6          # cond x eps ~ 1e-9


def test_refinement_reports_non_convergence(mid):
    """The same system with no_fallback: the call reports MIDAS_ERR_NOT_CONVERGED
    (NotConverged here) and returns the iterate with the smallest backward
    error, whose backward error it reports."""
    n = 600
    A, B = _ill_conditioned(n, 1e7, 5)
    _case(f"refinement not converged, bf16 nb=64, cond 1e7, n={n}, no_fallback")
    with pytest.raises(api.NotConverged) as e:
        mid.solve(A, B, api.SOLVE_BF16_IR, nb=64, no_fallback=True)
    X, info = e.value.result
    print(str(e.value), info)
    assert info["status"] == api.STATUS_NOT_CONVERGED
    assert info["info"] == api.ERR_NOT_CONVERGED
    assert info["berr"] > 1e-14
    assert np.all(np.isfinite(X))
    # the reported backward error is the returned iterate's
    assert _berr(A, X, B) == pytest.approx(info["berr"], rel=1e-6)


def test_workspace_reused_across_sizes(mid):
    """One context solving n = 1301, then 701, then 1301 again (the workspace
    grows once and is reused smaller) gives what fresh contexts give."""
    rng = np.random.default_rng(21)
    sys = {}
    for n in (1301, 701):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 40 * np.eye(n)
        sys[n] = (A, rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3)))
    _case("workspace reuse f16: 1301, 701, 1301")
    X1, _ = mid.solve(*sys[1301], api.SOLVE_F16_IR)
    X2, _ = mid.solve(*sys[701], api.SOLVE_F16_IR)
    X3, _ = mid.solve(*sys[1301], api.SOLVE_F16_IR)
    mid.release_workspace()
    X4, _ = mid.solve(*sys[701], api.SOLVE_F16_IR)
    fresh = api.Midas(0)
    try:
        F1, _ = fresh.solve(*sys[1301], api.SOLVE_F16_IR)
        F2, _ = fresh.solve(*sys[701], api.SOLVE_F16_IR)
    finally:
        fresh.close()
    print("1301 again vs first", _rel(X3, X1), "701 vs fresh", _rel(X2, F2))
    assert _rel(X1, F1) < 1e-15 and _rel(X3, F1) < 1e-15
    assert _rel(X2, F2) < 1e-15 and _rel(X4, F2) < 1e-15
    assert _rel(X1, np.linalg.solve(*sys[1301])) < 1e-10


@pytest.fixture(scope="module")
def dlca6():
    return g.dlca_aggregate(nmono=6, cells_per_mono=40, seed=5)


@pytest.mark.parametrize("use_aca", [False, True])
def test_cbfm_budgets_force_batches(mid, oracle, dlca6, use_aca):
    """Every batching budget of the CBFM set to one byte: each block LU chunk
    holds one block, each column strip one block column, each dense batch one
    block pair and each ACA batch one block column, so every batched branch
    runs as many batches as it has items.  Same K and field as the oracle and
    as the default budgets (one batch each at this size)."""
    s = dlca6
    _case(f"cbfm budgets 1 byte use_aca={use_aca} n={3 * s.n} blocks={len(s.blocks)} "
          f"sizes={[len(b) for b in s.blocks]}")
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    th_c, ph_c = g.directions(30)
    tth, tph = np.array([0.0, 35.0]), np.array([0.0, 80.0])
    E0, K0, _ = mid.compute_efields_cbfme(s, s.k0, s.blocks, s.ext, (th_c, ph_c), (tth, tph), use_aca=use_aca)
    mid.set_cbfm_budgets(1, 1, 1, 1)
    try:
        E1, K1, _ = mid.compute_efields_cbfme(s, s.k0, s.blocks, s.ext, (th_c, ph_c), (tth, tph), use_aca=use_aca)
    finally:
        mid.set_cbfm_budgets(0, 0, 0, 0)
    Eo, Ko, _ = oracle.cbfm(oc, s.blocks, s.ext, th_c, ph_c, tth, tph, use_aca=use_aca)
    print("K", list(K1), "oracle", Ko, "batched vs default", _rel(E1, E0), "vs oracle", _rel(E1, Eo))
    assert list(K1) == list(K0) == list(Ko)
    assert _rel(E1, E0) < 1e-12
    assert _rel(E1, Eo) < 1e-9


@pytest.mark.parametrize("block_mode", [0, 1, 2, 4, 5])
@pytest.mark.parametrize("sr", [False, True])
def test_cbfm_block_solve_modes(mid, oracle, dlca6, block_mode, sr):
    """The CBF block systems solved by the batched FP64 LU (0) or by a batched
    complex-FP32 LU with c32 / f16 / bf16 trailing updates (1 / 2 / 4) refined
    in FP64, or by the size rule (5): the same K per block and the same field
    as the oracle (dense and SR-masked blocks)."""
    s = dlca6
    _case(f"cbfm block mode {block_mode} sr={sr} n={3 * s.n} blocks={len(s.blocks)}")
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    th_c, ph_c = g.directions(30)
    tth, tph = np.array([0.0, 35.0]), np.array([0.0, 80.0])
    old = api.tune_set("cbfm_block_mode", block_mode)
    try:
        E, K, _ = mid.compute_efields_cbfme(s, s.k0, s.blocks, s.ext, (th_c, ph_c), (tth, tph), sr=sr, res_sr=0.01)
    finally:
        api.tune_set("cbfm_block_mode", old)
    Eo, Ko, _ = oracle.cbfm(oc, s.blocks, s.ext, th_c, ph_c, tth, tph, sr=sr, res_sr=0.01)
    print("K", list(K), "oracle", Ko, "vs oracle", _rel(E, Eo))
    assert list(K) == list(Ko)
    assert _rel(E, Eo) < 1e-9


@pytest.mark.parametrize("mode", [api.SOLVE_F16_IR, api.SOLVE_BF16_IR])
@pytest.mark.parametrize("n", [1300, 2600])
def test_two_panel_steps(mid, mode, n):
    """Outer steps of 1024 columns (two 512-wide panels factored on the panel
    stream, the second updated by the first; trailing updates with k = 1024):
    n = 1300 has one such step and a last single panel, n = 2600 two steps and
    a last step whose second panel is 40 columns wide.  Same solution as the
    one-panel steps and as LAPACK.  The matrix is diagonally weighted
    (cond ~ 5): on a plain random one (cond ~ 1e3-1e4) single-term FP16 / BF16
    factors are too coarse for refinement, with one- or two-panel steps alike,
    and the solve falls back to FP64."""
    rng = np.random.default_rng(30 + n)
    _case(f"two-panel LU steps {api.MODE_NAMES[mode]} n={n}")
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 2.5 * np.sqrt(2 * n) * np.eye(n)
    B = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    X2, i2 = mid.solve(A, B, mode, nb=1024)
    X1, i1 = mid.solve(A, B, mode, nb=512)
    Xo = np.linalg.solve(A, B)
    print("two-panel", i2, "one-panel", i1, "vs numpy", _rel(X2, Xo))
    assert i2["status"] == api.STATUS_CONVERGED and i2["info"] == 0
    assert i2["berr"] <= 1e-14
    assert _rel(X2, Xo) < 1e-10
    assert _rel(X2, X1) < 1e-10


@pytest.mark.parametrize("mode", [api.SOLVE_F16_IR, api.SOLVE_C32_IR, api.SOLVE_BF16X3_IR])
@pytest.mark.parametrize("n", [1300, 2600])
def test_diag_block_pivot_search(mid, mode, n):
    """lu_pivot_search = 2: every panel searches its pivots in its diagonal
    block only and forms L21 = A21 U11^-1 by a GEMM.  On a diagonally weighted
    matrix, where LAPACK's partial pivoting makes no interchange, the reported
    max |L21| certifies the same pivots (<= 1/sqrt 2) and the solution equals
    the whole-column search's and LAPACK's (one-panel steps, and for the
    single-term split mode two-panel steps with a ragged last step)."""
    import scipy.linalg as sla
    rng = np.random.default_rng(50 + n)
    _case(f"diagonal-block pivot search {api.MODE_NAMES[mode]} n={n}")
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 2.5 * np.sqrt(2 * n) * np.eye(n)
    B = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    _, piv = sla.lu_factor(A)
    assert np.array_equal(piv, np.arange(n))   # partial pivoting keeps the diagonal
    nb = 1024 if mode == api.SOLVE_F16_IR else 512
    olds = {"lu_pivot_search": api.tune_set("lu_pivot_search", 2),
            "lu_two_panel_min_n": api.tune_set("lu_two_panel_min_n", 0)}
    try:
        X2, i2 = mid.solve(A, B, mode, nb=nb)
        api.tune_set("lu_pivot_search", 0)
        X0, i0 = mid.solve(A, B, mode, nb=nb)
    finally:
        for k, v in olds.items():
            api.tune_set(k, v)
    Xo = np.linalg.solve(A, B)
    print("diagonal-block search", i2, "whole columns", i0, "vs numpy", _rel(X2, Xo))
    assert i2["pivot_search"] == 1 and i0["pivot_search"] == 0
    assert 0.0 < i2["max_l21"] <= 2 ** -0.5
    assert i2["status"] == api.STATUS_CONVERGED and i2["info"] == 0 and i2["berr"] <= 1e-14
    assert _rel(X2, Xo) < 1e-10
    assert _rel(X2, X0) < 1e-10


def test_diag_block_pivot_search_uncertified(mid):
    """A matrix that needs interchanges across diagonal blocks (plain random):
    the diagonal-block search reports max |L21| above 1/sqrt 2, and the solve
    still returns LAPACK's solution (refinement, or its FP64 fallback)."""
    rng = np.random.default_rng(61)
    n = 1500
    _case(f"diagonal-block pivot search, uncertified, c32 n={n}")
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    B = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    old = api.tune_set("lu_pivot_search", 2)
    try:
        X, info = mid.solve(A, B, api.SOLVE_C32_IR, nb=256)
    finally:
        api.tune_set("lu_pivot_search", old)
    print(info, "vs numpy", _rel(X, np.linalg.solve(A, B)))
    assert info["pivot_search"] == 1 and info["max_l21"] > 2 ** -0.5
    assert info["status"] in (api.STATUS_CONVERGED, api.STATUS_FP64_FALLBACK)
    assert _rel(X, np.linalg.solve(A, B)) < 1e-9


@pytest.mark.parametrize("mode", [api.SOLVE_F16_IR, api.SOLVE_C32_IR])
def test_mom_diag_block_pivot_search(mid, mode):
    """The MoM solve's default (lu_pivot_search = 1): the internal field of a
    Gaussian random sphere equals the whole-column search's, the certificate
    holds (the MoM matrix keeps its pivots on the diagonal), with two-panel
    steps forced on."""
    s = g.gaussian_random_sphere(size_param=3.0, target_cells=1400, eps=1.5 + 0.01j, seed=7)
    _case(f"MoM diagonal-block pivot search {api.MODE_NAMES[mode]} cells={s.n}")
    tth, tph = np.array([0.0, 40.0]), np.array([0.0, 30.0])
    rx = (np.array([0.0, 90.0]), np.array([0.0, 0.0]))
    olds = {"lu_pivot_search": api.tune_set("lu_pivot_search", 1),
            "lu_two_panel_min_n": api.tune_set("lu_two_panel_min_n", 0)}
    try:
        _, _, _, E1, i1 = mid.compute_efields_mom(s, s.k0, (tth, tph), rx, mode=mode)
        api.tune_set("lu_pivot_search", 0)
        _, _, _, E0, i0 = mid.compute_efields_mom(s, s.k0, (tth, tph), rx, mode=mode)
    finally:
        for k, v in olds.items():
            api.tune_set(k, v)
    print("diagonal-block", i1, "whole columns", i0, "field diff", _rel(E1, E0))
    assert i1["pivot_search"] == 1 and i0["pivot_search"] == 0
    assert 0.0 < i1["max_l21"] <= 2 ** -0.5
    assert i1["status"] == api.STATUS_CONVERGED and i1["berr"] <= 1e-14
    assert _rel(E1, E0) < 1e-12


@pytest.mark.parametrize("mode", [api.SOLVE_F16_IR, api.SOLVE_C32_IR])
@pytest.mark.parametrize("ntr", [1, 2])
def test_mom_matrix_free_residuals(mid, mode, ntr):
    """MoM mixed solves with few right-hand sides take ||Z||_inf and the FP64
    residuals of refinement from the cells (mom_matrix_free = 1, no FP64
    matrix stored): same norm, same sweeps and the same field as with the
    stored matrix."""
    s = g.gaussian_random_sphere(size_param=3.0, target_cells=1200, eps=1.7 + 0.05j, seed=11)
    _case(f"MoM matrix-free residuals {api.MODE_NAMES[mode]} ntr={ntr} cells={s.n}")
    tth, tph = np.array([0.0, 50.0])[:ntr], np.array([0.0, 20.0])[:ntr]
    rx = (np.array([0.0, 90.0]), np.array([0.0, 0.0]))
    old = api.tune_set("mom_matrix_free", 1)
    try:
        _, ce1, _, E1, i1 = mid.compute_efields_mom(s, s.k0, (tth, tph), rx, mode=mode)
        api.tune_set("mom_matrix_free", 0)
        _, ce0, _, E0, i0 = mid.compute_efields_mom(s, s.k0, (tth, tph), rx, mode=mode)
    finally:
        api.tune_set("mom_matrix_free", old)
    print("matrix-free", i1, "stored", i0, "field diff", _rel(E1, E0))
    assert abs(i1["anorm"] - i0["anorm"]) <= 1e-13 * i0["anorm"]
    assert i1["status"] == i0["status"] == api.STATUS_CONVERGED
    assert i1["iterations"] == i0["iterations"] and i1["berr"] <= 1e-14
    assert _rel(E1, E0) < 1e-13
    assert _rel(ce1, ce0) < 1e-13


def test_mom_matrix_free_fallback(mid):
    """A matrix-free mixed solve that cannot reach its tolerance (one sweep,
    tol 1e-30) fills the FP64 matrix only then and returns the FP64 LU's
    field, as the stored-matrix solve does."""
    s = g.gaussian_random_sphere(size_param=3.0, target_cells=900, eps=1.5 + 0.01j, seed=12)
    _case(f"MoM matrix-free fallback cells={s.n}")
    tth, tph = np.array([0.0]), np.array([0.0])
    rx = (np.array([0.0]), np.array([0.0]))
    old = api.tune_set("mom_matrix_free", 1)
    try:
        _, _, _, Ef, inf = mid.compute_efields_mom(s, s.k0, (tth, tph), rx, mode=api.SOLVE_F16_IR, max_iter=1,
                                                   tol=1e-30)
    finally:
        api.tune_set("mom_matrix_free", old)
    _, _, _, E64, _ = mid.compute_efields_mom(s, s.k0, (tth, tph), rx, mode=api.SOLVE_FP64)
    print(inf, "vs FP64", _rel(Ef, E64))
    assert inf["status"] == api.STATUS_FP64_FALLBACK
    assert _rel(Ef, E64) < 1e-13


@pytest.mark.parametrize("nrhs", [2, 8])
def test_refinement_graph_replays_match_eager(mid, oracle, nrhs):
    """The refinement sweeps replayed from one captured hipGraph (several
    replays, with the host's backward-error readback between them) give the
    eager launches' solution bit for bit, with the same sweep count."""
    s = g.sphere(size_param=2.0, nd=10)
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    Z = oracle.mom_fill(oc)
    th = np.linspace(0.0, 70.0, nrhs // 2)
    B = oracle.mom_rhs(oc, th, np.zeros(nrhs // 2))
    _case(f"refinement graph vs eager f16 n={3 * s.n} nrhs={nrhs}")
    old = api.tune_set("refine_graph", 1)
    try:
        Xg, ig = mid.solve(Z, B, api.SOLVE_F16_IR)
        api.tune_set("refine_graph", 0)
        Xe, ie = mid.solve(Z, B, api.SOLVE_F16_IR)
    finally:
        api.tune_set("refine_graph", old)
    print("graph", ig, "eager", ie)
    assert ig["iterations"] == ie["iterations"] >= 2
    np.testing.assert_array_equal(Xg, Xe)
    assert _rel(Xg, np.linalg.solve(Z, B)) < 1e-10


@pytest.mark.parametrize("sr,iws,one_per_batch", [(False, 1, False), (True, 2, False), (False, 1, True)])
def test_cbfm_graph_matches_eager(mid, dlca6, sr, iws, one_per_batch):
    """The CBF block batches (grouped fills, SR masks, excitations, batched
    FP64 LU and solve) launched as hipGraphs give the same K and fields, bit
    for bit, as the same launches issued one by one: on the first call (graphs
    instantiated), on a repeat (the kept executable graphs updated in place)
    and with another wavenumber (same launches, other arguments), with eager
    calls in between.  one_per_batch sets the LU budget to one byte, so every
    block is a batch with a graph of its own."""
    s = dlca6
    _case(f"cbfm graph vs eager sr={sr} iws={iws} one_per_batch={one_per_batch} n={3 * s.n} "
          f"blocks={len(s.blocks)}")
    th_c, ph_c = g.directions(30)
    tth, tph = np.array([0.0, 35.0]), np.array([0.0, 80.0])

    def run(k0, graph):
        old = api.tune_set("cbfm_graph", graph)
        if one_per_batch:
            mid.set_cbfm_budgets(1, 0, 0, 0)
        try:
            E, K, _ = mid.compute_efields_cbfme(s, k0, s.blocks, s.ext, (th_c, ph_c), (tth, tph), sr=sr,
                                                res_sr=0.01, iws=iws)
        finally:
            api.tune_set("cbfm_graph", old)
            if one_per_batch:
                mid.set_cbfm_budgets(0, 0, 0, 0)
        return E, list(K)

    k1 = 1.03 * s.k0
    ref = {s.k0: run(s.k0, 0), k1: run(k1, 0)}
    for k0 in (s.k0, s.k0, k1, s.k0):
        Eg, Kg = run(k0, 1)
        Ee, Ke = ref[k0]
        print("k0", k0, "K", Kg, "graph vs eager", _rel(Eg, Ee))
        assert Kg == Ke
        np.testing.assert_array_equal(Eg, Ee)
        # eager work between the graph launches
        Ex, Kx = run(k0, 0)
        assert Kx == Ke
        np.testing.assert_array_equal(Ex, Ee)
    assert not np.array_equal(ref[s.k0][0], ref[k1][0])


@pytest.mark.parametrize("name,value", [("panel_leaf", 32), ("panel_leaf", 128), ("trail_group_kib", 512),
                                        ("lu_nb_split", 768), ("lu_nb_fp64", 256), ("u12_split", 0),
                                        ("l21_split", 0), ("lu_pivot_search", 2), ("lu_lookahead", 0)])
def test_tuned_parameters_keep_results(mid, name, value):
    """Every schedule parameter the autotuner varies changes the time, not the
    solution: FP64 for lu_nb_fp64, the FP16 split mode (with two-panel steps
    forced on) for the others (lu_pivot_search = 2: the diagonal-block search
    with its split-precision L21)."""
    rng = np.random.default_rng(40)
    n = 2600
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 2.5 * np.sqrt(2 * n) * np.eye(n)
    B = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    mode = api.SOLVE_FP64 if name == "lu_nb_fp64" else api.SOLVE_F16_IR
    _case(f"tuning {name} = {value}, {api.MODE_NAMES[mode]} n={n}")
    olds = {name: api.tune_set(name, value), "lu_two_panel_min_n": api.tune_set("lu_two_panel_min_n", 0)}
    try:
        X, info = mid.solve(A, B, mode)
    finally:
        for k, v in olds.items():
            api.tune_set(k, v)
    print(info)
    assert info["status"] == api.STATUS_CONVERGED and info["info"] == 0
    assert _rel(X, np.linalg.solve(A, B)) < 1e-10
