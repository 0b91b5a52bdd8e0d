"""GPU unit tests of the matrix-core building blocks against numpy."""
import ctypes as C

import numpy as np
import pytest

from midas_amd import api

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mid():
    m = api.Midas(0)
    yield m
    m.close()


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _probe_expected(which):
    K = {0: 4, 1: 4, 2: 32, 3: 32, 4: 16, 5: 16}[which]
    M = 32 if which >= 4 else 16
    i = np.arange(M)[:, None]
    k = np.arange(K)[None, :]
    A = ((i + 2 * k) % 7 - 3).astype(float)
    kk = np.arange(K)[:, None]
    j = np.arange(M)[None, :]
    B = ((3 * kk + j) % 5 - 2).astype(float)
    return A @ B


@pytest.mark.parametrize("which", [0, 1, 2, 3, 4, 5])
def test_mfma_lane_layout(mid, which):
    print(f"\n[case] mfma probe {which}", flush=True)
    M = 32 if which >= 4 else 16
    out = np.zeros(M * M)
    r = mid.lib.midas_test_mfma_probe(mid.ctx, C.c_int(which), _p(out))
    assert r == 0
    D = out.reshape(M, M, order="F")
    E = _probe_expected(which)
    if not np.array_equal(D, E):
        print("got\n", D[:4, :6], "\nexpected\n", E[:4, :6])
        print("transposed match:", np.array_equal(D, E.T))
    assert np.array_equal(D, E)


def _gemm(mid, prec, A, B, C0, alpha, beta, tA, tB, k=None, n=None):
    m = C0.shape[0]
    n = n or C0.shape[1]
    k = k if k is not None else (A.shape[0] if tA else A.shape[1])
    Cm = np.asfortranarray(C0.copy())
    a = np.array([alpha.real, alpha.imag])
    b = np.array([beta.real, beta.imag])
    A = np.asfortranarray(A)
    B = np.asfortranarray(B)
    r = mid.lib.midas_test_gemm(mid.ctx, C.c_int(prec), C.c_int(m), C.c_int(n), C.c_int(k), _p(a), _p(A),
                                C.c_long(A.shape[0]), C.c_int(tA), _p(B), C.c_long(B.shape[0]), C.c_int(tB), _p(b),
                                _p(Cm), C.c_long(m))
    assert r == 0, mid.lib.midas_last_error()
    return Cm


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("shape", [(64, 64, 16), (100, 37, 53), (1, 5, 200), (130, 2, 300)])
@pytest.mark.parametrize("tA,tB", [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 1)])
def test_gemm(mid, prec, shape, tA, tB):
    m, n, k = shape
    print(f"\n[case] gemm prec={prec} m={m} n={n} k={k} tA={tA} tB={tB}", flush=True)
    rng = np.random.default_rng(m * 1000 + n + k)
    cr = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    A = cr(k, m) if tA else cr(m, k)
    B = cr(n, k) if tB else cr(k, n)
    C0 = cr(m, n)
    alpha, beta = 0.7 - 0.2j, -1.1 + 0.4j
    got = _gemm(mid, prec, A, B, C0, alpha, beta, tA, tB)
    op = lambda X, t: X if t == 0 else (X.T if t == 1 else X.conj().T)   # 2: conjugate transpose
    ref = alpha * (op(A, tA) @ op(B, tB)) + beta * C0
    err = np.abs(got - ref).max() / (np.abs(A).max() * np.abs(B).max() * k + np.abs(C0).max())
    print("max scaled err", err)
    assert err < (1e-15 if prec == 64 else 3e-7)


@pytest.fixture(params=[0, 1, 2, 3, 4])
def zgemm_tile(request):
    api.tune_set("zgemm_tile", request.param)
    yield request.param
    api.tune_set("zgemm_tile", 2)


@pytest.mark.parametrize("shape", [(128, 128, 8), (64, 64, 16), (100, 37, 53), (1, 5, 200), (130, 2, 300),
                                   (257, 300, 129), (384, 256, 1024), (129, 131, 7), (5, 3, 0)])
def test_zgemm_nn(mid, zgemm_tile, shape):
    """The FP64 ZGEMM of the LU trailing updates and many-RHS residuals, every
    tile shape: edges of every kind (partial tiles in m, n and k, k below one
    stage, k = 0) against numpy, to FP64 rounding."""
    m, n, k = shape
    print(f"\n[case] zgemm_nn tile {zgemm_tile} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(m * 1000 + n + k)
    cr = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    A, B, C0 = cr(m, max(k, 1)), cr(max(k, 1), n), cr(m, n)
    alpha, beta = 0.7 - 0.2j, -1.1 + 0.4j
    got = _gemm(mid, 65, A, B, C0, alpha, beta, 0, 0, k=k)
    ref = alpha * (A[:, :k] @ B[:k, :]) + beta * C0
    err = np.abs(got - ref).max() / (np.abs(A).max() * np.abs(B).max() * max(k, 1) + np.abs(C0).max())
    print("max scaled err", err)
    assert err < 1e-15
    # the trailing update's form: C - A B
    got = _gemm(mid, 65, A, B, C0, -1.0 + 0j, 1.0 + 0j, 0, 0, k=k)
    assert np.abs(got - (C0 - A[:, :k] @ B[:k, :])).max() < 1e-15 * (np.abs(A).max() * np.abs(B).max() * max(k, 1)
                                                                     + np.abs(C0).max())


@pytest.fixture(params=[0, 1, 2, 3, 4, 5])
def cgemm_tile(request):
    old = api.tune_set("cgemm_tile", request.param)
    yield request.param
    api.tune_set("cgemm_tile", old)


@pytest.mark.parametrize("shape", [(128, 128, 8), (100, 37, 53), (1, 5, 200), (257, 300, 129), (129, 131, 7)])
def test_cgemm_nn(mid, cgemm_tile, shape):
    """The complex-FP32 products without transposes in every tile shape of the
    FP32 kernel (5: the 64 x 64 k_cgemm): partial tiles in m, n and k, against
    numpy on the operands rounded to complex64, to FP32 rounding."""
    m, n, k = shape
    print(f"\n[case] cgemm_nn tile {cgemm_tile} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(5 * m + n + k)
    cr = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    A, B, C0 = cr(m, k), cr(k, n), cr(m, n)
    alpha, beta = 0.7 - 0.2j, -1.1 + 0.4j
    got = _gemm(mid, 32, A, B, C0, alpha, beta, 0, 0)
    r = lambda X: X.astype(np.complex64).astype(np.complex128)
    ref = alpha * (r(A) @ r(B)) + beta * r(C0)
    err = np.abs(got - ref).max() / (np.abs(A).max() * np.abs(B).max() * k + np.abs(C0).max())
    print("max scaled err", err)
    assert err < 3e-7


@pytest.mark.parametrize("shape", [(200, 150, 40), (128, 256, 300)])
def test_zgemm_nn_batched(mid, zgemm_tile, shape):
    """A strided batch of three products (the CBFM block LUs' trailing updates)."""
    m, n, k = shape
    print(f"\n[case] zgemm_nn tile {zgemm_tile} batch 3 m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(7 + m + n + k)
    cr = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    A, B, C0 = cr(m, 3 * k), cr(k, 3 * n), cr(m, 3 * n)
    alpha, beta = -1.0 + 0j, 1.0 + 0j
    got = _gemm(mid, 66, A, B, C0, alpha, beta, 0, 0, k=k, n=n)
    for q in range(3):
        ref = C0[:, q * n:(q + 1) * n] - A[:, q * k:(q + 1) * k] @ B[:, q * n:(q + 1) * n]
        err = np.abs(got[:, q * n:(q + 1) * n] - ref).max() / (np.abs(A).max() * np.abs(B).max() * k)
        assert err < 1e-15, (q, err)


@pytest.mark.parametrize("mode,tol", [(1, 4e-3), (3, 1e-5), (11, 3e-2), (13, 2e-4)])
def test_trailing_split(mid, mode, tol):
    m, n, k = 300, 250, 96
    print(f"\n[case] trailing split mode={mode}", flush=True)
    rng = np.random.default_rng(mode)
    cr = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    A = cr(m, k)
    A /= np.abs(A).max() * 1.5      # |Re|,|Im| <= 1 as for LU multipliers
    B = cr(k, n) * np.exp(rng.normal(size=(1, n)) * 3)   # columns of very different scale
    C0 = cr(m, n)
    Cm = np.asfortranarray(C0.copy())
    A, B = np.asfortranarray(A), np.asfortranarray(B)
    r = mid.lib.midas_test_trailing(mid.ctx, C.c_int(mode), C.c_int(m), C.c_int(n), C.c_int(k), _p(A), C.c_long(m),
                                    _p(B), C.c_long(k), _p(Cm), C.c_long(m))
    assert r == 0
    ref = C0 - A @ B
    # error relative to |A||B| per column (the scale the split precision resolves)
    colscale = (np.abs(A).max() * np.abs(B).max(0) * k)[None, :]
    err = (np.abs(Cm - ref) / (colscale + np.abs(C0).max() * 1e-7)).max()
    print("mode", mode, "scaled err", err)
    assert err < tol


@pytest.mark.parametrize("one_wg", ["0", "1"])
@pytest.mark.parametrize("shape", [(257, 130, 45), (640, 384, 512), (129, 1, 16)])
@pytest.mark.parametrize("mode,tol", [(1, 4e-3), (11, 3e-2)])
def test_trailing_split_kernels(mid, mode, tol, shape, one_wg, monkeypatch):
    """Single-term split update: the two-workgroups-per-CU kernel (default) and
    the one-workgroup kernel (MIDAS_TRAIL_1WG=1) on ragged shapes, k not a
    multiple of the stage depth, and k at the panel width 512."""
    m, n, k = shape
    print(f"\n[case] trailing split kernels mode={mode} shape={shape} one_wg={one_wg}", flush=True)
    rng = np.random.default_rng(m + n + k)
    cr = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    A = cr(m, k)
    A /= np.abs(A).max() * 1.5
    B = cr(k, n) * np.exp(rng.normal(size=(1, n)) * 3)
    C0 = cr(m, n)
    Cm = np.asfortranarray(C0.copy())
    A, B = np.asfortranarray(A), np.asfortranarray(B)
    monkeypatch.setenv("MIDAS_TRAIL_1WG", one_wg)
    r = mid.lib.midas_test_trailing(mid.ctx, C.c_int(mode), C.c_int(m), C.c_int(n), C.c_int(k), _p(A), C.c_long(m),
                                    _p(B), C.c_long(k), _p(Cm), C.c_long(m))
    monkeypatch.delenv("MIDAS_TRAIL_1WG")
    assert r == 0
    ref = C0 - A @ B
    colscale = (np.abs(A).max() * np.abs(B).max(0) * k)[None, :]
    err = (np.abs(Cm - ref) / (colscale + np.abs(C0).max() * 1e-7)).max()
    print("scaled err", err)
    assert err < tol


@pytest.mark.parametrize("mode,tol", [(1, 4e-3), (11, 3e-2)])
def test_trailing_split_in_larger_matrix(mid, mode, tol):
    """Single-term split update of a block inside larger arrays (lda, ldc
    above m), k = 1024 as in the two-panel LU steps, full and edge tiles: the
    rows of C below the block and the columns right of it stay untouched."""
    m, n, k, ld, ncol = 520, 390, 1024, 611, 400
    print(f"\n[case] trailing split in a larger matrix mode={mode}", flush=True)
    rng = np.random.default_rng(7 + mode)
    cr = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    f32 = lambda x: x.astype(np.complex64).astype(complex)   # FP32-exact: the round trip keeps them
    Abig = np.asfortranarray(f32(cr(ld, k) / 3.0))
    B = np.asfortranarray(f32(cr(k, n) * np.exp(rng.normal(size=(1, n)))))
    Cbig = np.asfortranarray(f32(cr(ld, ncol)))
    C0 = Cbig.copy()
    r = mid.lib.midas_test_trailing(mid.ctx, C.c_int(mode), C.c_int(m), C.c_int(n), C.c_int(k), _p(Abig), C.c_long(ld),
                                    _p(B), C.c_long(k), _p(Cbig), C.c_long(ld))
    assert r == 0
    A = Abig[:m]
    ref = C0[:m, :n] - A @ B
    colscale = (np.abs(A).max() * np.abs(B).max(0) * k)[None, :]
    err = (np.abs(Cbig[:m, :n] - ref) / (colscale + np.abs(C0).max() * 1e-7)).max()
    print("scaled err", err)
    assert err < tol
    assert np.array_equal(Cbig[m:, :], C0[m:, :]) and np.array_equal(Cbig[:, n:], C0[:, n:])


def _heev(mid, Gs, cut=0.0, path=0):
    b, n, _ = Gs.shape
    G = np.ascontiguousarray(np.stack([np.asfortranarray(g) for g in Gs]).transpose(0, 2, 1))  # column-major each
    D = np.zeros(b * n)
    V = np.zeros(b * n * n, dtype=complex)
    r = mid.lib.midas_test_heev(mid.ctx, C.c_int(n), C.c_int(b), _p(G), _p(D), _p(V), C.c_double(cut), C.c_int(path))
    assert r == 0, mid.lib.midas_last_error()
    return D.reshape(b, n), V.reshape(b, n, n).transpose(0, 2, 1)


def _heev_cases(n, rng):
    cr = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    A = cr(n, n)
    out = {"random": A + A.conj().T, "real": (A + A.T).real.astype(complex),
           "diagonal": np.diag(rng.normal(size=n)).astype(complex)}
    # the CBF Gram matrices: X^H X of a rank-deficient X, zero-padded
    r = max(1, n // 3)
    X = cr(2 * n, r) @ cr(r, n - n // 5) * np.logspace(0, -8, n - n // 5)
    Gm = np.zeros((n, n), dtype=complex)
    Gm[:n - n // 5, :n - n // 5] = X.conj().T @ X
    out["gram"] = Gm
    # rank-deficient without padding: n - r eigenvalues at the rounding level
    # of the large ones (the tail that a deflation test relative to the
    # neighbouring diagonal entries never converges on)
    Xr = cr(2 * n, r) @ cr(r, n)
    out["rankdef"] = Xr.conj().T @ Xr
    # clusters: repeated eigenvalues under a random unitary
    Q, _ = np.linalg.qr(cr(n, n))
    lam = np.repeat(np.arange(1, n // 4 + 2, dtype=float), 4)[:n]
    out["clusters"] = (Q * lam) @ Q.conj().T
    return out


@pytest.mark.parametrize("n,path", [(1, 1), (2, 1), (7, 1), (64, 1), (182, 1), (300, 1),
                                    (1, 2), (2, 2), (7, 2), (64, 2), (182, 2), (300, 2), (870, 2)])
def test_heev_batched(mid, n, path):
    """The batched Hermitian eigensolver of the CBF subspaces (both kernels)
    against LAPACK (numpy eigh): eigenvalues to n eps |G|, residual
    G V - V diag(D) and orthonormality of V to rounding, and the invariant
    subspace of a gap-separated leading group (what the CBF basis keeps) equal
    to LAPACK's."""
    rng = np.random.default_rng(100 + n)
    cases = _heev_cases(n, rng)
    names = list(cases)
    Gs = np.stack([0.5 * (cases[k] + cases[k].conj().T) for k in names])
    print(f"\n[case] heev path {path} n={n} batch {len(names)}: {names}", flush=True)
    D, V = _heev(mid, Gs, path=path)
    for q, name in enumerate(names):
        G = Gs[q]
        nrm = max(np.abs(G).max(), 1e-300)
        lam = np.linalg.eigvalsh(G)
        assert np.all(np.diff(D[q]) >= 0), name
        eerr = np.abs(D[q] - lam).max() / nrm
        res = np.abs(G @ V[q] - V[q] * D[q]).max() / nrm
        orth = np.abs(V[q].conj().T @ V[q] - np.eye(n)).max()
        print(f"  {name:9s} eig {eerr:.1e} res {res:.1e} orth {orth:.1e}")
        assert eerr < 1e-14 * max(n, 8), name
        assert res < 1e-14 * max(n, 8), name
        assert orth < 1e-14 * max(n, 8), name
        if name in ("gram", "rankdef") and n >= 7:
            # leading eigenvectors above a relative gap: same span as LAPACK's
            w, U = np.linalg.eigh(G)
            k = int(np.sum(w > 1e-6 * w[-1]))
            P, Pr = V[q][:, -k:] @ V[q][:, -k:].conj().T, U[:, -k:] @ U[:, -k:].conj().T
            assert np.abs(P - Pr).max() < 1e-10, name


@pytest.mark.parametrize("n", [300, 844])
def test_heev_large_leading_pairs(mid, n):
    """The large-order kernel as the CBF subspaces call it: only the pairs with
    lambda >= cut lambda_max (cut = thr^2 / 4 for thr = 1e-3).  The count of
    eigenvalues above thr^2 lambda_max (the CBF count K_b) and those
    eigenvalues equal LAPACK's, the uncomputed ones are 0, and the kept
    eigenvectors span LAPACK's leading invariant subspace."""
    rng = np.random.default_rng(7 * n)
    cr = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    Gs = []
    for decay in (1e-3, 1e-5, 1e-7):   # CBF-like spectra: smooth decay into a rank-deficient tail
        X = cr(2 * n, n) * np.logspace(0, np.log10(decay), n)
        Gs.append(X.conj().T @ X)
    Gs = np.stack(Gs)
    thr = 1e-3
    print(f"\n[case] heev large leading pairs n={n}", flush=True)
    D, V = _heev(mid, Gs, cut=0.25 * thr * thr, path=2)
    for q in range(len(Gs)):
        w, U = np.linalg.eigh(Gs[q])
        K = int(np.sum(np.sqrt(np.maximum(w, 0) / w[-1]) >= thr))
        Kg = int(np.sum(np.sqrt(np.maximum(D[q], 0) / D[q][-1]) >= thr))
        kv = int(np.sum(D[q] != 0))
        print(f"  decay case {q}: K {K} got {Kg}, computed {kv}")
        assert Kg == K and kv >= K
        assert np.abs(D[q][-kv:] - w[-kv:]).max() < 1e-13 * w[-1]
        assert np.all(D[q][:-kv] == 0) and np.all(V[q][:, :-kv] == 0)
        Vk = V[q][:, -K:]
        assert np.abs(Gs[q] @ Vk - Vk * D[q][-K:]).max() < 1e-12 * w[-1]
        P, Pr = Vk @ Vk.conj().T, U[:, -K:] @ U[:, -K:].conj().T
        assert np.abs(P - Pr).max() < 1e-8


@pytest.mark.parametrize("n", [1, 7, 182, 870])
def test_pchol_rank_reduction(mid, n):
    """Pivoted Cholesky of the CBF Gram matrices: G = L L^H to the stopping
    tolerance, the exact rank of an exactly rank-deficient X^H X, a smoothly
    decaying spectrum cut where its remaining diagonal falls below 100 eps,
    and the zero matrix (rank 0)."""
    rng = np.random.default_rng(11 + n)
    cr = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    r0 = max(1, n // 4)
    X1 = cr(2 * n, r0) @ cr(r0, n)
    X2 = cr(2 * n, n) * np.logspace(0, -9, n)
    Gs = np.stack([X1.conj().T @ X1, X2.conj().T @ X2, np.zeros((n, n), complex)])
    G = np.ascontiguousarray(Gs.transpose(0, 2, 1))
    L = np.zeros(3 * n * n, dtype=complex)
    rk = np.zeros(3, dtype=np.int32)
    tol = 100 * np.finfo(float).eps
    print(f"\n[case] pchol n={n}", flush=True)
    r = mid.lib.midas_test_pchol(mid.ctx, C.c_int(n), C.c_int(3), _p(G), _p(L), rk.ctypes.data_as(C.c_void_p),
                                 C.c_double(tol))
    assert r == 0, mid.lib.midas_last_error()
    L = L.reshape(3, n, n).transpose(0, 2, 1)
    print("ranks", list(rk))
    assert rk[0] == min(r0, n) and rk[2] == 0 and np.all(L[2] == 0)
    for q in range(2):
        Gq, Lq = Gs[q], L[q]
        assert np.all(Lq[:, rk[q]:] == 0)
        nrm = np.abs(np.diag(Gq)).max()
        err = np.abs(Gq - Lq @ Lq.conj().T).max() / nrm
        print(f"  case {q}: rank {rk[q]}, |G - L L^H| / max diag {err:.1e}")
        assert err < 100 * n * np.finfo(float).eps


def _getrf(mid, prec, A, nb, search=0):
    n = A.shape[0]
    F = np.asfortranarray(A.astype(np.complex128))
    piv = np.zeros(n, dtype=np.int32)
    lm = C.c_double(-1.0)
    r = mid.lib.midas_test_getrf(mid.ctx, C.c_int(prec), C.c_int(n), _p(F), C.c_long(n), C.c_int(nb),
                                 C.c_int(search), _p(piv), C.byref(lm))
    assert r == 0, mid.lib.midas_last_error()
    return F, piv, lm.value


@pytest.mark.parametrize("n,nb", [(37, 64), (300, 64), (700, 128), (1100, 256)])
def test_getrf_pivots_match_lapack(mid, n, nb):
    """FP64 factorisation: the same pivot rows as LAPACK's zgetrf (search by
    |re| + |im|) and the same factors."""
    import scipy.linalg as sl
    print(f"\n[case] getrf fp64 n={n} nb={nb}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    F, piv, _ = _getrf(mid, 64, A, nb)
    lu, p = sl.lu_factor(A)
    assert np.array_equal(piv, p)
    assert np.abs(F - lu).max() <= 1e-10 * np.abs(lu).max()


@pytest.mark.parametrize("n,nb", [(300, 64), (1100, 256), (1500, 512)])
def test_getrf_fp32_partial_pivoting(mid, n, nb):
    """FP32 factorisation of a real matrix: every multiplier |l_ij| <= 1 (the
    pivot is the column's largest entry) and P A = L U to FP32 rounding."""
    print(f"\n[case] getrf fp32 n={n} nb={nb}", flush=True)
    rng = np.random.default_rng(7 + n)
    A = rng.standard_normal((n, n))
    F, piv, _ = _getrf(mid, 32, A, nb)
    L = np.tril(F, -1) + np.eye(n)
    U = np.triu(F)
    assert np.abs(L).max() <= 1 + 1e-6
    PA = A.astype(np.complex128).copy()
    for j in range(n):
        PA[[j, piv[j]]] = PA[[piv[j], j]]
    assert np.abs(L @ U - PA).max() <= 2e-4 * np.abs(A).max()


def test_getrf_diag_block_search(mid):
    """Diagonal-block pivot search on a diagonally dominant matrix: certified
    (max |L21| <= 1/sqrt(2)) and equal to LAPACK's factors, no interchanges."""
    import scipy.linalg as sl
    n, nb = 1200, 512
    print(f"\n[case] getrf diag-block search n={n} nb={nb}", flush=True)
    rng = np.random.default_rng(3)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / n + np.diag(np.full(n, 4.0 + 1j))
    F, piv, lm = _getrf(mid, 64, A, nb, search=1)
    lu, p = sl.lu_factor(A)
    assert np.array_equal(p, np.arange(n)) and np.array_equal(piv, p)
    assert 0 < lm <= 0.70710678
    assert np.abs(F - lu).max() <= 1e-10 * np.abs(lu).max()
