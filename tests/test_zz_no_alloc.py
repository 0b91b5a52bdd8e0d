"""The allocation-free entry point midas_dev_mom_solve_noalloc against
midas_dev_mom_solve on the same device-resident inputs: two calls in a row,
the second smaller, for a mixed-precision mode and for FP64; device memory
in use does not grow across its calls; sizes beyond its limits are refused
before any work."""
import ctypes as C

import numpy as np
import pytest

from midas_amd import api, geometry as g

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    return C.CDLL("libamdhip64.so.7")


@pytest.fixture(scope="module")
def mid():
    m = api.Midas(0)
    yield m
    m.close()


def _case(name):
    print(f"\n[case] {name}", flush=True)


class DevArrays:
    """Device copies of host arrays, freed together."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.ptrs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.empty(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), C.c_int(H2D)) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), C.c_int(D2H)) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _free_bytes(hip):
    f, t = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def _solve(mid, hip, s, tth, tph, mode, noalloc):
    """One call of the chosen entry point on fresh device inputs; returns
    (X (3n x 2ntr), info dict, device bytes in use before minus after)."""
    d = DevArrays(hip)
    try:
        n3, nr = 3 * s.n, 2 * len(tth)
        dx, dy, dz, dsc = (d.put(np.asarray(a, np.float64)) for a in (s.x, s.y, s.z, s.sc))
        deps = d.put(np.ascontiguousarray(s.eps, np.complex128))
        dth, dph = d.put(np.asarray(tth, np.float64)), d.put(np.asarray(tph, np.float64))
        Z, X = d.empty(16 * n3 * n3), d.empty(16 * n3 * nr)
        cells = api.cells_t(s.n, dx, dy, dz, dsc, deps, None, None, None, None)
        opts = api.solve_opts_t(30, 0.0, 0, 0)
        info = api.solve_info_t()
        assert hip.hipDeviceSynchronize() == 0
        before = _free_bytes(hip)
        if noalloc:
            r = mid.lib.midas_dev_mom_solve_noalloc(mid.ctx, C.byref(cells), C.c_double(s.k0), C.c_int(api.Q_ALL),
                                                    C.c_int(len(tth)), dth, dph, C.c_int(mode), C.byref(opts), Z,
                                                    C.c_long(n3), X, C.c_long(n3), C.byref(info))
        else:
            tf = C.c_double(0)
            r = mid.lib.midas_dev_mom_solve(mid.ctx, C.byref(cells), C.c_double(s.k0), C.c_int(api.Q_ALL),
                                            C.c_int(len(tth)), dth, dph, C.c_int(mode), C.byref(opts), Z,
                                            C.c_long(n3), X, C.c_long(n3), C.byref(info), C.byref(tf))
        assert r == 0, mid.lib.midas_last_error().decode()
        assert hip.hipDeviceSynchronize() == 0
        grown = before - _free_bytes(hip)
        Xh = d.get(X, (nr, n3), np.complex128).T
        return Xh, info.as_dict(), grown
    finally:
        d.free()


@pytest.mark.parametrize("mode", [api.SOLVE_F16_IR, api.SOLVE_FP64])
def test_noalloc_matches_dev_mom_solve(mid, hip, mode):
    big = g.sphere(size_param=2.0, nd=10)            # 552 cells, 1,656 unknowns
    small = g.sphere(size_param=1.5, nd=8)           # 280 cells, 840 unknowns
    cases = [(big, np.array([0.0, 40.0]), np.array([0.0, 10.0])), (small, np.array([25.0]), np.array([60.0]))]
    for s, tth, tph in cases:
        _case(f"noalloc vs dev_mom_solve {api.MODE_NAMES[mode]} n={3 * s.n} rhs={2 * len(tth)}")
        Xr, ir, _ = _solve(mid, hip, s, tth, tph, mode, noalloc=False)
        Xn, inn, grown = _solve(mid, hip, s, tth, tph, mode, noalloc=True)
        err = np.linalg.norm(Xn - Xr) / np.linalg.norm(Xr)
        print("rel diff", err, "noalloc", inn, "device bytes grown", grown, flush=True)
        assert inn["info"] == 0 and inn["status"] == api.STATUS_CONVERGED
        if mode != api.SOLVE_FP64:
            assert inn["berr"] <= 1e-14
        assert err < 1e-12
        assert grown < (8 << 20)        # a workspace of this size would take > 20 MB


def test_noalloc_refuses_beyond_limits(mid, hip):
    s = g.sphere(size_param=2.0, nd=11)              # more than 682 cells
    assert 3 * s.n > 2048
    _case(f"noalloc limits n={3 * s.n}")
    d = DevArrays(hip)
    try:
        ptrs = [d.put(np.asarray(a, np.float64)) for a in (s.x, s.y, s.z, s.sc)]
        deps = d.put(np.ascontiguousarray(s.eps, np.complex128))
        dth = d.put(np.zeros(1))
        Z, X = d.empty(16), d.empty(16)
        cells = api.cells_t(s.n, *ptrs, deps, None, None, None, None)
        info = api.solve_info_t()
        r = mid.lib.midas_dev_mom_solve_noalloc(mid.ctx, C.byref(cells), C.c_double(s.k0), C.c_int(api.Q_ALL),
                                                C.c_int(1), dth, dth, C.c_int(api.SOLVE_F16_IR), None, Z,
                                                C.c_long(3 * s.n), X, C.c_long(3 * s.n), C.byref(info))
    finally:
        d.free()
    assert r < 0
    assert "at most 2048" in mid.lib.midas_last_error().decode()
