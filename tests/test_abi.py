"""CPU checks of the C-ABI library and the host-side code: the built library
loads without a GPU and exports every function the public headers declare,
the ctypes mirror binds exactly those, the geometry generators produce the
workloads they promise, and bench.py refuses multi-GPU runs.  No compute call
is made (there is no GPU here)."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("midas_gpu.h", "midas_gpu_test.h")]


def declared(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(midas_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    path = os.path.join(ROOT, "midas_amd", "libmidas_gpu.so")
    if not os.path.exists(path):
        pytest.skip("libmidas_gpu.so not built (run __graft_entry__.build())")
    return C.CDLL(path)


def test_headers_declare_entry_points():
    names = declared(HEADERS[0])
    for must in ("midas_init", "midas_mom_fill", "midas_solve", "midas_compute_efields_mom",
                 "midas_compute_efields_cbfm", "midas_aca_block", "midas_aca_matrix", "midas_block_fill",
                 "midas_plane_wave", "midas_scattering_matrix", "midas_cext_cabs"):
        assert must in names


@pytest.mark.parametrize("header", HEADERS, ids=os.path.basename)
def test_library_exports_every_declared_symbol(lib, header):
    missing = [n for n in declared(header) if not hasattr(lib, n)]
    assert not missing, f"declared in {os.path.basename(header)} but not exported: {missing}"


def test_library_exports_are_unmangled_c(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "midas_amd", "libmidas_gpu.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    pub = set(declared(HEADERS[0])) | set(declared(HEADERS[1]))
    assert pub <= exported


def test_api_mirror_binds_declared_names():
    from midas_amd import api
    assert set(api.EXPORTS) <= set(declared(HEADERS[0])) | set(declared(HEADERS[1]))
    assert set(declared(HEADERS[0])) <= set(api.EXPORTS)


def test_no_oracle_in_product():
    """The product never imports or links the oracle."""
    pkg = os.path.join(ROOT, "midas_amd")
    for dirpath, _, files in os.walk(pkg):
        if "_build" in dirpath:
            continue
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".f90")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.sub(r"#.*|//.*|!.*", "", src).lower(), f
    out = subprocess.run(["ldd", os.path.join(ROOT, "midas_amd", "libmidas_gpu.so")], capture_output=True,
                         text=True).stdout
    assert "oracle" not in out and "midas_ref" not in out


def test_geometry_workloads():
    from midas_amd import geometry as g
    s = g.sphere(size_param=2.0, nd=16)
    assert s.n == 2176                                 # the reference lattice rule at Nd = 16
    assert g.sphere(size_param=2.0, nd=15).n == 1791
    d = g.dlca_aggregate(nmono=8, cells_per_mono=40, seed=3)
    assert len(d.blocks) == 8
    assert sorted(np.concatenate(d.blocks).tolist()) == list(range(d.n))
    for b in d.blocks:                                 # blocks are contiguous cell ranges
        assert np.array_equal(b, np.arange(b[0], b[0] + len(b)))
    for b, e in zip(d.blocks, d.ext):                  # extension cells lie outside their block
        assert not set(e.tolist()) & set(b.tolist())


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stdout


def _flang():
    import shutil
    for c in (shutil.which("flang"), "/opt/rocm/lib/llvm/bin/flang", "/opt/rocm/llvm/bin/flang"):
        if c and os.path.exists(c):
            return c
    pytest.skip("flang not available")


def test_fortran_binding_module_compiles(tmp_path):
    """The iso_c_binding module the drop-in uses compiles on its own."""
    fc = _flang()
    r = subprocess.run([fc, "-c", os.path.join(ROOT, "midas_amd", "fortran", "midas_gpu_c.f90")], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_fortran_dropin_compiles_against_reference_modules(tmp_path):
    """The drop-in subroutines compile against the reference's own modules
    (Initialization, common_variables) as oracle/_ref builds them."""
    ref = os.path.join(ROOT, "oracle", "_ref")
    if not os.path.exists(os.path.join(ref, "initialization.mod")):
        pytest.skip("oracle/_ref modules not built (needs the reference tree)")
    fc = _flang()
    for f in ("midas_gpu_c.f90", "midas_dropin.f90"):
        r = subprocess.run([fc, "-c", f"-I{ref}", f"-I{tmp_path}", os.path.join(ROOT, "midas_amd", "fortran", f)],
                           cwd=tmp_path, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_cbf_direction_grid_matches_reference(oracle):
    """The package's CBF direction grid equals the reference's getTransmitters_CBFM
    grid (pinned in the oracle tests against the compiled routine)."""
    from midas_amd import geometry as g
    for ni in (94, 200, 422):
        th, ph = g.cbf_directions_uniform(ni)
        tho, pho = oracle.transmitters_un(ni)
        assert len(th) == len(tho)
        assert np.abs(th - tho).max() < 1e-12 and np.abs(ph - pho).max() < 1e-12


def test_tune_parameters_roundtrip(lib):
    """The schedule parameters the autotuner varies: readable, settable within
    their ranges, refused outside them and for unknown names (no GPU needed)."""
    from midas_amd import api
    for name in api.TUNE_NAMES:
        v = api.tune_get(name)
        assert v == api.tune_set(name, v)
    old = api.tune_set("panel_leaf", 32)
    assert api.tune_get("panel_leaf") == 32
    api.tune_set("panel_leaf", old)
    with pytest.raises(api.MidasError):
        api.tune_set("panel_leaf", 4)
    with pytest.raises(api.MidasError):
        api.tune_get("no_such_parameter")
