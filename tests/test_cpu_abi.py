"""The C-ABI library without a GPU: it loads, exports every function include/galx.h declares (and the ctypes
binding declares the same set), and argument checks that run before any device call return GALX_E_ARG."""
import ctypes as C
import os
import re

import pytest

from galaexi_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "galx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(galx_[a-z_0-9]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} missing: run python -m galaexi_amd.build")
    return _lib.lib()


def test_header_and_binding_agree():
    assert header_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_header_symbol(L):
    raw = C.CDLL(_lib.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(raw, f)]
    assert not missing, missing


def test_constants_match_header():
    txt = open(os.path.join(ROOT, "include", "galx.h")).read()
    for name in ("GALX_FORM_WEAK", "GALX_FORM_SPLIT_PI", "GALX_RIEMANN_LF", "GALX_RIEMANN_ROE", "GALX_RIEMANN_ROEEF",
                 "GALX_PREC_FP64", "GALX_PREC_MIXED"):
        val = int(re.search(rf"#define {name} (\d+)", txt).group(1))
        assert getattr(_lib, name) == val, name


def test_version_and_degrees(L):
    assert L.galx_version().decode().startswith("galx")
    assert L.galx_supported_degrees(_lib.GALX_PREC_FP64) & (1 << 7)
    assert L.galx_supported_degrees(_lib.GALX_PREC_MIXED) & (1 << 7)


@pytest.mark.parametrize("field,value", [("form", 7), ("riemann", 2), ("precision", 5), ("kappa", 1.0), ("N", 40)])
def test_create_rejects_bad_config(L, field, value):
    cfg = _lib.GalxConfig(N=5, form=1, riemann=0, precision=0, kappa=1.4, R=1.0, Pr=0.72, mu0=1e-3, CFLScale=0.9,
                          DFLScale=0.9, device=0)
    setattr(cfg, field, value)
    h = C.c_void_p()
    assert L.galx_create(C.byref(cfg), C.byref(h)) == -1
    assert h.value is None
    assert L.galx_last_error()


def test_null_context_is_an_argument_error(L):
    assert L.galx_rhs(None, 0.0) == -1
    assert L.galx_lserk_steps(None, 0.0, 1e-3, 1) == -1
    assert L.galx_rhs_vjp(None, None, None) == -1
    assert L.galx_lserk_adjoint(None, None) == -1
    assert L.galx_destroy(None) == 0


def test_fortran_binding_compiles_and_matches_header(tmp_path):
    """include/galx_mod.f90 (the ISO_C_BINDING module for a Fortran host) compiles with ROCm flang, names only
    functions the C header declares, and a program using it links against the library and gets GALX_E_ARG for
    a bad configuration."""
    import shutil
    import subprocess

    src = os.path.join(ROOT, "include", "galx_mod.f90")
    names = set(re.findall(r"NAME='(galx_[a-z_0-9]+)'", open(src).read()))
    assert names and names <= set(header_functions())
    flang = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    if not os.path.exists(flang):
        pytest.skip("flang not available")
    prog = tmp_path / "use_galx.f90"
    prog.write_text("""PROGRAM use_galx
USE ISO_C_BINDING
USE MOD_GALX
IMPLICIT NONE
TYPE(galx_config) :: cfg
TYPE(C_PTR) :: ctx
INTEGER(C_INT) :: rc
cfg = galx_config(7, GALX_FORM_SPLIT_PI, GALX_RIEMANN_LF, GALX_PREC_FP64, 1.4d0, 71.42857d0, 0.72d0, &
                  1.d0/1600.d0, 0.5d0, 0.5d0, 0)
cfg%form = 9
rc = galx_create(cfg, ctx)
PRINT *, rc
IF (rc .NE. -1) STOP 1
END PROGRAM use_galx
""")
    r = subprocess.run([flang, "-c", src, "-o", str(tmp_path / "galx_mod.o"), "-J", str(tmp_path)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    libdir = os.path.dirname(_lib.LIB_PATH)
    r = subprocess.run([flang, str(prog), str(tmp_path / "galx_mod.o"), "-I", str(tmp_path), "-L", libdir, "-lgalx",
                        f"-Wl,-rpath,{libdir}", "-o", str(tmp_path / "use_galx")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(tmp_path / "use_galx")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.split()[0] == "-1", (r.stdout, r.stderr)


def test_fortran_host_binds_the_library():
    """The Fortran host test program (built against include/galx_mod.f90) loads libgalx.so and reads the version and
    the degree masks through the ISO_C_BINDING interfaces; no device is touched."""
    import os
    import subprocess

    host = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "galx_host")
    if not os.path.exists(host):
        pytest.skip("Fortran host not built (no Fortran compiler at build time)")
    from galaexi_amd import _lib

    res = subprocess.run([host, "--version"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    L = _lib.lib()
    assert L.galx_version().decode() in res.stdout
    assert f"fp64 mask {L.galx_supported_degrees(0)}" in res.stdout
    assert f"mixed mask {L.galx_supported_degrees(1)}" in res.stdout


def _detached_operator(N, nElems=None):
    """A DGOperator with no library context: the host-side checks below run before any library call."""
    from galaexi_amd import dg

    op = dg.DGOperator.__new__(dg.DGOperator)
    op._h = None
    op.N, op.n = N, N + 1
    op.nElems = op.nSides = nElems
    return op


def test_set_source_needs_coordinates_once_a_mesh_is_set():
    """ExactFunc 4/41/42/43 read 3*nDOF node coordinates: without them (or with the wrong shape) the host mirror
    raises instead of handing the library a short array."""
    import numpy as np

    op = _detached_operator(3, nElems=8)
    for case in (4, 41, 42, 43):
        with pytest.raises(ValueError, match="Elem_xGP"):
            op.set_source(case, (0.3, 0.3, 0.3))
        with pytest.raises(ValueError, match="shape"):
            op.set_source(case, (0.3, 0.3, 0.3), np.zeros((3, 4, 4, 4, 7), order="F"))


def test_set_mesh_checks_degree_and_shapes():
    from galaexi_amd import mesh as M

    m = M.box_mesh((2, 2, 2), 3)
    with pytest.raises(ValueError, match="N=3"):
        _detached_operator(4).set_mesh(m)
    m.sJ = m.sJ[:, :, :, :7]
    with pytest.raises(ValueError, match="sJ"):
        _detached_operator(3).set_mesh(m)


def test_surf_int_noalloc_refuses_oversize_before_touching_the_device():
    """galx_surf_int_noalloc checks its staging size against its fixed scratch before any HIP call, so the
    refusal is testable without a GPU: GALX_E_ARG and a message naming the scratch."""
    L = _lib.lib()
    N, nElems, nSides = 7, 4096, 12288
    dummy_i = (C.c_int * 1)()
    dummy_d = (C.c_double * 1)()
    rc = L.galx_surf_int_noalloc(0, 5, N, nElems, nSides, dummy_i, dummy_i, dummy_d, dummy_d, 1.0, 1.0, dummy_d)
    assert rc == _lib.GALX_E_ARG
    assert b"scratch" in L.galx_last_error()
