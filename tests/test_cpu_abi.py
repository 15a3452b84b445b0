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
