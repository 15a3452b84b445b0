"""ctypes wrapper around liborc.so, the FP64 CPU restatement of the operator.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the CPU-baseline leg of the
benchmark.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NODE_G, NODE_GL, NODE_CL, NODE_VISU = 1, 2, 3, 4
RIEMANN_LF, RIEMANN_ROE, RIEMANN_ROEEF = 0, 1, 3


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")
        L.orc_nodes_and_weights.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.orc_derivative_matrix.argtypes = [C.c_int, C.c_int, dp]
        L.orc_dg_basis.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, dp]
        L.orc_prolong_to_face.argtypes = [C.c_int] * 4 + [ip, ip, dp, dp, dp]
        L.orc_surfint.argtypes = [C.c_int] * 4 + [ip, ip, dp, dp, C.c_double, C.c_double, dp]
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int] * 5 + [C.c_double] * 4
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_mesh.argtypes = [C.c_void_p, ip, ip, ip] + [dp] * 8
        L.orc_rhs.argtypes = [C.c_void_p, dp, dp]
        L.orc_rhs_t.argtypes = [C.c_void_p, dp, dp, C.c_double]
        L.orc_set_source.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.orc_source.argtypes = [C.c_void_p, C.c_double, dp]
        L.orc_array.restype = C.POINTER(C.c_double)
        L.orc_array.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long)]
        L.orc_lserk_step.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_double]
        L.orc_calc_dt.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.POINTER(C.c_double),
                                  C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_lserk_coefs.argtypes = [dp, dp, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        _LIB = L
    return _LIB


def F(a, dtype=np.float64):
    return np.asfortranarray(a, dtype=dtype)


def nodes_and_weights(N, node_type):
    x, w, wb = (np.zeros(N + 1, order="F") for _ in range(3))
    assert lib().orc_nodes_and_weights(N, node_type, x, w, wb) == 0
    return x, w, wb


def derivative_matrix(N, node_type):
    D = np.zeros((N + 1, N + 1), order="F")
    assert lib().orc_derivative_matrix(N, node_type, D) == 0
    return D


def dg_basis(N):
    n = N + 1
    x, w, lm, lp = (np.zeros(n, order="F") for _ in range(4))
    D, DhT, DVS = (np.zeros((n, n), order="F") for _ in range(3))
    lib().orc_dg_basis(N, x, w, D, DhT, lm, lp, DVS)
    return dict(xGP=x, wGP=w, D=D, D_Hat_T=DhT, L_HatMinus=lm, L_HatPlus=lp, DVolSurf=DVS)


def prolong_to_face(N, nElems, nSides, SideToElem, S2V2, Uvol, nVar):
    n = N + 1
    Um = np.zeros((nVar, n, n, nSides), order="F")
    Us = np.zeros((nVar, n, n, nSides), order="F")
    lib().orc_prolong_to_face(nVar, N, nElems, nSides, F(SideToElem, np.int32), F(S2V2, np.int32), F(Uvol), Um, Us)
    return Um, Us


def surfint(N, nElems, nSides, SideToElem, S2V2, Fm, Fs, LHm0, LHpN, nVar, Ut=None):
    n = N + 1
    if Ut is None:
        Ut = np.zeros((nVar, n, n, n, nElems), order="F")
    Ut = F(Ut).copy(order="F")
    lib().orc_surfint(nVar, N, nElems, nSides, F(SideToElem, np.int32), F(S2V2, np.int32), F(Fm), F(Fs),
                      LHm0, LHpN, Ut)
    return Ut


def lserk_coefs():
    A, b, c = (np.zeros(5, order="F") for _ in range(3))
    lib().orc_lserk_coefs(A, b, c)
    return A, b, c


def set_threads(n):
    return lib().orc_set_threads(int(n))


class Operator:
    """The FP64 operator on a mesh (galaexi_amd.mesh.Mesh-like object with reference-layout arrays)."""

    def __init__(self, mesh, split=True, riemann=RIEMANN_LF, kappa=1.4, R=287.058, Pr=0.72, mu0=1e-4):
        self.mesh = mesh
        self.N, self.n = mesh.N, mesh.N + 1
        self.h = lib().orc_create(mesh.N, mesh.nElems, mesh.nSides, int(split), int(riemann), kappa, R, Pr, mu0)
        if not self.h:
            raise RuntimeError("orc_create failed")
        m = mesh
        lib().orc_set_mesh(self.h, F(m.ElemToSide, np.int32), F(m.SideToElem, np.int32), F(m.S2V2, np.int32),
                           F(m.Metrics_fTilde), F(m.Metrics_gTilde), F(m.Metrics_hTilde), F(m.sJ), F(m.NormVec),
                           F(m.TangVec1), F(m.TangVec2), F(m.SurfElem))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def rhs(self, U, t=0.0):
        Ut = np.zeros_like(F(U), order="F")
        lib().orc_rhs_t(self.h, F(U), Ut, float(t))
        return Ut

    def set_source(self, exactFunc, advVel, Elem_xGP):
        self._xgp = F(Elem_xGP)  # the C side keeps its own copy
        lib().orc_set_source(self.h, int(exactFunc), F(np.asarray(advVel, dtype=np.float64).reshape(3)), self._xgp)

    def source(self, t):
        S = np.zeros((5, self.n, self.n, self.n, self.mesh.nElems), order="F")
        lib().orc_source(self.h, float(t), S)
        return S

    def array(self, which):
        cnt = C.c_long(0)
        p = lib().orc_array(self.h, which, C.byref(cnt))
        return np.ctypeslib.as_array(p, shape=(cnt.value,)).copy()

    def lserk_step(self, U, Ut_tmp, t, dt):
        lib().orc_lserk_step(self.h, U, Ut_tmp, t, dt)

    def calc_dt(self, U, cfl=0.9, dfl=0.9):
        dt, err, dc, dv = C.c_double(), C.c_int(), C.c_double(), C.c_double()
        lib().orc_calc_dt(self.h, F(U), cfl, dfl, C.byref(dt), C.byref(err), C.byref(dc), C.byref(dv))
        return dt.value, err.value, dc.value, dv.value
