"""Host-side mirror of the reference's operator interface for the DGSEM hot path, over the C ABI.

Reference interface replaced (all state lives in Fortran module globals there):
  DGTimeDerivative_weakForm(t)     src/dg/dg.f90:255-425
  TimeStepByLSERKW2(t)             src/timedisc/timestep.f90:49-120
  CALCTIMESTEP(errType)            src/equations/navierstokes/calctimestep.f90:98-186
  ProlongToFaceCons / SurfIntCons  src/interpolation/prolongtoface.t90:116-162, src/dg/surfint.t90:237-290
Arrays are numpy arrays in the reference's Fortran layouts (order='F').
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import GalxConfig, check, lib

FORMS = {"weak": _lib.GALX_FORM_WEAK, "split": _lib.GALX_FORM_SPLIT_PI}
RIEMANNS = {"lf": _lib.GALX_RIEMANN_LF, "roe": _lib.GALX_RIEMANN_ROE, "roeentropyfix": _lib.GALX_RIEMANN_ROEEF}
PRECISIONS = {"fp64": _lib.GALX_PREC_FP64, "mixed": _lib.GALX_PREC_MIXED}
# the exact functions whose manufactured source CalcSource implements (exactfunc.f90:131-139)
SOURCE_FUNCS = (4, 41, 42, 43)


def _dptr(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous and a.ndim == 1)
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a):
    assert a.dtype == np.int32 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _F(a, dtype=np.float64):
    return np.asfortranarray(a, dtype=dtype)


class DGOperator:
    """One rank's element share of the DG operator on one GPU (a galx context)."""

    def __init__(self, N, form="split", riemann="lf", precision="fp64", kappa=1.4, R=287.058, Pr=0.72, mu0=1e-4,
                 CFLScale=0.9, DFLScale=0.9, device=0):
        cfg = GalxConfig(N=N, form=FORMS[form], riemann=RIEMANNS[riemann.lower()], precision=PRECISIONS[precision],
                         kappa=kappa, R=R, Pr=Pr, mu0=mu0, CFLScale=CFLScale, DFLScale=DFLScale, device=device)
        h = C.c_void_p()
        check(lib().galx_create(C.byref(cfg), C.byref(h)), "galx_create")
        self._h = h
        self.N, self.n = N, N + 1
        self.cfg = dict(N=N, form=form, riemann=riemann, precision=precision, kappa=kappa, R=R, Pr=Pr, mu0=mu0,
                        CFLScale=CFLScale, DFLScale=DFLScale, device=device)
        self.nElems = self.nSides = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().galx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # InitMesh counterpart
    def set_mesh(self, mesh):
        """Copy the mesh (reference layouts) to the device.  The mesh must be built for this operator's degree:
        galx_set_mesh reads every array at the sizes the context's N implies."""
        m = mesh
        if m.N != self.N:
            raise ValueError(f"mesh is built for N={m.N}, the operator for N={self.N}")
        n, nE, nS = self.n, int(m.nElems), int(m.nSides)
        shapes = {"ElemToSide": (3, 6, nE), "SideToElem": (5, nS), "S2V2": (2, n, n, 5, 6),
                  "Metrics_fTilde": (3, n, n, n, nE), "Metrics_gTilde": (3, n, n, n, nE),
                  "Metrics_hTilde": (3, n, n, n, nE), "sJ": (n, n, n, nE), "NormVec": (3, n, n, nS),
                  "TangVec1": (3, n, n, nS), "TangVec2": (3, n, n, nS), "SurfElem": (n, n, nS)}
        for name, shape in shapes.items():
            got = np.shape(getattr(m, name))
            if tuple(got) != shape:
                raise ValueError(f"mesh.{name} has shape {tuple(got)}, expected {shape}")
        keep = [_F(m.ElemToSide, np.int32), _F(m.SideToElem, np.int32), _F(m.S2V2, np.int32),
                _F(m.Metrics_fTilde), _F(m.Metrics_gTilde), _F(m.Metrics_hTilde), _F(m.sJ), _F(m.NormVec),
                _F(m.TangVec1), _F(m.TangVec2), _F(m.SurfElem)]
        check(lib().galx_set_mesh(self._h, m.nElems, m.nSides, _iptr(keep[0]), _iptr(keep[1]), _iptr(keep[2]),
                                  *[_dptr(a) for a in keep[3:]]), "galx_set_mesh")
        self.nElems, self.nSides = m.nElems, m.nSides

    def _shape(self):
        return (5, self.n, self.n, self.n, self.nElems)

    def set_state(self, U):
        U = _F(U)
        assert U.shape == self._shape(), (U.shape, self._shape())
        check(lib().galx_set_state(self._h, _dptr(U)), "galx_set_state")

    # InitExactFunc's doCalcSource + CalcSource (src/equations/navierstokes/idealgas/exactfunc.f90:131-139,915-1111)
    def set_source(self, exactFunc, AdvVel=(0.0, 0.0, 0.0), Elem_xGP=None):
        """Manufactured source of ExactFunc 4, 41, 42 or 43 on every later RHS evaluation, at the stage time;
        any other exactFunc switches it off.  Elem_xGP(3,0:N,0:N,0:N,nElems): the node coordinates, required
        when the source is switched on."""
        adv = _F(np.asarray(AdvVel, dtype=np.float64).reshape(3))
        on = int(exactFunc) in SOURCE_FUNCS
        if Elem_xGP is None:
            if on and self.nElems is not None:
                raise ValueError(f"ExactFunc {exactFunc} needs the node coordinates Elem_xGP")
            x = np.zeros(1, order="F")  # not read: no source, or no mesh yet (the library refuses the call)
        else:
            x = _F(Elem_xGP)
            if self.nElems is not None and x.shape != (3, self.n, self.n, self.n, self.nElems):
                raise ValueError(f"Elem_xGP has shape {x.shape}, expected {(3, self.n, self.n, self.n, self.nElems)}")
        check(lib().galx_set_source(self._h, int(exactFunc), _dptr(adv), _dptr(x)), "galx_set_source")

    def get_state(self):
        U = np.zeros(self._shape(), order="F")
        check(lib().galx_get_state(self._h, _dptr(U)), "galx_get_state")
        return U

    def set_state_device(self, ptr):
        check(lib().galx_set_state_device(self._h, C.c_void_p(ptr)), "galx_set_state_device")

    def get_state_device(self, ptr):
        check(lib().galx_get_state_device(self._h, C.c_void_p(ptr)), "galx_get_state_device")

    def get_ut_device(self, ptr):
        check(lib().galx_get_ut_device(self._h, C.c_void_p(ptr)), "galx_get_ut_device")

    # DGTimeDerivative_weakForm
    def DGTimeDerivative_weakForm(self, t=0.0):
        check(lib().galx_rhs(self._h, float(t)), "galx_rhs")
        Ut = np.zeros(self._shape(), order="F")
        check(lib().galx_get_ut(self._h, _dptr(Ut)), "galx_get_ut")
        return Ut

    def rhs_only(self, t=0.0):
        """DGTimeDerivative_weakForm without copying Ut back."""
        check(lib().galx_rhs(self._h, float(t)), "galx_rhs")

    # TimeStepByLSERKW2
    def TimeStepByLSERKW2(self, t, dt, nSteps=1):
        check(lib().galx_lserk_steps(self._h, float(t), float(dt), int(nSteps)), "galx_lserk_steps")
        return t + nSteps * dt

    def lserk_stage(self, t, dt, iStage):
        """Stage iStage (1..5) of TimeStepByLSERKW2."""
        check(lib().galx_lserk_stage(self._h, float(t), float(dt), int(iStage)), "galx_lserk_stage")

    def get_gradients(self, nVarLifting=4):
        """Lifted gradients (gradUx, gradUy, gradUz) of the current state, each (nVarLifting,0:N,0:N,0:N,nElems)
        over the lifting variables: (u,v,w,T) for nVarLifting=4 (PP_OPTLIFT), (rho,u,v,w,T) for 5 (the
        reference's default lifting)."""
        n = self.n
        out = [np.zeros((nVarLifting, n, n, n, self.nElems), order="F") for _ in range(3)]
        check(lib().galx_get_lifted_gradients(self._h, int(nVarLifting), *[_dptr(a) for a in out]),
              "galx_get_lifted_gradients")
        return tuple(out)

    # CALCTIMESTEP
    def CalcTimeStep(self):
        dt, err = C.c_double(), C.c_int()
        check(lib().galx_calc_dt(self._h, C.byref(dt), C.byref(err)), "galx_calc_dt")
        return dt.value, err.value

    # discrete adjoint (split PI + LF; FP64 VJPs, along a mixed-precision trajectory in a mixed context)
    def rhs_vjp(self, lam):
        """Ubar = (dR/dU)^T lam at the current state (host arrays in the U layout)."""
        lam = _F(lam)
        assert lam.shape == self._shape()
        Ubar = np.zeros(self._shape(), order="F")
        check(lib().galx_rhs_vjp(self._h, _dptr(lam), _dptr(Ubar)), "galx_rhs_vjp")
        return Ubar

    def rhs_vjp_device(self, lam_ptr, ubar_ptr):
        check(lib().galx_rhs_vjp_device(self._h, C.c_void_p(lam_ptr), C.c_void_p(ubar_ptr)), "galx_rhs_vjp_device")

    def set_adjoint_storage(self, mode):
        """What the reverse sweep keeps between checkpoints: "stages" (each segment's stage inputs; fastest) or
        "steps" (each segment's step inputs, the stages of a step recomputed before its VJPs; least memory)."""
        modes = {"stages": _lib.GALX_ADJ_STAGES, "steps": _lib.GALX_ADJ_STEPS}
        if mode not in modes:
            raise ValueError(f"adjoint storage {mode!r}: one of {sorted(modes)}")
        check(lib().galx_set_adjoint_storage(self._h, modes[mode]), "galx_set_adjoint_storage")

    def forward_ckpt(self, t, dt, nSteps, ckptEvery):
        """TimeStepByLSERKW2 over nSteps with the trajectory recorded for lserk_adjoint."""
        check(lib().galx_lserk_forward_ckpt(self._h, float(t), float(dt), int(nSteps), int(ckptEvery)),
              "galx_lserk_forward_ckpt")
        return t + nSteps * dt

    def lserk_adjoint(self, Ubar_final):
        """dJ/dU(initial) from dJ/dU(final) over the recorded trajectory."""
        Ubar = _F(Ubar_final).copy(order="F")
        assert Ubar.shape == self._shape()
        check(lib().galx_lserk_adjoint(self._h, _dptr(Ubar)), "galx_lserk_adjoint")
        return Ubar

    def lserk_adjoint_device(self, ubar_ptr):
        check(lib().galx_lserk_adjoint_device(self._h, C.c_void_p(ubar_ptr)), "galx_lserk_adjoint_device")

    def step_vjp(self, t, dt, Ubar_next):
        """dJ/dU_n from dJ/dU_{n+1} for one LSERK step of size dt at the current state (which is kept)."""
        Ubar = _F(Ubar_next).copy(order="F")
        assert Ubar.shape == self._shape()
        check(lib().galx_step_vjp(self._h, float(t), float(dt), _dptr(Ubar)), "galx_step_vjp")
        return Ubar

    def step_vjp_device(self, t, dt, ubar_ptr):
        check(lib().galx_step_vjp_device(self._h, float(t), float(dt), C.c_void_p(ubar_ptr)), "galx_step_vjp_device")

    def timers(self, on=True):
        check(lib().galx_timers_enable(self._h, int(on)))

    def timer(self, name):
        ms, cnt = C.c_double(), C.c_long()
        check(lib().galx_timer_read(self._h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def timers_reset(self):
        check(lib().galx_timers_reset(self._h))

    def memory_bytes(self):
        b = C.c_longlong()
        check(lib().galx_memory_bytes(self._h, C.byref(b)))
        return b.value


def ProlongToFaceCons(N, nElems, nSides, SideToElem, S2V2, Uvol, device=0):
    """EvalElemFaceGL_Device on the GPU for host arrays: Uvol(nVar,0:N,0:N,0:N,nElems) -> (U_master, U_slave)."""
    Uvol = _F(Uvol)
    nVar = Uvol.shape[0]
    n = N + 1
    Um = np.zeros((nVar, n, n, nSides), order="F")
    Us = np.zeros((nVar, n, n, nSides), order="F")
    s2e, s2v2 = _F(SideToElem, np.int32), _F(S2V2, np.int32)
    check(lib().galx_prolong_to_face(device, nVar, N, nElems, nSides, _iptr(s2e), _iptr(s2v2), _dptr(Uvol),
                                     _dptr(Um), _dptr(Us)), "galx_prolong_to_face")
    return Um, Us


def SurfIntCons(N, nElems, nSides, SideToElem, S2V2, Flux_master, Flux_slave, L_HatMinus0, L_HatPlusN, Ut=None,
                device=0):
    """SurfInt_Device on the GPU for host arrays; returns the updated Ut(nVar,0:N,0:N,0:N,nElems)."""
    Fm, Fs = _F(Flux_master), _F(Flux_slave)
    nVar = Fm.shape[0]
    n = N + 1
    Ut = np.zeros((nVar, n, n, n, nElems), order="F") if Ut is None else _F(Ut).copy(order="F")
    s2e, s2v2 = _F(SideToElem, np.int32), _F(S2V2, np.int32)
    check(lib().galx_surf_int(device, nVar, N, nElems, nSides, _iptr(s2e), _iptr(s2v2), _dptr(Fm), _dptr(Fs),
                              float(L_HatMinus0), float(L_HatPlusN), _dptr(Ut)), "galx_surf_int")
    return Ut
