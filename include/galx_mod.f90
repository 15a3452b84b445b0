! galx_mod.f90 -- ISO_C_BINDING interface to include/galx.h for a Fortran host (GALAEXI).
! Field order and types of galx_config follow include/galx.h; arrays are passed in the reference layouts
! (U(5,0:N,0:N,0:N,nElems), ElemToSide(3,6,nElems), SideToElem(5,nSides), S2V2(2,0:N,0:N,0:4,6), ...).
! Replaces the module-global data flow into DGTimeDerivative_weakForm (src/dg/dg.f90:255-425),
! TimeStepByLSERKW2 (src/timedisc/timestep.f90:49-120) and CALCTIMESTEP
! (src/equations/navierstokes/calctimestep.f90:98-186).
MODULE MOD_GALX
USE ISO_C_BINDING
IMPLICIT NONE
INTEGER(C_INT),PARAMETER :: GALX_FORM_WEAK=0, GALX_FORM_SPLIT_PI=1
INTEGER(C_INT),PARAMETER :: GALX_RIEMANN_LF=0, GALX_RIEMANN_ROE=1, GALX_RIEMANN_ROEEF=3
INTEGER(C_INT),PARAMETER :: GALX_PREC_FP64=0, GALX_PREC_MIXED=1
INTEGER(C_INT),PARAMETER :: GALX_ADJ_STAGES=0, GALX_ADJ_STEPS=1
TYPE,BIND(C) :: galx_config
  INTEGER(C_INT) :: N, form, riemann, precision
  REAL(C_DOUBLE) :: kappa, R, Pr, mu0, CFLScale, DFLScale
  INTEGER(C_INT) :: device
END TYPE galx_config
INTERFACE
  INTEGER(C_INT) FUNCTION galx_create(cfg,ctx) BIND(C,NAME='galx_create')
    IMPORT :: C_INT, C_PTR, galx_config
    TYPE(galx_config),INTENT(IN) :: cfg
    TYPE(C_PTR),INTENT(OUT)      :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_destroy(ctx) BIND(C,NAME='galx_destroy')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_set_mesh(ctx,nElems,nSides,ElemToSide,SideToElem,S2V2,Mf,Mg,Mh,sJ, &
                                        NormVec,TangVec1,TangVec2,SurfElem) BIND(C,NAME='galx_set_mesh')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE            :: ctx
    INTEGER(C_INT),VALUE         :: nElems, nSides
    INTEGER(C_INT),INTENT(IN)    :: ElemToSide(*), SideToElem(*), S2V2(*)
    REAL(C_DOUBLE),INTENT(IN)    :: Mf(*), Mg(*), Mh(*), sJ(*), NormVec(*), TangVec1(*), TangVec2(*), SurfElem(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_set_state(ctx,U) BIND(C,NAME='galx_set_state')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE         :: ctx
    REAL(C_DOUBLE),INTENT(IN) :: U(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_get_state(ctx,U) BIND(C,NAME='galx_get_state')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: ctx
    REAL(C_DOUBLE),INTENT(OUT) :: U(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_set_source(ctx,exactFunc,AdvVel,Elem_xGP) BIND(C,NAME='galx_set_source')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE         :: ctx
    INTEGER(C_INT),VALUE      :: exactFunc
    REAL(C_DOUBLE),INTENT(IN) :: AdvVel(3), Elem_xGP(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_rhs(ctx,t) BIND(C,NAME='galx_rhs')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE    :: ctx
    REAL(C_DOUBLE),VALUE :: t
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_get_ut(ctx,Ut) BIND(C,NAME='galx_get_ut')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: ctx
    REAL(C_DOUBLE),INTENT(OUT) :: Ut(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_lserk_steps(ctx,t,dt,nSteps) BIND(C,NAME='galx_lserk_steps')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE    :: ctx
    REAL(C_DOUBLE),VALUE :: t, dt
    INTEGER(C_INT),VALUE :: nSteps
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_lserk_stage(ctx,t,dt,iStage) BIND(C,NAME='galx_lserk_stage')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE    :: ctx
    REAL(C_DOUBLE),VALUE :: t, dt
    INTEGER(C_INT),VALUE :: iStage
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_get_gradients(ctx,gradUx,gradUy,gradUz) BIND(C,NAME='galx_get_gradients')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: ctx
    REAL(C_DOUBLE),INTENT(OUT) :: gradUx(*), gradUy(*), gradUz(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_calc_dt(ctx,dt,errType) BIND(C,NAME='galx_calc_dt')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE           :: ctx
    REAL(C_DOUBLE),INTENT(OUT)  :: dt
    INTEGER(C_INT),INTENT(OUT)  :: errType
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_rhs_vjp(ctx,lam,Ubar) BIND(C,NAME='galx_rhs_vjp')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: ctx
    REAL(C_DOUBLE),INTENT(IN)  :: lam(*)
    REAL(C_DOUBLE),INTENT(OUT) :: Ubar(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_lserk_forward_ckpt(ctx,t,dt,nSteps,ckptEvery) BIND(C,NAME='galx_lserk_forward_ckpt')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE    :: ctx
    REAL(C_DOUBLE),VALUE :: t, dt
    INTEGER(C_INT),VALUE :: nSteps, ckptEvery
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_get_lifted_gradients(ctx,nVarLifting,gradUx,gradUy,gradUz) &
      BIND(C,NAME='galx_get_lifted_gradients')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE          :: ctx
    INTEGER(C_INT),VALUE       :: nVarLifting
    REAL(C_DOUBLE),INTENT(OUT) :: gradUx(*), gradUy(*), gradUz(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_set_adjoint_storage(ctx,mode) BIND(C,NAME='galx_set_adjoint_storage')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE    :: ctx
    INTEGER(C_INT),VALUE :: mode
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_lserk_adjoint(ctx,Ubar) BIND(C,NAME='galx_lserk_adjoint')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE             :: ctx
    REAL(C_DOUBLE),INTENT(INOUT)  :: Ubar(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_step_vjp(ctx,t,dt,Ubar) BIND(C,NAME='galx_step_vjp')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR),VALUE             :: ctx
    REAL(C_DOUBLE),VALUE          :: t, dt
    REAL(C_DOUBLE),INTENT(INOUT)  :: Ubar(*)
  END FUNCTION
  TYPE(C_PTR) FUNCTION galx_version() BIND(C,NAME='galx_version')
    IMPORT :: C_PTR
  END FUNCTION
  INTEGER(C_INT) FUNCTION galx_supported_degrees(precision) BIND(C,NAME='galx_supported_degrees')
    IMPORT :: C_INT
    INTEGER(C_INT),VALUE :: precision
  END FUNCTION
  TYPE(C_PTR) FUNCTION galx_last_error() BIND(C,NAME='galx_last_error')
    IMPORT :: C_PTR
  END FUNCTION
END INTERFACE
END MODULE MOD_GALX
