! galx_host.f90 -- a Fortran host that drives the engine through MOD_GALX (include/galx_mod.f90) the way GALAEXI's
! time loop would: InitDG/InitMesh -> galx_create + galx_set_mesh (src/dg/dg.f90:67-174, src/mesh/mesh.f90:307-451),
! d_U = U -> galx_set_state (dg.f90:166-169), DGTimeDerivative_weakForm(t) -> galx_rhs + galx_get_ut
! (dg.f90:255-425), CALCTIMESTEP -> galx_calc_dt (calctimestep.f90:98-186), TimeStepByLSERKW2 -> galx_lserk_steps
! (timestep.f90:49-120), U = d_U -> galx_get_state (timedisc_func.f90:348-349).
!
! Usage:  galx_host --version     prints the library version and the compiled degree masks (no device needed)
!         galx_host DIR           reads DIR/in.bin, writes DIR/out.bin (unformatted stream, native byte order)
!   in.bin : int32 N, nElems, nSides, form, riemann, nSteps; float64 kappa, R, Pr, mu0, CFLScale, DFLScale, t;
!            int32 ElemToSide(3,6,nElems), SideToElem(5,nSides), S2V2(2,0:N,0:N,0:4,6);
!            float64 Metrics_fTilde, Metrics_gTilde, Metrics_hTilde (3,0:N,0:N,0:N,nElems), sJ(0:N,0:N,0:N,nElems),
!            NormVec, TangVec1, TangVec2 (3,0:N,0:N,nSides), SurfElem(0:N,0:N,nSides), U(5,0:N,0:N,0:N,nElems)
!   out.bin: float64 Ut(5,0:N,0:N,0:N,nElems) at t; float64 dt and int32 errType of CALCTIMESTEP;
!            float64 U(5,0:N,0:N,0:N,nElems) after nSteps steps of that dt from t
PROGRAM galx_host
USE ISO_C_BINDING
USE MOD_GALX
IMPLICIT NONE
CHARACTER(LEN=4096)         :: arg
INTEGER(C_INT)              :: hdr(6), N, nElems, nSides, nSteps, errType
REAL(C_DOUBLE)              :: par(7), t, dt
INTEGER(C_INT),ALLOCATABLE  :: E2S(:), S2E(:), S2V2(:)
REAL(C_DOUBLE),ALLOCATABLE  :: Mf(:), Mg(:), Mh(:), sJ(:), nv(:), t1(:), t2(:), se(:), U(:), Ut(:)
TYPE(galx_config)           :: cfg
TYPE(C_PTR)                 :: ctx
INTEGER                     :: nv3, nf2, u_in, u_out

CALL GET_COMMAND_ARGUMENT(1, arg)
IF (TRIM(arg) == '--version') THEN
  WRITE(*,'(A)') 'galx_host: '//c_str(galx_version())
  WRITE(*,'(A,I0,A,I0)') 'degrees fp64 mask ', galx_supported_degrees(GALX_PREC_FP64), &
                         ' mixed mask ', galx_supported_degrees(GALX_PREC_MIXED)
  STOP
END IF
IF (LEN_TRIM(arg) == 0) THEN
  WRITE(*,'(A)') 'usage: galx_host --version | galx_host DIR'
  STOP 2
END IF

OPEN(NEWUNIT=u_in, FILE=TRIM(arg)//'/in.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
READ(u_in) hdr, par
N = hdr(1); nElems = hdr(2); nSides = hdr(3); nSteps = hdr(6)
nv3 = (N+1)**3*nElems
nf2 = (N+1)**2*nSides
ALLOCATE(E2S(18*nElems), S2E(5*nSides), S2V2(2*(N+1)**2*30))
ALLOCATE(Mf(3*nv3), Mg(3*nv3), Mh(3*nv3), sJ(nv3), nv(3*nf2), t1(3*nf2), t2(3*nf2), se(nf2), U(5*nv3), Ut(5*nv3))
READ(u_in) E2S, S2E, S2V2, Mf, Mg, Mh, sJ, nv, t1, t2, se, U
CLOSE(u_in)

cfg%N = N; cfg%form = hdr(4); cfg%riemann = hdr(5); cfg%precision = GALX_PREC_FP64
cfg%kappa = par(1); cfg%R = par(2); cfg%Pr = par(3); cfg%mu0 = par(4)
cfg%CFLScale = par(5); cfg%DFLScale = par(6); cfg%device = 0
t = par(7)

CALL check(galx_create(cfg, ctx), 'galx_create')
CALL check(galx_set_mesh(ctx, nElems, nSides, E2S, S2E, S2V2, Mf, Mg, Mh, sJ, nv, t1, t2, se), 'galx_set_mesh')
CALL check(galx_set_state(ctx, U), 'galx_set_state')
CALL check(galx_rhs(ctx, t), 'galx_rhs')
CALL check(galx_get_ut(ctx, Ut), 'galx_get_ut')
CALL check(galx_calc_dt(ctx, dt, errType), 'galx_calc_dt')
CALL check(galx_lserk_steps(ctx, t, dt, nSteps), 'galx_lserk_steps')
CALL check(galx_get_state(ctx, U), 'galx_get_state')
CALL check(galx_destroy(ctx), 'galx_destroy')

OPEN(NEWUNIT=u_out, FILE=TRIM(arg)//'/out.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
WRITE(u_out) Ut, dt, errType, U
CLOSE(u_out)
WRITE(*,'(A,ES12.5,A,I0)') 'galx_host: done, dt ', dt, ' errType ', errType

CONTAINS

! the NUL-terminated C string at p as a Fortran string
FUNCTION c_str(p) RESULT(s)
  TYPE(C_PTR),INTENT(IN)                  :: p
  CHARACTER(LEN=:),ALLOCATABLE            :: s
  CHARACTER(KIND=C_CHAR),POINTER          :: c(:)
  INTEGER                                 :: k
  s = ''
  IF (.NOT. C_ASSOCIATED(p)) RETURN
  CALL C_F_POINTER(p, c, [4096])
  k = 1
  DO WHILE (k <= 4096)
    IF (c(k) == C_NULL_CHAR) EXIT
    s = s//c(k)
    k = k + 1
  END DO
END FUNCTION c_str

! a nonzero galx return code stops the host with the library's message, as the reference's Abort would
SUBROUTINE check(rc, what)
  INTEGER(C_INT),INTENT(IN) :: rc
  CHARACTER(*),INTENT(IN)   :: what
  IF (rc == 0) RETURN
  WRITE(*,'(A,I0,A)') 'galx_host: '//what//' returned ', rc, ': '//c_str(galx_last_error())
  STOP 1
END SUBROUTINE check

END PROGRAM galx_host
