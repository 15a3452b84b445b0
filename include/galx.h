/*
 * galx.h -- C ABI of the MI355X (gfx950) DGSEM right-hand-side engine.
 *
 * Drop-in for the GALAEXI hot path: DGTimeDerivative_weakForm (src/dg/dg.f90:255-425), the low-storage RK
 * step TimeStepByLSERKW2 (src/timedisc/timestep.f90:49-120) and CALCTIMESTEP
 * (src/equations/navierstokes/calctimestep.f90:98-186).  The reference passes data through Fortran module
 * globals; here the host hands over the same arrays, in the same Fortran layouts, once per mesh and state.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative GALX_E* code on failure; galx_last_error() gives a
 *    message.  Nothing throws across the ABI.
 *  - Host arrays use the reference layouts (column-major, 0-based offsets, 1-based IDs inside connectivity):
 *      U(5,0:N,0:N,0:N,nElems), Metrics_{f,g,h}Tilde(3,0:N,0:N,0:N,nElems), sJ(0:N,0:N,0:N,nElems),
 *      NormVec/TangVec1/TangVec2(3,0:N,0:N,nSides), SurfElem(0:N,0:N,nSides),
 *      ElemToSide(3,6,nElems), SideToElem(5,nSides), S2V2(2,0:N,0:N,0:4,6)
 *    (src/dg/dg.f90:106-152, src/mesh/mesh_vars.f90:62-70, src/flexi.h:98-115, src/mesh/mappings.f90:75-235).
 *  - A context is single-threaded.  All work goes to one HIP stream owned by the context; every call returns
 *    after that stream has finished unless stated otherwise.
 *  - Only periodic conforming meshes (nBCSides = 0, no mortars), Gauss-Lobatto nodes, constant viscosity,
 *    BR1 lifting in strong form.
 */
#ifndef GALX_H
#define GALX_H
#ifdef __cplusplus
extern "C" {
#endif

#define GALX_OK 0
#define GALX_E_ARG (-1)      /* invalid argument or unsupported configuration */
#define GALX_E_STATE (-2)    /* call order violated (e.g. rhs before set_mesh) */
#define GALX_E_HIP (-3)      /* HIP runtime error */
#define GALX_E_NOMEM (-4)    /* device allocation failed */
#define GALX_E_MESH (-5)     /* inconsistent connectivity */

/* volume integral form: VolInt_weakForm (src/dg/volint.f90:129-188) or SPLIT_DG with the Pirozzoli flux
 * (src/dg/volint.f90:211-353, src/equations/navierstokes/splitflux.f90:669-761) */
#define GALX_FORM_WEAK 0
#define GALX_FORM_SPLIT_PI 1
/* Riemann solver, numbered as the reference's -DRIEMANN (src/CMakeLists.txt:100-107) */
#define GALX_RIEMANN_LF 0
#define GALX_RIEMANN_ROE 1
#define GALX_RIEMANN_ROEEF 3
/* arithmetic: plain FP64, or mixed (FP32 fluxes, lifting, volume/surface integrals and traces; FP64 U,
 * metrics, Jacobian, LSERK registers and time-step reduction) */
#define GALX_PREC_FP64 0
#define GALX_PREC_MIXED 1

typedef struct galx_config {
    int N;           /* polynomial degree (PP_N); supported set: galx_supported_degrees() */
    int form;        /* GALX_FORM_* */
    int riemann;     /* GALX_RIEMANN_* */
    int precision;   /* GALX_PREC_* */
    double kappa;    /* EOS_Vars(EOS_KAPPA) */
    double R;        /* EOS_Vars(EOS_R) */
    double Pr;       /* EOS_Vars(EOS_PR) */
    double mu0;      /* EOS_Vars(EOS_MU0) */
    double CFLScale; /* ini CFLScale, before the alpha(N)/(2N+1) scaling (src/timedisc/timedisc_func.f90:480-550) */
    double DFLScale; /* ini DFLScale */
    int device;      /* HIP device ordinal */
} galx_config;

typedef struct galx_ctx galx_ctx;

/* Library version string and the degrees N the kernels are compiled for (bit N set). */
const char *galx_version(void);
unsigned galx_supported_degrees(int precision);
const char *galx_last_error(void);

/* InitDG/InitEquation/InitTimeDisc counterpart: allocate a context (src/dg/dg.f90:67-174). */
int galx_create(const galx_config *cfg, galx_ctx **out);
int galx_destroy(galx_ctx *ctx);

/* InitMesh counterpart (src/mesh/mesh.f90:307-451): copy connectivity and geometry (host pointers, reference
 * layouts, sized for the context's N) into the device layout.  Validates SideToElem/ElemToSide consistency.  A
 * failed call leaves nothing allocated and may be repeated; a successful one cannot. */
int galx_set_mesh(galx_ctx *ctx, int nElems, int nSides, const int *ElemToSide, const int *SideToElem,
                  const int *S2V2, const double *Metrics_fTilde, const double *Metrics_gTilde,
                  const double *Metrics_hTilde, const double *sJ, const double *NormVec, const double *TangVec1,
                  const double *TangVec2, const double *SurfElem);

/* CalcSource (src/equations/navierstokes/idealgas/exactfunc.f90:915-1111) as switched on by InitExactFunc
 * (exactfunc.f90:131-139): for exactFunc 4, 41, 42, 43 every later RHS evaluation adds the manufactured source of
 * that exact function at its stage time (timestep.f90:87-90); any other value switches it off.  AdvVel(3) and
 * Elem_xGP(3,0:N,0:N,0:N,nElems) (the node coordinates, reference layout) are host arrays; needs the mesh.  When
 * the source is switched on, Elem_xGP must hold 3*nElems*(N+1)^3 doubles; switching off reads neither array.
 * Discards a trajectory recorded by galx_lserk_forward_ckpt. */
int galx_set_source(galx_ctx *ctx, int exactFunc, const double *AdvVel, const double *Elem_xGP);

/* d_U = U and U = d_U (src/dg/dg.f90:166-169, src/timedisc/timedisc_func.f90:348-349), host arrays. */
int galx_set_state(galx_ctx *ctx, const double *U);
int galx_get_state(galx_ctx *ctx, double *U);
/* Same with device pointers in the reference layout (zero host copy). */
int galx_set_state_device(galx_ctx *ctx, const double *dU);
int galx_get_state_device(galx_ctx *ctx, double *dU);

/* DGTimeDerivative_weakForm(t) (src/dg/dg.f90:255-425): Ut from the current state.  t enters only through the
 * manufactured source (galx_set_source); without one the periodic operator has no explicit time dependence. */
int galx_rhs(galx_ctx *ctx, double t);
/* Ut of the last galx_rhs in the reference layout (host / device pointer). */
int galx_get_ut(galx_ctx *ctx, double *Ut);
int galx_get_ut_device(galx_ctx *ctx, double *dUt);

/* TimeStepByLSERKW2(t) with Carpenter RK4-5 (src/timedisc/timestep.f90:49-120): advances the state by dt.
 * Ut_tmp is kept on the device between calls (zeroed by galx_set_state). */
int galx_lserk_step(galx_ctx *ctx, double t, double dt);
/* nSteps steps with a fixed dt, without host synchronisation between steps. */
int galx_lserk_steps(galx_ctx *ctx, double t, double dt, int nSteps);

/* One stage iStage = 1..5 of TimeStepByLSERKW2 (src/timedisc/timestep.f90:85-119, CurrentStage = iStage); five
 * calls in order equal one galx_lserk_step.  Lets the host act between stages as the reference's loop does. */
int galx_lserk_stage(galx_ctx *ctx, double t, double dt, int iStage);

/* Lifted gradients of the current state, as MOD_Lifting_Vars holds them after DGTimeDerivative_weakForm
 * (src/dg/lifting/lifting_br1.t90:47-167) with the PP_OPTLIFT lifting variables (u,v,w,T)
 * (src/equations/navierstokes/idealgas/eos.h:145-152): gradUx/y/z(4,0:N,0:N,0:N,nElems), FP64, host / device. */
int galx_get_gradients(galx_ctx *ctx, double *gradUx, double *gradUy, double *gradUz);
int galx_get_gradients_device(galx_ctx *ctx, double *dGradUx, double *dGradUy, double *dGradUz);
/* The same for either lifting-variable set of the reference (eos.h:133-152): nVarLifting = 5 is its default
 * (PP_OPTLIFT = 0), gradUx/y/z(5,0:N,0:N,0:N,nElems) over (rho,u,v,w,T), the density gradient lifted with the same
 * BR1 operator (FP64); nVarLifting = 4 is galx_get_gradients. */
int galx_get_lifted_gradients(galx_ctx *ctx, int nVarLifting, double *gradUx, double *gradUy, double *gradUz);
int galx_get_lifted_gradients_device(galx_ctx *ctx, int nVarLifting, double *dGradUx, double *dGradUy,
                                     double *dGradUz);

/* CALCTIMESTEP(errType) (src/equations/navierstokes/calctimestep.f90:98-186), FP64 reduction.
 * errType: 0 ok, 1 non-positive or NaN density, 2 convective / 3 viscous time step is NaN. */
int galx_calc_dt(galx_ctx *ctx, double *dt, int *errType);

/* Unit-level entry points mirroring the reference routines exercised by its unit tests (host pointers in the
 * reference layouts; work runs on the GPU of `device`).  GL nodes, non-EXACT_MM. */
/* ProlongToFaceCons -> EvalElemFaceGL_Device (src/interpolation/prolongtoface.t90:116-162,268-344) */
int galx_prolong_to_face(int device, int nVar, int N, int nElems, int nSides, const int *SideToElem,
                         const int *S2V2, const double *Uvol, double *Uface_master, double *Uface_slave);
/* SurfIntCons -> SurfInt_Device (src/dg/surfint.t90:237-290,352-586); Ut is updated in place. */
int galx_surf_int(int device, int nVar, int N, int nElems, int nSides, const int *SideToElem, const int *S2V2,
                  const double *Flux_master, const double *Flux_slave, double L_HatMinus0, double L_HatPlusN,
                  double *Ut);

/* galx_surf_int_noalloc: galx_surf_int with no device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The inputs and outputs are staged through a fixed device array of
 * GALX_NOALLOC_SCRATCH_BYTES inside the library (held on every GPU the library runs on until the process ends)
 * instead of per-call hipMalloc/hipFree; the call allocates no device memory and calls no allocating routine.
 * Requests whose staging (ElemToSide, S2V2_inv, both flux arrays and Ut, each rounded up to 256 bytes) exceeds
 * that size return GALX_E_ARG.  All calls on one GPU share its scratch: calls that overlap in time on the same GPU
 * (from several host threads) race and must be serialised by the caller. */
#define GALX_NOALLOC_SCRATCH_BYTES (16u << 20)
int galx_surf_int_noalloc(int device, int nVar, int N, int nElems, int nSides, const int *SideToElem,
                          const int *S2V2, const double *Flux_master, const double *Flux_slave,
                          double L_HatMinus0, double L_HatPlusN, double *Ut);

/* Discrete adjoint (form SPLIT_PI, riemann LF; GALX_E_ARG otherwise).  The VJPs are FP64 in both precisions;
 * a mixed-precision context records and recomputes its trajectory in the mixed form and accumulates the adjoint
 * in FP64 along it.  The reference has no
 * adjoint; its nearest interface is the matrix-free Jacobian-vector product of the implicit module,
 * MatrixVector (src/implicit/implicit.f90:480-528), which forms (R(U+eps v) - R(U))/eps by finite differences.
 * These entry points give the exact transpose of that Jacobian: Ubar = (dR/dU)^T lam, to rounding.
 * galx_rhs_vjp: at the current state, lam and Ubar in the U layout (host / device pointers). */
int galx_rhs_vjp(galx_ctx *ctx, const double *lam, double *Ubar);
int galx_rhs_vjp_device(galx_ctx *ctx, const double *dLam, double *dUbar);
/* galx_lserk_forward_ckpt: TimeStepByLSERKW2 for nSteps steps of fixed dt from the current state, recording
 * the trajectory for the reverse sweep: the state every ckptEvery steps is kept on the device, together with
 * room for what the sweep recomputes between checkpoints (galx_set_adjoint_storage).  The context ends at the
 * final state. */
int galx_lserk_forward_ckpt(galx_ctx *ctx, double t, double dt, int nSteps, int ckptEvery);
/* galx_set_adjoint_storage: what the reverse sweep keeps between checkpoints, with C = min(ckptEvery, nSteps).
 *  GALX_ADJ_STAGES (default): each segment is recomputed once and its 5*C stage inputs are kept (in the FP64 form
 *    with their trace records): device memory ~ (nSteps/ckptEvery + 5*C + 1) copies of U (FP64 form: plus 5*C
 *    trace-record sets of 0.75 U each at N=7).
 *  GALX_ADJ_STEPS: each segment is recomputed once keeping its C step inputs, and each step's stages are
 *    recomputed before its VJPs (4 more RHS evaluations per step, FP64 trace records recomputed per VJP):
 *    ~ (nSteps/ckptEvery + C + 5) copies of U.  The same VJPs on the same stage inputs: the gradient is bitwise
 *    the GALX_ADJ_STAGES one.
 * Changing the mode discards a recorded trajectory.  No counterpart in the reference (which has no adjoint). */
#define GALX_ADJ_STAGES 0
#define GALX_ADJ_STEPS 1
int galx_set_adjoint_storage(galx_ctx *ctx, int mode);
/* galx_lserk_adjoint: reverse sweep over the recorded trajectory (segments recomputed from their
 * checkpoints).  Ubar holds dJ/dU(final) on entry and dJ/dU(initial) on exit (U layout, host / device).
 * The context is left at the final state of the trajectory, which stays recorded.  galx_set_state[_device],
 * galx_lserk_step[s], galx_lserk_stage and galx_set_source discard the recording: the sweep then returns
 * GALX_E_STATE. */
int galx_lserk_adjoint(galx_ctx *ctx, double *Ubar);
int galx_lserk_adjoint_device(galx_ctx *ctx, double *dUbar);
/* galx_step_vjp: the adjoint of one TimeStepByLSERKW2 step (src/timedisc/timestep.f90:49-120) of size dt at the
 * current state U_n: Ubar holds dJ/dU_{n+1} on entry and dJ/dU_n = (dU_{n+1}/dU_n)^T dJ/dU_{n+1} on exit (U
 * layout, host / device).  The context stays at U_n; a trajectory recorded before is discarded.  A multi-step
 * run is galx_lserk_forward_ckpt followed by galx_lserk_adjoint. */
int galx_step_vjp(galx_ctx *ctx, double t, double dt, double *Ubar);
int galx_step_vjp_device(galx_ctx *ctx, double t, double dt, double *dUbar);

/* Kernel timers: when enabled, every kernel launch of the context is bracketed with HIP events on the
 * context stream; galx_timer_read returns the accumulated milliseconds and launch count of a kernel family
 * ("trace", "lift", "rhs", "dt", "adjoint"). */
int galx_timers_enable(galx_ctx *ctx, int on);
int galx_timer_read(galx_ctx *ctx, const char *name, double *ms, long *count);
int galx_timers_reset(galx_ctx *ctx);
/* Device-memory footprint of the context in bytes. */
int galx_memory_bytes(galx_ctx *ctx, long long *bytes);

#ifdef __cplusplus
}
#endif
#endif
