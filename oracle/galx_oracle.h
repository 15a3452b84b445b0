/*
 * galx_oracle.h -- FP64 CPU restatement of the GALAEXI DGSEM right-hand side and LSERK step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only tests/, the smoke entry
 * point and the cpu_baseline leg of bench.py may load it.  The product path (galaexi_amd) never
 * links, imports or calls anything in oracle/.
 *
 * All arrays use the reference's Fortran column-major layouts with 0-based offsets and the
 * reference's 1-based element/side IDs inside the connectivity arrays:
 *   U(5,0:N,0:N,0:N,nElems), UPrim(6,...), Ut(5,...)
 *   Metrics_{f,g,h}Tilde(3,0:N,0:N,0:N,nElems), sJ(0:N,0:N,0:N,nElems)
 *   NormVec/TangVec1/TangVec2(3,0:N,0:N,nSides), SurfElem(0:N,0:N,nSides)
 *   ElemToSide(3,6,nElems) = [SideID, flip, isMaster], SideToElem(5,nSides) = [Elem,nbElem,locSide,nbLocSide,flip]
 *   S2V2(2,0:N,0:N,0:4,6)
 * (src/dg/dg.f90:106-152, src/mesh/mesh_vars.f90:62-70, src/flexi.h:98-115, src/mesh/mappings.f90:75-235)
 */
#ifndef GALX_ORACLE_H
#define GALX_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* node types for orc_nodes_and_weights (src/interpolation/interpolation_vars.f90:40-44) */
enum { ORC_NODE_G = 1, ORC_NODE_GL = 2, ORC_NODE_CL = 3, ORC_NODE_VISU = 4 };
/* Riemann selectors, numbering as the reference's -DRIEMANN= (src/CMakeLists.txt:100-107) */
enum { ORC_RIEMANN_LF = 0, ORC_RIEMANN_ROE = 1, ORC_RIEMANN_ROEEF = 3 };

typedef struct orc_ctx orc_ctx;

/* basis (src/interpolation/basis.f90, src/interpolation/interpolation.f90:199-318) */
int orc_nodes_and_weights(int N, int nodeType, double *x, double *w, double *wBary);
int orc_derivative_matrix(int N, int nodeType, double *D);      /* D(0:N,0:N), column-major */
/* D, D_Hat_T, L_HatMinus, L_HatPlus, DVolSurf for GL nodes (src/dg/dg.f90:181-242) */
int orc_dg_basis(int N, double *xGP, double *wGP, double *D, double *D_Hat_T,
                 double *L_HatMinus, double *L_HatPlus, double *DVolSurf);

/* host twins used by the unit tests, all sides 1..nSides (src/interpolation/prolongtoface.t90:20-110,
 * src/dg/surfint.t90:19-232); GL nodes, non-EXACT_MM */
int orc_prolong_to_face(int nVar, int N, int nElems, int nSides, const int *SideToElem, const int *S2V2,
                        const double *Uvol, double *Uface_master, double *Uface_slave);
int orc_surfint(int nVar, int N, int nElems, int nSides, const int *SideToElem, const int *S2V2,
                const double *Flux_master, const double *Flux_slave, double L_HatMinus0, double L_HatPlusN,
                double *Ut);

/* operator context */
orc_ctx *orc_create(int N, int nElems, int nSides, int splitForm, int riemann,
                    double kappa, double R, double Pr, double mu0);
void orc_destroy(orc_ctx *c);
int orc_set_mesh(orc_ctx *c, const int *ElemToSide, const int *SideToElem, const int *S2V2,
                 const double *Mf, const double *Mg, const double *Mh, const double *sJ,
                 const double *NormVec, const double *TangVec1, const double *TangVec2, const double *SurfElem);
/* DGTimeDerivative_weakForm (src/dg/dg.f90:255-425) */
int orc_rhs(orc_ctx *c, const double *U, double *Ut);
/* the same at time t, which enters only through the manufactured source */
int orc_rhs_t(orc_ctx *c, const double *U, double *Ut, double t);
/* CalcSource of exact function 4, 41, 42 or 43 (other values: none) with Elem_xGP(3,0:N,0:N,0:N,nElems) */
int orc_set_source(orc_ctx *c, int exactFunc, const double *advVel, const double *Elem_xGP);
/* the physical-space source alone, S(5,0:N,0:N,0:N,nElems) at time t */
int orc_source(orc_ctx *c, double t, double *S);
/* access to the intermediate arrays of the last orc_rhs call (reference layouts); which:
 * 0 U_master 1 U_slave 2 UPrim 3 gradUx 4 gradUy 5 gradUz 6 gradUx_master 7 gradUx_slave
 * 8 Flux_master 9 gradUy_master 10 gradUz_master 11 gradUy_slave 12 gradUz_slave */
const double *orc_array(orc_ctx *c, int which, long *count);
/* TimeStepByLSERKW2 with Carpenter RK4-5 (src/timedisc/timestep.f90:49-120) */
int orc_lserk_step(orc_ctx *c, double *U, double *Ut_tmp, double t, double dt);
/* CALCTIMESTEP with CFL/DFL scaling (src/equations/navierstokes/calctimestep.f90:51-296,
 * src/timedisc/timedisc_func.f90:480-550); cfl/dfl are the ini values CFLScale/DFLScale */
int orc_calc_dt(orc_ctx *c, const double *U, double cfl, double dfl, double *dt, int *errType,
                double *dtConv, double *dtVisc);
void orc_lserk_coefs(double *RKA, double *RKb, double *RKc);   /* 5 entries each, RKA[0]=RKc[0]=0 */
int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
