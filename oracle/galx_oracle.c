/*
 * galx_oracle.c -- FP64 CPU restatement of GALAEXI's DGTimeDerivative_weakForm + TimeStepByLSERKW2.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, the smoke entry point and the
 * cpu_baseline leg of the benchmark; never by the product.
 *
 * Each routine cites the reference lines it restates.  The reference is CUDA Fortran compiled
 * with -r8; this file follows the device-kernel semantics (what GALAEXI executes) and keeps the
 * reference's evaluation order where it is cheap to do so.
 */
#include "galx_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------------------------------ */
/* basis: src/interpolation/basis.f90                                                          */
/* ------------------------------------------------------------------------------------------ */

/* LegendrePolynomialAndDerivative (src/interpolation/basis.f90:195-234): L2-normalised, i.e. scaled by
 * sqrt(N+1/2) */
static void legendre_and_deriv(int N, double x, double *L, double *Lder)
{
    double l, ld;
    if (N == 0) { l = 1.0; ld = 0.0; }
    else if (N == 1) { l = x; ld = 1.0; }
    else {
        double Lm2 = 1.0, Lm1 = x, Ldm1 = 1.0;
        l = 0.0; ld = 0.0;
        for (int i = 2; i <= N; i++) {
            l = ((double)(2 * i - 1) * x * Lm1 - (double)(i - 1) * Lm2) / (double)i;
            ld = (double)i * Lm1 + x * Ldm1;
            Lm2 = Lm1; Lm1 = l; Ldm1 = ld;
        }
    }
    *L = l * sqrt((double)N + 0.5);
    *Lder = ld * sqrt((double)N + 0.5);
}

/* qAndLEvaluation (src/interpolation/basis.f90, ELEMENTAL SUBROUTINE qAndLEvaluation) */
static void q_and_L(int N, double x, double *q, double *qder, double *L)
{
    double Lm2 = 1.0, Lm1 = x, l = x;
    for (int i = 2; i <= N; i++) {
        l = ((double)(2 * i - 1) * x * Lm1 - (double)(i - 1) * Lm2) / (double)i;
        Lm2 = Lm1; Lm1 = l;
    }
    *q = (double)(2 * N + 1) / (double)(N + 1) * (x * l - Lm2);
    *qder = (double)(2 * N + 1) * l;
    *L = l;
}

/* LegGaussLobNodesAndWeights (src/interpolation/basis.f90:567-640) */
static void lgl_nodes(int N, double *x, double *w)
{
    x[0] = -1.0; x[N] = 1.0;
    w[0] = 2.0 / (double)(N * (N + 1)); w[N] = w[0];
    if (N > 1) {
        double c1 = PI / (double)N, c2 = 3.0 / ((double)(8 * N) * PI);
        for (int i = 1; i <= (N + 1) / 2 - 1; i++) {
            x[i] = -cos(c1 * ((double)i + 0.25) - c2 / ((double)i + 0.25));
            double q, qd, L;
            for (int it = 0; it <= 10; it++) {
                q_and_L(N, x[i], &q, &qd, &L);
                double dx = -q / qd;
                x[i] += dx;
                if (fabs(dx) < 1e-15 * fabs(x[i])) break;
            }
            q_and_L(N, x[i], &q, &qd, &L);
            x[N - i] = -x[i];
            w[i] = w[0] / (L * L);
            w[N - i] = w[i];
        }
    }
    if (N % 2 == 0) {
        double q, qd, L;
        x[N / 2] = 0.0;
        q_and_L(N, 0.0, &q, &qd, &L);
        w[N / 2] = w[0] / (L * L);
    }
}

/* LegendreGaussNodesAndWeights (src/interpolation/basis.f90:412-470) */
static void lg_nodes(int N, double *x, double *w)
{
    if (N == 0) { x[0] = 0.0; w[0] = 2.0; return; }
    if (N == 1) { x[0] = -sqrt(1.0 / 3.0); x[1] = -x[0]; w[0] = w[1] = 1.0; return; }
    double ct = 2.0 * atan(1.0) / (double)(N + 1);
    for (int i = 0; i <= (N + 1) / 2 - 1; i++) {
        x[i] = -cos(ct * (double)(2 * i + 1));
        double L, Ld;
        for (int it = 0; it <= 10; it++) {
            legendre_and_deriv(N + 1, x[i], &L, &Ld);
            double dx = -L / Ld;
            x[i] += dx;
            if (fabs(dx) < 1e-15 * fabs(x[i])) break;
        }
        legendre_and_deriv(N + 1, x[i], &L, &Ld);
        x[N - i] = -x[i];
        w[i] = (2.0 * N + 3) / ((1.0 - x[i] * x[i]) * Ld * Ld);
        w[N - i] = w[i];
    }
    if (N % 2 == 0) {
        double L, Ld;
        x[N / 2] = 0.0;
        legendre_and_deriv(N + 1, 0.0, &L, &Ld);
        w[N / 2] = (2.0 * N + 3) / (Ld * Ld);
    }
}

/* BarycentricWeights (src/interpolation/basis.f90, PPURE SUBROUTINE BarycentricWeights) */
static void bary_weights(int N, const double *x, double *wb)
{
    for (int i = 0; i <= N; i++) wb[i] = 1.0;
    for (int i = 1; i <= N; i++)
        for (int j = 0; j <= i - 1; j++) {
            wb[j] *= (x[j] - x[i]);
            wb[i] *= (x[i] - x[j]);
        }
    for (int i = 0; i <= N; i++) wb[i] = 1.0 / wb[i];
}

/* GetNodesAndWeights (src/interpolation/interpolation.f90:199-318) */
int orc_nodes_and_weights(int N, int type, double *x, double *w, double *wb)
{
    double wl[64];
    if (N < 0 || N > 60) return 1;
    if (!w) w = wl;
    switch (type) {
    case ORC_NODE_G: lg_nodes(N, x, w); break;
    case ORC_NODE_GL: lgl_nodes(N, x, w); break;
    case ORC_NODE_CL: /* ChebyGaussLobNodesAndWeights (src/interpolation/basis.f90:269-291) */
        for (int i = 0; i <= N; i++) x[i] = -cos((double)i / (double)N * PI);
        for (int i = 0; i <= N; i++) w[i] = PI / (double)N;
        w[0] *= 0.5; w[N] *= 0.5;
        break;
    case ORC_NODE_VISU:
        for (int i = 0; i <= N; i++) x[i] = 2.0 * (double)i / (double)N - 1.0;
        for (int i = 0; i <= N; i++) w[i] = 2.0 / (double)N;
        w[0] *= 0.5; w[N] *= 0.5;
        break;
    default: return 2;
    }
    if (wb) bary_weights(N, x, wb);
    return 0;
}

/* PolynomialDerivativeMatrix (src/interpolation/basis.f90:664-690) */
static void deriv_matrix(int N, const double *x, double *D)
{
    int n = N + 1;
    double wb[64];
    bary_weights(N, x, wb);
    for (int i = 0; i < n * n; i++) D[i] = 0.0;
    for (int l = 0; l <= N; l++)
        for (int i = 0; i <= N; i++)
            if (l != i) {
                D[i + n * l] = wb[l] / (wb[i] * (x[i] - x[l]));
                D[i + n * i] -= D[i + n * l];
            }
}

int orc_derivative_matrix(int N, int type, double *D)
{
    double x[64], w[64];
    int rc = orc_nodes_and_weights(N, type, x, w, NULL);
    if (rc) return rc;
    deriv_matrix(N, x, D);
    return 0;
}

/* InitDGbasis (src/dg/dg.f90:181-242) for GL nodes: D_Hat = -M^{-1} D^T M with diagonal M,
 * L_Hat = M^{-1} L, DVolSurf = D^T with the two corner corrections. */
int orc_dg_basis(int N, double *xGP, double *wGP, double *D, double *D_Hat_T,
                 double *L_HatMinus, double *L_HatPlus, double *DVolSurf)
{
    int n = N + 1;
    lgl_nodes(N, xGP, wGP);
    deriv_matrix(N, xGP, D);
    /* D_Hat(i,j) = -(1/w_i) * D(j,i) * w_j ; D_Hat_T(i,j) = D_Hat(j,i) */
    for (int i = 0; i <= N; i++)
        for (int j = 0; j <= N; j++) {
            double dh = -((1.0 / wGP[i]) * (D[j + n * i] * wGP[j]));
            D_Hat_T[j + n * i] = dh;
        }
    for (int i = 0; i <= N; i++)
        for (int j = 0; j <= N; j++) DVolSurf[i + n * j] = D[j + n * i];
    DVolSurf[0] = DVolSurf[0] + 1.0 / (2.0 * wGP[0]);
    DVolSurf[N + n * N] = DVolSurf[N + n * N] - 1.0 / (2.0 * wGP[N]);
    /* GL: L_Minus = e_0, L_Plus = e_N, L_Hat = M^{-1} L */
    for (int i = 0; i <= N; i++) {
        L_HatMinus[i] = (i == 0) ? 1.0 / wGP[0] : 0.0;
        L_HatPlus[i] = (i == N) ? 1.0 / wGP[N] : 0.0;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* mesh index helpers (src/flexi.h:98-115, src/mesh/mappings.f90)                              */
/* ------------------------------------------------------------------------------------------ */
enum { XI_MINUS = 5, XI_PLUS = 3, ETA_MINUS = 2, ETA_PLUS = 4, ZETA_MINUS = 1, ZETA_PLUS = 6 };

#define S2V2IDX(c, p, q, f, s, n) ((c) + 2 * ((p) + (n) * ((q) + (n) * ((f) + 5 * ((s) - 1)))))

/* volume node index (i,j,k) of face node (a,b) in the volume frame of locSide (GL trace) */
static inline void face_to_vol(int locSide, int a, int b, int N, int *i, int *j, int *k)
{
    switch (locSide) {
    case XI_MINUS: *i = 0; *j = a; *k = b; break;
    case XI_PLUS: *i = N; *j = a; *k = b; break;
    case ETA_MINUS: *i = a; *j = 0; *k = b; break;
    case ETA_PLUS: *i = a; *j = N; *k = b; break;
    case ZETA_MINUS: *i = a; *j = b; *k = 0; break;
    default: *i = a; *j = b; *k = N; break;
    }
}

/* ProlongToFace, host twin for GL (src/interpolation/prolongtoface.t90:20-110, EvalElemFaceGL) */
int orc_prolong_to_face(int nVar, int N, int nElems, int nSides, const int *SideToElem, const int *S2V2,
                        const double *Uvol, double *Um, double *Us)
{
    int n = N + 1;
    (void)nElems;
    for (int s = 0; s < nSides; s++) {
        int e = SideToElem[0 + 5 * s], nb = SideToElem[1 + 5 * s];
        if (e > 0) {
            int ls = SideToElem[2 + 5 * s];
            for (int q = 0; q <= N; q++)
                for (int p = 0; p <= N; p++) {
                    int a = S2V2[S2V2IDX(0, p, q, 0, ls, n)], b = S2V2[S2V2IDX(1, p, q, 0, ls, n)];
                    int i, j, k;
                    face_to_vol(ls, a, b, N, &i, &j, &k);
                    for (int v = 0; v < nVar; v++)
                        Um[v + nVar * (p + n * (q + n * (long)s))] =
                            Uvol[v + nVar * (i + n * (j + n * (k + n * (long)(e - 1))))];
                }
        }
        if (nb > 0) {
            int ls = SideToElem[3 + 5 * s], flip = SideToElem[4 + 5 * s];
            for (int q = 0; q <= N; q++)
                for (int p = 0; p <= N; p++) {
                    int a = S2V2[S2V2IDX(0, p, q, flip, ls, n)], b = S2V2[S2V2IDX(1, p, q, flip, ls, n)];
                    int i, j, k;
                    face_to_vol(ls, a, b, N, &i, &j, &k);
                    for (int v = 0; v < nVar; v++)
                        Us[v + nVar * (p + n * (q + n * (long)s))] =
                            Uvol[v + nVar * (i + n * (j + n * (k + n * (long)(nb - 1))))];
                }
        }
    }
    return 0;
}

/* SurfInt host twin for GL non-EXACT_MM (src/dg/surfint.t90:19-113,163-232) */
int orc_surfint(int nVar, int N, int nElems, int nSides, const int *SideToElem, const int *S2V2,
                const double *Fm, const double *Fs, double LHm, double LHp, double *Ut)
{
    int n = N + 1;
    (void)nElems;
    for (int s = 0; s < nSides; s++) {
        for (int side = 0; side < 2; side++) {
            int e = SideToElem[side + 5 * s];
            if (e <= 0) continue;
            int ls = SideToElem[(side ? 3 : 2) + 5 * s];
            int flip = side ? SideToElem[4 + 5 * s] : 0;
            const double *F = side ? Fs : Fm;
            double sgn = side ? -1.0 : 1.0;
            double LH = (ls == XI_MINUS || ls == ETA_MINUS || ls == ZETA_MINUS) ? LHm : LHp;
            for (int q = 0; q <= N; q++)
                for (int p = 0; p <= N; p++) {
                    int a = S2V2[S2V2IDX(0, p, q, flip, ls, n)], b = S2V2[S2V2IDX(1, p, q, flip, ls, n)];
                    int i, j, k;
                    face_to_vol(ls, a, b, N, &i, &j, &k);
                    for (int v = 0; v < nVar; v++) {
                        double ft = sgn * F[v + nVar * (p + n * (q + n * (long)s))];
                        double *u = &Ut[v + nVar * (i + n * (j + n * (k + n * (long)(e - 1))))];
                        *u = *u + ft * LH;
                    }
                }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* operator context                                                                            */
/* ------------------------------------------------------------------------------------------ */
struct orc_ctx {
    int N, n, nElems, nSides, split, riemann;
    double kappa, R, Pr, mu0, lambda;
    double xGP[64], wGP[64], D[64 * 64], D_T[64 * 64], D_Hat_T[64 * 64], L_HatMinus[64], L_HatPlus[64],
        DVolSurf[64 * 64];
    int *ElemToSide, *SideToElem, *S2V2, *S2V2_inv;
    double *Mf, *Mg, *Mh, *sJ, *NormVec, *TangVec1, *TangVec2, *SurfElem;
    /* work */
    double *UPrim, *U_master, *U_slave, *UPrim_master, *UPrim_slave;
    double *gradUx, *gradUy, *gradUz, *gxm, *gym, *gzm, *gxs, *gys, *gzs, *liftFlux;
    double *Flux_master, *Flux_slave, *f, *g, *h, *Ut_stage;
    /* manufactured source: exact function (0: none), AdvVel, node coordinates Elem_xGP(3,nDOF) */
    int src_func;
    double adv[3];
    double *Elem_xGP;
};

static void *xcalloc(long n, long sz) { return calloc((size_t)(n > 0 ? n : 1), (size_t)sz); }

orc_ctx *orc_create(int N, int nElems, int nSides, int split, int riemann, double kappa, double R,
                    double Pr, double mu0)
{
    if (N < 1 || N > 30 || nElems < 1 || nSides < 1) return NULL;
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    int n = N + 1;
    long nv = (long)n * n * n * nElems, nf = (long)n * n * nSides;
    c->N = N; c->n = n; c->nElems = nElems; c->nSides = nSides; c->split = split; c->riemann = riemann;
    c->kappa = kappa; c->R = R; c->Pr = Pr; c->mu0 = mu0;
    /* THERMAL_CONDUCTIVITY_EOS (src/equations/navierstokes/idealgas/eos.h:110) */
    c->lambda = mu0 * R * kappa / ((kappa - 1.0) * Pr);
    orc_dg_basis(N, c->xGP, c->wGP, c->D, c->D_Hat_T, c->L_HatMinus, c->L_HatPlus, c->DVolSurf);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) c->D_T[i + n * j] = c->D[j + n * i];
    c->ElemToSide = (int *)xcalloc(18L * nElems, sizeof(int));
    c->SideToElem = (int *)xcalloc(5L * nSides, sizeof(int));
    c->S2V2 = (int *)xcalloc(2L * n * n * 30, sizeof(int));
    c->S2V2_inv = (int *)xcalloc(2L * n * n * 30, sizeof(int));
    c->Mf = (double *)xcalloc(3 * nv, 8); c->Mg = (double *)xcalloc(3 * nv, 8); c->Mh = (double *)xcalloc(3 * nv, 8);
    c->sJ = (double *)xcalloc(nv, 8);
    c->NormVec = (double *)xcalloc(3 * nf, 8); c->TangVec1 = (double *)xcalloc(3 * nf, 8);
    c->TangVec2 = (double *)xcalloc(3 * nf, 8); c->SurfElem = (double *)xcalloc(nf, 8);
    c->UPrim = (double *)xcalloc(6 * nv, 8);
    c->U_master = (double *)xcalloc(5 * nf, 8); c->U_slave = (double *)xcalloc(5 * nf, 8);
    c->UPrim_master = (double *)xcalloc(6 * nf, 8); c->UPrim_slave = (double *)xcalloc(6 * nf, 8);
    c->gradUx = (double *)xcalloc(5 * nv, 8); c->gradUy = (double *)xcalloc(5 * nv, 8); c->gradUz = (double *)xcalloc(5 * nv, 8);
    c->gxm = (double *)xcalloc(5 * nf, 8); c->gym = (double *)xcalloc(5 * nf, 8); c->gzm = (double *)xcalloc(5 * nf, 8);
    c->gxs = (double *)xcalloc(5 * nf, 8); c->gys = (double *)xcalloc(5 * nf, 8); c->gzs = (double *)xcalloc(5 * nf, 8);
    c->liftFlux = (double *)xcalloc(5 * nf, 8);
    c->Flux_master = (double *)xcalloc(5 * nf, 8); c->Flux_slave = (double *)xcalloc(5 * nf, 8);
    c->f = (double *)xcalloc(5 * nv, 8); c->g = (double *)xcalloc(5 * nv, 8); c->h = (double *)xcalloc(5 * nv, 8);
    c->Ut_stage = (double *)xcalloc(5 * nv, 8);
    return c;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    void *p[] = {c->ElemToSide, c->SideToElem, c->S2V2, c->S2V2_inv, c->Mf, c->Mg, c->Mh, c->sJ, c->NormVec,
                 c->TangVec1, c->TangVec2, c->SurfElem, c->UPrim, c->U_master, c->U_slave, c->UPrim_master,
                 c->UPrim_slave, c->gradUx, c->gradUy, c->gradUz, c->gxm, c->gym, c->gzm, c->gxs, c->gys,
                 c->gzs, c->liftFlux, c->Flux_master, c->Flux_slave, c->f, c->g, c->h, c->Ut_stage,
                 c->Elem_xGP};
    for (unsigned i = 0; i < sizeof(p) / sizeof(p[0]); i++) free(p[i]);
    free(c);
}

int orc_set_mesh(orc_ctx *c, const int *E2S, const int *S2E, const int *S2V2, const double *Mf,
                 const double *Mg, const double *Mh, const double *sJ, const double *nv, const double *t1,
                 const double *t2, const double *se)
{
    int n = c->n;
    long nvol = (long)n * n * n * c->nElems, nf = (long)n * n * c->nSides;
    memcpy(c->ElemToSide, E2S, sizeof(int) * 18L * c->nElems);
    memcpy(c->SideToElem, S2E, sizeof(int) * 5L * c->nSides);
    memcpy(c->S2V2, S2V2, sizeof(int) * 2L * n * n * 30);
    /* S2V2_inv (src/mesh/mappings.f90:154-166) */
    for (int ls = 1; ls <= 6; ls++)
        for (int f = 0; f <= 4; f++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    int a = S2V2[S2V2IDX(0, i, j, f, ls, n)], b = S2V2[S2V2IDX(1, i, j, f, ls, n)];
                    c->S2V2_inv[S2V2IDX(0, a, b, f, ls, n)] = i;
                    c->S2V2_inv[S2V2IDX(1, a, b, f, ls, n)] = j;
                }
    memcpy(c->Mf, Mf, 24 * nvol); memcpy(c->Mg, Mg, 24 * nvol); memcpy(c->Mh, Mh, 24 * nvol);
    memcpy(c->sJ, sJ, 8 * nvol);
    memcpy(c->NormVec, nv, 24 * nf); memcpy(c->TangVec1, t1, 24 * nf); memcpy(c->TangVec2, t2, 24 * nf);
    memcpy(c->SurfElem, se, 8 * nf);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* physics (src/equations/navierstokes/...)                                                    */
/* ------------------------------------------------------------------------------------------ */

/* ConsToPrim (src/equations/navierstokes/idealgas/eos.f90:212-239) */
static inline void cons_to_prim(const double *cons, double *prim, double kappa, double R)
{
    double sRho = 1.0 / cons[0];
    prim[0] = cons[0];
    prim[1] = cons[1] * sRho;
    prim[2] = cons[2] * sRho;
    prim[3] = cons[3] * sRho;
    prim[4] = (kappa - 1.0) * (cons[4] - 0.5 * (cons[1] * prim[1] + cons[2] * prim[2] + cons[3] * prim[3]));
    prim[5] = prim[4] * sRho / R;
}

/* extended state indices (src/equations/navierstokes/idealgas/eos.h:119-135), 0-based:
 * 0 DENS 1..3 MOM 4 ENER 5 SRHO 6..8 VEL 9 PRES */
enum { X_DENS = 0, X_MOM1 = 1, X_ENER = 4, X_SRHO = 5, X_VEL1 = 6, X_PRES = 9 };

/* EvalEulerFlux1D_fast (src/equations/navierstokes/flux.f90:1123-1143) */
static inline void euler_flux_1d(const double *U, double *F)
{
    F[0] = U[X_MOM1];
    F[1] = U[X_MOM1] * U[X_VEL1] + U[X_PRES];
    F[2] = U[X_MOM1] * U[X_VEL1 + 1];
    F[3] = U[X_MOM1] * U[X_VEL1 + 2];
    F[4] = (U[X_ENER] + U[X_PRES]) * U[X_VEL1];
}

/* SplitSurfaceFluxPI (src/equations/navierstokes/splitflux.f90:735-761) */
static inline void split_surface_flux_pi(const double *L, const double *R, double *F)
{
    double HL = (L[X_ENER] + L[X_PRES]) / L[X_DENS];
    double HR = (R[X_ENER] + R[X_PRES]) / R[X_DENS];
    F[0] = 0.25 * (L[X_DENS] + R[X_DENS]) * (L[X_VEL1] + R[X_VEL1]);
    F[1] = 0.125 * (L[X_DENS] + R[X_DENS]) * ((L[X_VEL1] + R[X_VEL1]) * (L[X_VEL1] + R[X_VEL1])) +
           0.5 * (L[X_PRES] + R[X_PRES]);
    F[2] = 0.125 * (L[X_DENS] + R[X_DENS]) * (L[X_VEL1] + R[X_VEL1]) * (L[X_VEL1 + 1] + R[X_VEL1 + 1]);
    F[3] = 0.125 * (L[X_DENS] + R[X_DENS]) * (L[X_VEL1] + R[X_VEL1]) * (L[X_VEL1 + 2] + R[X_VEL1 + 2]);
    F[4] = 0.125 * (L[X_DENS] + R[X_DENS]) * (HL + HR) * (L[X_VEL1] + R[X_VEL1]);
}

/* Riemann_LF / Riemann_Roe / Riemann_RoeEntropyFix (src/equations/navierstokes/riemann.f90:707-989),
 * split or not (SPLIT_DG branch uses the PI surface flux as the central part). */
static void riemann_solver(const orc_ctx *c, const double *L, const double *R, const double *FL,
                           const double *FR, double *F)
{
    double kappa = c->kappa;
    if (c->riemann == ORC_RIEMANN_LF) {
        double cL = sqrt(kappa * L[X_PRES] * L[X_SRHO]), cR = sqrt(kappa * R[X_PRES] * R[X_SRHO]);
        double lmax = fmax(fabs(R[X_VEL1]), fabs(L[X_VEL1])) + fmax(cL, cR);
        if (!c->split) {
            for (int v = 0; v < 5; v++) F[v] = 0.5 * ((FL[v] + FR[v]) - lmax * (R[v] - L[v]));
        } else {
            split_surface_flux_pi(L, R, F);
            for (int v = 0; v < 5; v++) F[v] = F[v] - 0.5 * lmax * (R[v] - L[v]);
        }
        return;
    }
    double HL = (L[X_ENER] + L[X_PRES]) * L[X_SRHO], HR = (R[X_ENER] + R[X_PRES]) * R[X_SRHO];
    double sqL = sqrt(L[X_DENS]), sqR = sqrt(R[X_DENS]);
    double ss = 1.0 / (sqL + sqR);
    double rv[3];
    for (int d = 0; d < 3; d++) rv[d] = (sqR * R[X_VEL1 + d] + sqL * L[X_VEL1 + d]) * ss;
    double absVel = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
    double RoeH = (sqR * HR + sqL * HL) * ss;
    /* ROEC_RIEMANN_H (eos.h:76) */
    double Roec = sqrt((kappa - 1.0) * (RoeH - 0.5 * absVel));
    double a[5] = {rv[0] - Roec, rv[0], rv[0], rv[0], rv[0] + Roec};
    double r1[5] = {1.0, a[0], rv[1], rv[2], RoeH - rv[0] * Roec};
    double r2[5] = {1.0, rv[0], rv[1], rv[2], 0.5 * absVel};
    double r3[5] = {0.0, 0.0, 1.0, 0.0, rv[1]};
    double r4[5] = {0.0, 0.0, 0.0, 1.0, rv[2]};
    double r5[5] = {1.0, a[4], rv[1], rv[2], RoeH + rv[0] * Roec};
    double al[5], aa[5];
    if (c->riemann == ORC_RIEMANN_ROE) {
        double dU[6];
        for (int v = 0; v < 5; v++) dU[v] = R[v] - L[v];
        dU[5] = dU[4] - (dU[2] - rv[1] * dU[0]) * rv[1] - (dU[3] - rv[2] * dU[0]) * rv[2];
        al[2] = dU[2] - rv[1] * dU[0];
        al[3] = dU[3] - rv[2] * dU[0];
        /* ALPHA2_RIEMANN_H (eos.h:77) */
        al[1] = (kappa - 1.0) / (Roec * Roec) * (dU[0] * (RoeH - rv[0] * rv[0]) - dU[5] + rv[0] * dU[1]);
        al[0] = 0.5 / Roec * (dU[0] * (rv[0] + Roec) - dU[1] - Roec * al[1]);
        al[4] = dU[0] - al[0] - al[1];
        for (int v = 0; v < 5; v++) aa[v] = fabs(a[v]);
    } else {
        double cL = sqrt(kappa * L[X_PRES] * L[X_SRHO]), cR = sqrt(kappa * R[X_PRES] * R[X_SRHO]);
        double RoeDens = sqrt(L[X_DENS] * R[X_DENS]);
        double dU[5] = {R[X_DENS] - L[X_DENS], R[X_VEL1] - L[X_VEL1], R[X_VEL1 + 1] - L[X_VEL1 + 1],
                        R[X_VEL1 + 2] - L[X_VEL1 + 2], R[X_PRES] - L[X_PRES]};
        double tmp = 0.5 / (Roec * Roec);
        al[0] = tmp * (dU[4] - RoeDens * Roec * dU[1]);
        al[1] = dU[0] - dU[4] * 2.0 * tmp;
        al[2] = RoeDens * dU[2];
        al[3] = RoeDens * dU[3];
        al[4] = tmp * (dU[4] + RoeDens * Roec * dU[1]);
        double aL[5] = {L[X_VEL1] - cL, L[X_VEL1], L[X_VEL1], L[X_VEL1], L[X_VEL1] + cL};
        double aR[5] = {R[X_VEL1] - cR, R[X_VEL1], R[X_VEL1], R[X_VEL1], R[X_VEL1] + cR};
        for (int v = 0; v < 5; v++) {
            double da = fmax(0.0, fmax(a[v] - aL[v], aR[v] - a[v]));
            if (fabs(a[v]) < da) aa[v] = 0.5 * (a[v] * a[v] / da + da);
            else aa[v] = fabs(a[v]);
        }
    }
    if (!c->split) {
        for (int v = 0; v < 5; v++)
            F[v] = 0.5 * ((FL[v] + FR[v]) - al[0] * aa[0] * r1[v] - al[1] * aa[1] * r2[v] -
                          al[2] * aa[2] * r3[v] - al[3] * aa[3] * r4[v] - al[4] * aa[4] * r5[v]);
    } else {
        split_surface_flux_pi(L, R, F);
        for (int v = 0; v < 5; v++)
            F[v] = F[v] - 0.5 * (al[0] * aa[0] * r1[v] + al[1] * aa[1] * r2[v] + al[2] * aa[2] * r3[v] +
                                 al[3] * aa[3] * r4[v] + al[4] * aa[4] * r5[v]);
    }
}

/* Riemann driver: rotate into normal frame, solve, rotate back (src/equations/navierstokes/riemann.f90:216-289) */
static void riemann_point(const orc_ctx *c, const double *UL, const double *UR, const double *PL,
                          const double *PR, const double *nv, const double *t1, const double *t2, double *Fout)
{
    double L[10], R[10], FL[5] = {0}, FR[5] = {0}, F[5];
    L[X_DENS] = UL[0]; L[X_SRHO] = 1.0 / L[X_DENS]; L[X_ENER] = UL[4]; L[X_PRES] = PL[4];
    L[X_VEL1] = PL[1] * nv[0] + PL[2] * nv[1] + PL[3] * nv[2];
    L[X_VEL1 + 1] = PL[1] * t1[0] + PL[2] * t1[1] + PL[3] * t1[2];
    L[X_VEL1 + 2] = PL[1] * t2[0] + PL[2] * t2[1] + PL[3] * t2[2];
    L[X_MOM1] = L[X_DENS] * L[X_VEL1]; L[X_MOM1 + 1] = L[X_DENS] * L[X_VEL1 + 1]; L[X_MOM1 + 2] = L[X_DENS] * L[X_VEL1 + 2];
    R[X_DENS] = UR[0]; R[X_SRHO] = 1.0 / R[X_DENS]; R[X_ENER] = UR[4]; R[X_PRES] = PR[4];
    R[X_VEL1] = PR[1] * nv[0] + PR[2] * nv[1] + PR[3] * nv[2];
    R[X_VEL1 + 1] = PR[1] * t1[0] + PR[2] * t1[1] + PR[3] * t1[2];
    R[X_VEL1 + 2] = PR[1] * t2[0] + PR[2] * t2[1] + PR[3] * t2[2];
    R[X_MOM1] = R[X_DENS] * R[X_VEL1]; R[X_MOM1 + 1] = R[X_DENS] * R[X_VEL1 + 1]; R[X_MOM1 + 2] = R[X_DENS] * R[X_VEL1 + 2];
    if (!c->split) { euler_flux_1d(L, FL); euler_flux_1d(R, FR); }
    riemann_solver(c, L, R, FL, FR, F);
    Fout[0] = F[0];
    for (int d = 0; d < 3; d++) Fout[1 + d] = nv[d] * F[1] + t1[d] * F[2] + t2[d] * F[3];
    Fout[4] = F[4];
}

/* EvalNormalDiffFlux3D (src/equations/navierstokes/flux.f90:779-847); lifting vars (rho,u,v,w,T) */
static inline void normal_diff_flux(const double *P, const double *gx, const double *gy, const double *gz,
                                    const double *nv, double mu, double lambda, double *f)
{
    const double s23 = 2.0 / 3.0, s43 = 4.0 / 3.0;
    double txx = mu * (s43 * gx[1] - s23 * gy[2] - s23 * gz[3]);
    double tyy = mu * (-s23 * gx[1] + s43 * gy[2] - s23 * gz[3]);
    double tzz = mu * (-s23 * gx[1] - s23 * gy[2] + s43 * gz[3]);
    double txy = mu * (gy[1] + gx[2]);
    double txz = mu * (gz[1] + gx[3]);
    double tyz = mu * (gz[2] + gy[3]);
    f[0] = 0.0;
    f[1] = -nv[0] * txx - nv[1] * txy - nv[2] * txz;
    f[2] = -nv[0] * txy - nv[1] * tyy - nv[2] * tyz;
    f[3] = -nv[0] * txz - nv[1] * tyz - nv[2] * tzz;
    f[4] = -nv[0] * (txx * P[1] + txy * P[2] + txz * P[3] + lambda * gx[4]) -
           nv[1] * (txy * P[1] + tyy * P[2] + tyz * P[3] + lambda * gy[4]) -
           nv[2] * (txz * P[1] + tyz * P[2] + tzz * P[3] + lambda * gz[4]);
}

/* EvalTransformedEulerDiffFlux3D (flux.f90:509-614) when withEuler, else EvalTransformedDiffFlux3D
 * (flux.f90:619-700) */
static void transformed_flux(int withEuler, const double *U, const double *P, const double *gx,
                             const double *gy, const double *gz, const double *Mf, const double *Mg,
                             const double *Mh, double mu, double lambda, double *f, double *g, double *h)
{
    const double s23 = 2.0 / 3.0, s43 = 4.0 / 3.0;
    double txx = mu * (s43 * gx[1] - s23 * gy[2] - s23 * gz[3]);
    double tyy = mu * (-s23 * gx[1] + s43 * gy[2] - s23 * gz[3]);
    double tzz = mu * (-s23 * gx[1] - s23 * gy[2] + s43 * gz[3]);
    double txy = mu * (gy[1] + gx[2]);
    double txz = mu * (gz[1] + gx[3]);
    double tyz = mu * (gz[2] + gy[3]);
    double qx = txx * P[1] + txy * P[2] + txz * P[3] + lambda * gx[4];
    double qy = txy * P[1] + tyy * P[2] + tyz * P[3] + lambda * gy[4];
    double qz = txz * P[1] + tyz * P[2] + tzz * P[3] + lambda * gz[4];
    const double *M[3] = {Mf, Mg, Mh};
    double *o[3] = {f, g, h};
    for (int d = 0; d < 3; d++) {
        const double *m = M[d];
        double *r = o[d];
        if (withEuler) {
            double Ep = (U[4] + P[4]) / U[0];
            double Mmom = m[0] * U[1] + m[1] * U[2] + m[2] * U[3];
            r[0] = Mmom;
            r[1] = Mmom * P[1] + m[0] * P[4] - m[0] * txx - m[1] * txy - m[2] * txz;
            r[2] = Mmom * P[2] + m[1] * P[4] - m[0] * txy - m[1] * tyy - m[2] * tyz;
            r[3] = Mmom * P[3] + m[2] * P[4] - m[0] * txz - m[1] * tyz - m[2] * tzz;
            r[4] = Mmom * Ep - m[0] * qx - m[1] * qy - m[2] * qz;
        } else {
            r[0] = 0.0;
            r[1] = -m[0] * txx - m[1] * txy - m[2] * txz;
            r[2] = -m[0] * txy - m[1] * tyy - m[2] * tyz;
            r[3] = -m[0] * txz - m[1] * tyz - m[2] * tzz;
            r[4] = -m[0] * qx - m[1] * qy - m[2] * qz;
        }
    }
}

/* SplitVolumeFluxPI (src/equations/navierstokes/splitflux.f90:669-730) */
static inline void split_volume_flux_pi(const double *UR, const double *PR, const double *U, const double *P,
                                        const double *MR, const double *M, double *Flux)
{
    double HRef = (UR[4] + PR[4]) / UR[0];
    double H = (U[4] + P[4]) / U[0];
    double sr = UR[0] + U[0], su = PR[1] + P[1], sv = PR[2] + P[2], sw = PR[3] + P[3], sp = PR[4] + P[4];
    double ft[5], gt[5], ht[5];
    ft[0] = 0.5 * sr * su;
    ft[1] = 0.25 * sr * (su * su) + sp;
    ft[2] = 0.25 * sr * su * sv;
    ft[3] = 0.25 * sr * su * sw;
    ft[4] = 0.25 * sr * su * (HRef + H);
    gt[0] = 0.5 * sr * sv;
    gt[1] = ft[2];
    gt[2] = 0.25 * sr * (sv * sv) + sp;
    gt[3] = 0.25 * sr * sv * sw;
    gt[4] = 0.25 * sr * sv * (HRef + H);
    ht[0] = 0.5 * sr * sw;
    ht[1] = ft[3];
    ht[2] = gt[3];
    ht[3] = 0.25 * sr * (sw * sw) + sp;
    ht[4] = 0.25 * sr * sw * (HRef + H);
    for (int v = 0; v < 5; v++)
        Flux[v] = 0.5 * (MR[0] + M[0]) * ft[v] + 0.5 * (MR[2] + M[2]) * ht[v] + 0.5 * (MR[1] + M[1]) * gt[v];
}

/* ------------------------------------------------------------------------------------------ */
/* DG operator pieces                                                                           */
/* ------------------------------------------------------------------------------------------ */
#define VIDX(nv, v, i, j, k, e, n) ((v) + (nv) * ((i) + (n) * ((j) + (n) * ((k) + (n) * (long)(e)))))
#define FIDX(nv, v, p, q, s, n) ((v) + (nv) * ((p) + (n) * ((q) + (n) * (long)(s))))
static const int LIFT2PRIM[5] = {0, 1, 2, 3, 5}; /* PRIM_LIFT (eos.h:138) */

/* EvalElemFaceGL_Device semantics for all elements/local sides (prolongtoface.t90:268-344) */
static void prolong_gl(const orc_ctx *c, int nVar, const double *Uvol, double *Um, double *Us)
{
    int N = c->N, n = c->n;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < c->nElems; e++)
        for (int ls = 1; ls <= 6; ls++) {
            int side = c->ElemToSide[0 + 3 * ((ls - 1) + 6 * e)] - 1;
            int flip = c->ElemToSide[1 + 3 * ((ls - 1) + 6 * e)];
            double *dst = flip == 0 ? Um : Us;
            for (int p = 0; p <= N; p++)
                for (int q = 0; q <= N; q++) {
                    int a = c->S2V2[S2V2IDX(0, p, q, flip, ls, n)], b = c->S2V2[S2V2IDX(1, p, q, flip, ls, n)];
                    int i, j, k;
                    face_to_vol(ls, a, b, N, &i, &j, &k);
                    for (int v = 0; v < nVar; v++)
                        dst[FIDX(nVar, v, p, q, side, n)] = Uvol[VIDX(nVar, v, i, j, k, e, n)];
                }
        }
}

/* SurfInt_Device / SurfInt_Device_Single_Flux (surfint.t90:352-725), GL non-EXACT_MM.
 * single!=0: one flux array, sign +1 for both sides (strong lifting).  applyJac: Ut = sJ*(Ut+loc). */
static void surfint_device(const orc_ctx *c, int nVar, const double *Fm, const double *Fs, int single,
                           int applyJac, double *Ut)
{
    int N = c->N, n = c->n;
    double LHm = c->L_HatMinus[0], LHp = c->L_HatPlus[N];
#pragma omp parallel for schedule(static)
    for (int e = 0; e < c->nElems; e++)
        for (int k = 0; k <= N; k++)
            for (int j = 0; j <= N; j++)
                for (int i = 0; i <= N; i++) {
                    double loc[5] = {0, 0, 0, 0, 0};
                    for (int ls = 1; ls <= 6; ls++) {
                        int side = c->ElemToSide[0 + 3 * ((ls - 1) + 6 * e)] - 1;
                        int flip = c->ElemToSide[1 + 3 * ((ls - 1) + 6 * e)];
                        int a, b, on;
                        double LH;
                        switch (ls) {
                        case XI_MINUS: on = (i == 0); a = j; b = k; LH = LHm; break;
                        case ETA_MINUS: on = (j == 0); a = i; b = k; LH = LHm; break;
                        case ZETA_MINUS: on = (k == 0); a = i; b = j; LH = LHm; break;
                        case XI_PLUS: on = (i == N); a = j; b = k; LH = LHp; break;
                        case ETA_PLUS: on = (j == N); a = i; b = k; LH = LHp; break;
                        default: on = (k == N); a = i; b = j; LH = LHp; break;
                        }
                        if (!on) continue;
                        int p = c->S2V2_inv[S2V2IDX(0, a, b, flip, ls, n)];
                        int q = c->S2V2_inv[S2V2IDX(1, a, b, flip, ls, n)];
                        if (single) {
                            for (int v = 0; v < nVar; v++) loc[v] = loc[v] + 1.0 * Fm[FIDX(nVar, v, p, q, side, n)] * LH;
                        } else if (flip == 0) {
                            for (int v = 0; v < nVar; v++) loc[v] = loc[v] + Fm[FIDX(nVar, v, p, q, side, n)] * LH;
                        } else {
                            for (int v = 0; v < nVar; v++) loc[v] = loc[v] - Fs[FIDX(nVar, v, p, q, side, n)] * LH;
                        }
                    }
                    double sj = c->sJ[i + n * (j + n * (k + n * (long)e))];
                    for (int v = 0; v < nVar; v++) {
                        double *u = &Ut[VIDX(nVar, v, i, j, k, e, n)];
                        *u = applyJac ? sj * (*u + loc[v]) : *u + loc[v];
                    }
                }
}

/* ApplyDMatrix_Kernel (src/dg/applydmatrix.t90:19-75) */
static void apply_dmatrix(const orc_ctx *c, const double *f, const double *g, const double *h, int overwrite,
                          double *Ut)
{
    int N = c->N, n = c->n;
    const double *DT = c->D_Hat_T;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < c->nElems; e++)
        for (int k = 0; k <= N; k++)
            for (int j = 0; j <= N; j++)
                for (int i = 0; i <= N; i++) {
                    double loc[5] = {0, 0, 0, 0, 0};
                    for (int l = 0; l <= N; l++)
                        for (int v = 0; v < 5; v++)
                            loc[v] = loc[v] + DT[l + n * i] * f[VIDX(5, v, l, j, k, e, n)] +
                                     DT[l + n * k] * h[VIDX(5, v, i, j, l, e, n)] +
                                     DT[l + n * j] * g[VIDX(5, v, i, l, k, e, n)];
                    for (int v = 0; v < 5; v++) {
                        double *u = &Ut[VIDX(5, v, i, j, k, e, n)];
                        *u = overwrite ? loc[v] : *u + loc[v];
                    }
                }
}

/* Lifting_BR1 (src/dg/lifting/lifting_br1.t90:47-167), strong form, non-conservative volume
 * integral (lifting_volint.t90:262-328), periodic (no BC sides). */
static void lifting_br1(orc_ctx *c)
{
    int N = c->N, n = c->n;
    long nf = (long)n * n * c->nSides;
    /* Lifting_FillFlux (lifting_fillflux.t90:39-93), sig=-1 */
#pragma omp parallel for schedule(static)
    for (long s = 0; s < c->nSides; s++)
        for (int q = 0; q <= N; q++)
            for (int p = 0; p <= N; p++) {
                long fi = p + n * (q + n * s);
                for (int iv = 0; iv < 5; iv++) {
                    int pv = LIFT2PRIM[iv];
                    c->liftFlux[5 * fi + iv] =
                        0.5 * c->SurfElem[fi] * (-1.0 * c->UPrim_master[6 * fi + pv] + c->UPrim_slave[6 * fi + pv]);
                }
                /* Lifting_FillFlux_NormVec (lifting_fillflux.t90:212-256) */
                for (int iv = 0; iv < 5; iv++) {
                    c->gxm[5 * fi + iv] = c->liftFlux[5 * fi + iv] * c->NormVec[3 * fi + 0];
                    c->gym[5 * fi + iv] = c->liftFlux[5 * fi + iv] * c->NormVec[3 * fi + 1];
                    c->gzm[5 * fi + iv] = c->liftFlux[5 * fi + iv] * c->NormVec[3 * fi + 2];
                }
            }
    (void)nf;
    /* Lifting_VolInt_Nonconservative_GPU_Kernel */
#pragma omp parallel for schedule(static)
    for (int e = 0; e < c->nElems; e++)
        for (int k = 0; k <= N; k++)
            for (int j = 0; j <= N; j++)
                for (int i = 0; i <= N; i++) {
                    double gxi[5] = {0}, get[5] = {0}, gze[5] = {0};
                    for (int l = 0; l <= N; l++)
                        for (int iv = 0; iv < 5; iv++) {
                            int pv = LIFT2PRIM[iv];
                            gxi[iv] = gxi[iv] + c->D_T[l + n * i] * c->UPrim[VIDX(6, pv, l, j, k, e, n)];
                            get[iv] = get[iv] + c->D_T[l + n * j] * c->UPrim[VIDX(6, pv, i, l, k, e, n)];
                            gze[iv] = gze[iv] + c->D_T[l + n * k] * c->UPrim[VIDX(6, pv, i, j, l, e, n)];
                        }
                    const double *mf = &c->Mf[VIDX(3, 0, i, j, k, e, n)];
                    const double *mg = &c->Mg[VIDX(3, 0, i, j, k, e, n)];
                    const double *mh = &c->Mh[VIDX(3, 0, i, j, k, e, n)];
                    for (int iv = 0; iv < 5; iv++) {
                        c->gradUx[VIDX(5, iv, i, j, k, e, n)] = mf[0] * gxi[iv] + mg[0] * get[iv] + mh[0] * gze[iv];
                        c->gradUy[VIDX(5, iv, i, j, k, e, n)] = mf[1] * gxi[iv] + mg[1] * get[iv] + mh[1] * gze[iv];
                        c->gradUz[VIDX(5, iv, i, j, k, e, n)] = mf[2] * gxi[iv] + mg[2] * get[iv] + mh[2] * gze[iv];
                    }
                }
    /* SurfIntLifting x,y,z with doApplyJacobian (lifting_br1.t90:120-124) */
    surfint_device(c, 5, c->gxm, c->gxm, 1, 1, c->gradUx);
    surfint_device(c, 5, c->gym, c->gym, 1, 1, c->gradUy);
    surfint_device(c, 5, c->gzm, c->gzm, 1, 1, c->gradUz);
    /* ProlongToFaceLifting (lifting_br1.t90:152-156) */
    prolong_gl(c, 5, c->gradUx, c->gxm, c->gxs);
    prolong_gl(c, 5, c->gradUy, c->gym, c->gys);
    prolong_gl(c, 5, c->gradUz, c->gzm, c->gzs);
}

/* InitExactFunc's doCalcSource (src/equations/navierstokes/idealgas/exactfunc.f90:131-139): exact functions 4,
 * 41, 42, 43 carry a source, any other switches it off. */
int orc_set_source(orc_ctx *c, int exactFunc, const double *advVel, const double *Elem_xGP)
{
    long nDOF = (long)c->n * c->n * c->n * c->nElems;
    if (exactFunc != 4 && exactFunc != 41 && exactFunc != 42 && exactFunc != 43) {
        c->src_func = 0;
        return 0;
    }
    if (!c->Elem_xGP) c->Elem_xGP = (double *)xcalloc(3 * nDOF, 8);
    memcpy(c->Elem_xGP, Elem_xGP, (size_t)(3 * nDOF) * sizeof(double));
    for (int d = 0; d < 3; d++) c->adv[d] = advVel[d];
    c->src_func = exactFunc;
    return 0;
}

/* Ut_src of CalcSource_CUDA_Kernel (exactfunc.f90:946-1111), parabolic 3D build, at one node x and time t:
 * the physical-space source rate of the exact function. */
static void source_point(const orc_ctx *c, const double *x, double t, double *S)
{
    const double Frequency = 1.0, Amplitude = 0.1, Cc = 2.0;
    const double kappa = c->kappa, mu0 = c->mu0, Pr = c->Pr;
    const double pi = acos(-1.0);
    const double Omega = pi * Frequency;
    for (int v = 0; v < 5; v++) S[v] = 0.0;
    if (c->src_func == 4) { /* exactfunc.f90:991-1022 */
        double a = c->adv[0] * 2.0 * pi, tmp[6];
        tmp[0] = -a + 3.0 * Omega;
        tmp[1] = -a + 0.5 * Omega * (1.0 + kappa * 5.0);
        tmp[2] = Amplitude * Omega * (kappa - 1.0);
        tmp[3] = 0.5 * ((9.0 + kappa * 15.0) * Omega - 8.0 * a);
        tmp[4] = Amplitude * (3.0 * Omega * kappa - a);
        tmp[5] = 3.0 * mu0 * kappa * Omega * Omega / Pr;
        for (int q = 0; q < 6; q++) tmp[q] *= Amplitude;
        double arg = Omega * (x[0] + x[1] + x[2]) - a * t;
        double cs = cos(arg), sn = sin(arg), sn2 = 2.0 * sn * cs;
        S[0] = tmp[0] * cs;
        for (int v = 1; v <= 3; v++) S[v] = tmp[1] * cs + tmp[2] * sn2;
        S[4] = tmp[3] * cs + tmp[4] * sn2 + tmp[5] * sn;
    } else { /* 41, 42, 43: exactfunc.f90:1024-1101 */
        int d = c->src_func - 41;
        double A = Amplitude, a = c->adv[d] * 2.0 * pi, xd = x[d];
        double cs = cos(Omega * xd - a * t), sn = sin(Omega * xd - a * t);
        double sn2 = sin(2.0 * Omega * xd - 2.0 * a * t);
        S[0] = (-A * a + A * Omega) * cs;
        S[1 + d] = (-A * A * Omega + A * A * Omega * kappa) * sn2 +
                   (-A * a + 2.0 * A * Omega * kappa * Cc - 0.5 * A * Omega * kappa + 1.5 * A * Omega -
                    2.0 * A * Omega * Cc) * cs;
        S[4] = mu0 * kappa * A * sn * Omega * Omega / Pr + (-A * A * a + A * A * Omega * kappa) * sn2 +
               0.5 * (-4.0 * A * a * Cc + 4.0 * A * Omega * kappa * Cc - A * Omega * kappa + A * Omega) * cs;
    }
}

/* the source alone: S(5,nDOF) at time t (zero without one) */
int orc_source(orc_ctx *c, double t, double *S)
{
    long nDOF = (long)c->n * c->n * c->n * c->nElems;
    if (!c->src_func) {
        memset(S, 0, (size_t)(5 * nDOF) * sizeof(double));
        return 0;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nDOF; i++) source_point(c, &c->Elem_xGP[3 * i], t, &S[5 * i]);
    return 0;
}

int orc_rhs(orc_ctx *c, const double *U, double *Ut) { return orc_rhs_t(c, U, Ut, 0.0); }

/* DGTimeDerivative_weakForm(t) (src/dg/dg.f90:255-425) */
int orc_rhs_t(orc_ctx *c, const double *U, double *Ut, double t)
{
    int N = c->N, n = c->n;
    long nDOF = (long)n * n * n * c->nElems, nfn = (long)n * n * c->nSides;
    double mu = c->mu0, lambda = c->lambda;
    /* 2. ProlongToFaceCons */
    prolong_gl(c, 5, U, c->U_master, c->U_slave);
    /* 3. ConsToPrim volume */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nDOF; i++) cons_to_prim(&U[5 * i], &c->UPrim[6 * i], c->kappa, c->R);
    /* 5. GetPrimitiveStateSurface (equation.f90:198-261) */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nfn; i++) {
        cons_to_prim(&c->U_master[5 * i], &c->UPrim_master[6 * i], c->kappa, c->R);
        cons_to_prim(&c->U_slave[5 * i], &c->UPrim_slave[6 * i], c->kappa, c->R);
    }
    /* 6. Lifting */
    lifting_br1(c);
    /* 8. VolInt */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nDOF; i++)
        transformed_flux(!c->split, &U[5 * i], &c->UPrim[6 * i], &c->gradUx[5 * i], &c->gradUy[5 * i],
                         &c->gradUz[5 * i], &c->Mf[3 * i], &c->Mg[3 * i], &c->Mh[3 * i], mu, lambda,
                         &c->f[5 * i], &c->g[5 * i], &c->h[5 * i]);
    apply_dmatrix(c, c->f, c->g, c->h, 1, Ut);
    if (c->split) {
        /* VolInt_splitForm_Kernel (src/dg/volint.f90:265-353) */
#pragma omp parallel for schedule(static)
        for (int e = 0; e < c->nElems; e++)
            for (int k = 0; k <= N; k++)
                for (int j = 0; j <= N; j++)
                    for (int i = 0; i <= N; i++) {
                        double tmp[5] = {0, 0, 0, 0, 0}, F[5];
                        long me = i + n * (j + n * (k + n * (long)e));
                        for (int l = 0; l <= N; l++) {
                            long o = l + n * (j + n * (k + n * (long)e));
                            split_volume_flux_pi(&U[5 * me], &c->UPrim[6 * me], &U[5 * o], &c->UPrim[6 * o],
                                                 &c->Mf[3 * me], &c->Mf[3 * o], F);
                            for (int v = 0; v < 5; v++) tmp[v] = tmp[v] + c->DVolSurf[l + n * i] * F[v];
                        }
                        for (int l = 0; l <= N; l++) {
                            long o = i + n * (l + n * (k + n * (long)e));
                            split_volume_flux_pi(&U[5 * me], &c->UPrim[6 * me], &U[5 * o], &c->UPrim[6 * o],
                                                 &c->Mg[3 * me], &c->Mg[3 * o], F);
                            for (int v = 0; v < 5; v++) tmp[v] = tmp[v] + c->DVolSurf[l + n * j] * F[v];
                        }
                        for (int l = 0; l <= N; l++) {
                            long o = i + n * (j + n * (l + n * (long)e));
                            split_volume_flux_pi(&U[5 * me], &c->UPrim[6 * me], &U[5 * o], &c->UPrim[6 * o],
                                                 &c->Mh[3 * me], &c->Mh[3 * o], F);
                            for (int v = 0; v < 5; v++) tmp[v] = tmp[v] + c->DVolSurf[l + n * k] * F[v];
                        }
                        for (int v = 0; v < 5; v++) Ut[5 * me + v] = Ut[5 * me + v] + tmp[v];
                    }
    }
    /* 11. FillFlux (src/dg/fillflux.f90:45-172): Riemann, ViscousFlux, *SurfElem, copy */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nfn; i++) {
        double F[5], fv[5];
        riemann_point(c, &c->U_master[5 * i], &c->U_slave[5 * i], &c->UPrim_master[6 * i], &c->UPrim_slave[6 * i],
                      &c->NormVec[3 * i], &c->TangVec1[3 * i], &c->TangVec2[3 * i], F);
        /* ViscousFlux_Kernel_CUDA (riemann.f90:638-700) */
        normal_diff_flux(&c->UPrim_master[6 * i], &c->gxm[5 * i], &c->gym[5 * i], &c->gzm[5 * i],
                         &c->NormVec[3 * i], mu, lambda, fv);
        for (int v = 0; v < 5; v++) F[v] = F[v] + 0.5 * fv[v];
        normal_diff_flux(&c->UPrim_slave[6 * i], &c->gxs[5 * i], &c->gys[5 * i], &c->gzs[5 * i],
                         &c->NormVec[3 * i], mu, lambda, fv);
        for (int v = 0; v < 5; v++) F[v] = F[v] + 0.5 * fv[v];
        for (int v = 0; v < 5; v++) {
            c->Flux_master[5 * i + v] = F[v] * c->SurfElem[i];
            c->Flux_slave[5 * i + v] = c->Flux_master[5 * i + v];
        }
    }
    /* 11.5 SurfIntCons (no Jacobian) */
    surfint_device(c, 5, c->Flux_master, c->Flux_slave, 0, 0, Ut);
    /* 12. Ut = -Ut; 13. CalcSource: Ut += Ut_src/sJ (exactfunc.f90:1106-1108); 14. ApplyJacobian
     * (applyjacobian.t90:196-224) */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nDOF; i++) {
        double sj = c->sJ[i], S[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (c->src_func) source_point(c, &c->Elem_xGP[3 * i], t, S);
        for (int v = 0; v < 5; v++) {
            double u = Ut[5 * i + v] * -1.0;
            if (c->src_func) u = u + S[v] / sj;
            Ut[5 * i + v] = u * sj;
        }
    }
    return 0;
}

const double *orc_array(orc_ctx *c, int which, long *count)
{
    long nv = (long)c->n * c->n * c->n * c->nElems, nf = (long)c->n * c->n * c->nSides;
    switch (which) {
    case 0: *count = 5 * nf; return c->U_master;
    case 1: *count = 5 * nf; return c->U_slave;
    case 2: *count = 6 * nv; return c->UPrim;
    case 3: *count = 5 * nv; return c->gradUx;
    case 4: *count = 5 * nv; return c->gradUy;
    case 5: *count = 5 * nv; return c->gradUz;
    case 6: *count = 5 * nf; return c->gxm;
    case 7: *count = 5 * nf; return c->gxs;
    case 8: *count = 5 * nf; return c->Flux_master;
    case 9: *count = 5 * nf; return c->gym;
    case 10: *count = 5 * nf; return c->gzm;
    case 11: *count = 5 * nf; return c->gys;
    case 12: *count = 5 * nf; return c->gzs;
    default: *count = 0; return NULL;
    }
}

/* Carpenter-Kennedy RK4-5 (src/timedisc/timedisc_vars.f90:213-227) */
void orc_lserk_coefs(double *A, double *b, double *cc)
{
    A[0] = 0.0;
    A[1] = 567301805773.0 / 1357537059087.0;
    A[2] = 2404267990393.0 / 2016746695238.0;
    A[3] = 3550918686646.0 / 2091501179385.0;
    A[4] = 1275806237668.0 / 842570457699.0;
    b[0] = 1432997174477.0 / 9575080441755.0;
    b[1] = 5161836677717.0 / 13612068292357.0;
    b[2] = 1720146321549.0 / 2090206949498.0;
    b[3] = 3134564353537.0 / 4481467310338.0;
    b[4] = 2277821191437.0 / 14882151754819.0;
    cc[0] = 0.0;
    cc[1] = 1432997174477.0 / 9575080441755.0;
    cc[2] = 2526269341429.0 / 6820363962896.0;
    cc[3] = 2006345519317.0 / 3224310063776.0;
    cc[4] = 2802321613138.0 / 2924317926251.0;
}

/* TimeStepByLSERKW2 + VAXPB_OUT_VAXPB_IN (src/timedisc/timestep.f90:49-120, src/globals/vector.f90:163-183) */
int orc_lserk_step(orc_ctx *c, double *U, double *Ut_tmp, double t, double dt)
{
    double A[5], b[5], cc[5], bdt[5];
    long nTot = 5L * c->n * c->n * c->n * c->nElems;
    orc_lserk_coefs(A, b, cc);
    for (int s = 0; s < 5; s++) bdt[s] = b[s] * dt;
    for (int s = 0; s < 5; s++) {
        double tStage = (s == 0) ? t : t + cc[s] * dt;
        orc_rhs_t(c, U, c->Ut_stage, tStage);
        double mRKA = (s == 0) ? 0.0 : -1.0 * A[s];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nTot; i++) {
            Ut_tmp[i] = Ut_tmp[i] * mRKA + c->Ut_stage[i];
            U[i] = U[i] + Ut_tmp[i] * bdt[s];
        }
    }
    return 0;
}

/* CFLScaleAlpha / DFLScaleAlpha for GL nodes, Carpenter RK4-5 (timedisc_vars.f90:105-109,188-227) */
static const double CFL_ALPHA_GL[15] = {4.7497, 3.4144, 2.8451, 2.4739, 2.2027, 1.9912, 1.8225, 1.6830,
                                        1.5682, 1.4692, 1.3849, 1.3106, 1.2454, 1.1880, 1.1362};
static const double DFL_ALPHA_GL[10] = {4.50, 1.95, 1.18, 0.79, 0.56, 0.42, 0.32, 0.25, 0.20, 0.17};

/* CALCTIMESTEP / CalcMaxEigenvalue / InitCalctimestep (calctimestep.f90:51-296) with FillCFL_DFL
 * (timedisc_func.f90:480-550).  The reduction implements the intended min_e(CFL*2/lambda_e). */
int orc_calc_dt(orc_ctx *c, const double *U, double cfl, double dfl, double *dt, int *errType,
                double *dtConv, double *dtVisc)
{
    int N = c->N, n = c->n;
    int Nc = N < 15 ? N : 15, Nd = N < 10 ? N : 10;
    double CFLScale = cfl * CFL_ALPHA_GL[Nc - 1] / (2.0 * N + 1.0);
    double DFLScale = dfl * DFL_ALPHA_GL[Nd - 1] * 1.853 / ((2.0 * N + 1.0) * (2.0 * N + 1.0));
    double KappasPr = c->kappa / c->Pr;
    double KPmax = KappasPr > 4.0 / 3.0 ? KappasPr : 4.0 / 3.0;
    double tc = 1e300, tv = 1e300;
    int err = 0;
    for (int e = 0; e < c->nElems; e++) {
        double mx[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k <= N; k++)
            for (int j = 0; j <= N; j++)
                for (int i = 0; i <= N; i++) {
                    long id = i + n * (j + n * (k + n * (long)e));
                    const double *u = &U[5 * id];
                    double srho = 1.0 / u[0];
                    double vel[3] = {u[1] * srho, u[2] * srho, u[3] * srho};
                    double p = (c->kappa - 1.0) * (u[4] - 0.5 * (vel[0] * u[1] + vel[1] * u[2] + vel[2] * u[3]));
                    double cs = sqrt(c->kappa * p * srho);
                    double sj = c->sJ[id];
                    double vsJ[3] = {vel[0] * sj, vel[1] * sj, vel[2] * sj};
                    const double *M[3] = {&c->Mf[3 * id], &c->Mg[3 * id], &c->Mh[3 * id]};
                    if (!(u[0] > 0.0)) err = 1;
                    for (int d = 0; d < 3; d++) {
                        const double *m = M[d];
                        double adv = sj * sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
                        double lc = fabs(m[0] * vsJ[0] + m[1] * vsJ[1] + m[2] * vsJ[2]) + cs * adv;
                        if (lc > mx[d] || lc != lc) mx[d] = lc;
                        double a0 = m[0] * sj, a1 = m[1] * sj, a2 = m[2] * sj;
                        double visc = KPmax * (a0 * a0 + a1 * a1 + a2 * a2);
                        double lv = c->mu0 * srho * visc;
                        if (lv > mx[3 + d] || lv != lv) mx[3 + d] = lv;
                    }
                }
        double lcon = mx[0] + mx[1] + mx[2], lvis = mx[3] + mx[4] + mx[5];
        double a = CFLScale * 2.0 / lcon, b = DFLScale * 4.0 / lvis;
        if (a != a) err = err ? err : 2;
        if (b != b) err = err ? err : 3;
        if (a < tc) tc = a;
        if (b < tv) tv = b;
    }
    *dtConv = tc; *dtVisc = tv;
    *dt = tc < tv ? tc : tv;
    *errType = err;
    return 0;
}

int orc_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}
