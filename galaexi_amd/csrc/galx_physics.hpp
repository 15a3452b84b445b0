// galx_physics.hpp -- pointwise compressible Navier-Stokes physics for the gfx950 kernels.
//
// Templated on the arithmetic type T (double for the FP64 form, float for the mixed form).  Each routine
// restates the reference routine cited next to it and keeps its operation order, so that the FP64 form
// agrees with the CPU restatement to rounding.
//
// Pressure offset: states carry p' = p - pref and E' = E - pref/(kappa-1) for one constant pref per context.
// A constant pressure has zero discrete divergence on free-stream preserving meshes, so p' can replace p in
// every additive pressure term of the momentum fluxes; the full values are restored where they enter
// nonlinearly (sound speed: p'+pref, enthalpy: E'+p'+hshift with hshift = kappa*pref/(kappa-1)).  The FP64
// form uses pref = 0 (reference arithmetic); the mixed form uses the mean initial pressure so that the FP32
// values hold the pressure fluctuation at full relative precision.
#pragma once
#include <hip/hip_runtime.h>

namespace galx {

enum { RIEMANN_LF = 0, RIEMANN_ROE = 1, RIEMANN_ROEEF = 3 };

// rR = 1/R and rkm1 = 1/(kappa-1), formed once on the host: the pointwise routines multiply by them where the
// reference divides by R and by kappa-1
template <typename T>
struct Gas {
    T kappa, R, mu, lambda, pref, hshift, rR, rkm1;
};

// 1/x in FP64 from the hardware reciprocal (v_rcp_f64) refined by two Newton steps: for the normal, finite densities
// it is taken of, the IEEE quotient (bit for bit at all 8.4M values of [1e-3, 1e3) checked on MI355X), in five
// instructions instead of the scaled division sequence (about twelve, two of them long-latency)
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// square root, reciprocal and quotient.  FP64: IEEE square root and quotient, the reciprocal of rcp64.  FP32 (the
// mixed form): the hardware instructions (v_sqrt_f32, v_rcp_f32, about 1 ulp), whose error is far below the
// rounding of the FP32 fluxes they feed, at one instruction instead of a correction sequence (the mixed kernels
// are VALU-bound)
__device__ __forceinline__ double gsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float gsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double grcp(double x) { return rcp64(x); }
__device__ __forceinline__ float grcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double gdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float gdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double gabs(double x) { return fabs(x); }
__device__ __forceinline__ float gabs(float x) { return fabsf(x); }
__device__ __forceinline__ double gmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float gmax(float a, float b) { return fmaxf(a, b); }

// ConsToPrim (src/equations/navierstokes/idealgas/eos.f90:212-239): prim = (rho,u,v,w,p,T)
template <typename T>
__device__ __forceinline__ void cons_to_prim(const T* c, T* p, const Gas<T>& g)
{
    T sRho = grcp(c[0]);
    p[0] = c[0];
    p[1] = c[1] * sRho;
    p[2] = c[2] * sRho;
    p[3] = c[3] * sRho;
    p[4] = (g.kappa - T(1)) * (c[4] - T(0.5) * (c[1] * p[1] + c[2] * p[2] + c[3] * p[3]));
    p[5] = p[4] * sRho * g.rR;
}

// face state from a trace record (rho, rho*u, rho*v, rho*w, p'): Uc = (rho, rho*u, rho*v, rho*w, E'),
// P = (rho, u, v, w, p', T); ConsToPrim (eos.f90:212-239) with the pressure taken from the record
template <typename T>
__device__ __forceinline__ void trace_state(const T* tr, const Gas<T>& g, T* Uc, T* P)
{
    T sRho = grcp(tr[0]);
    P[0] = tr[0];
    P[1] = tr[1] * sRho;
    P[2] = tr[2] * sRho;
    P[3] = tr[3] * sRho;
    P[4] = tr[4];
    P[5] = (tr[4] + g.pref) * sRho * g.rR;
    Uc[0] = tr[0];
    Uc[1] = tr[1];
    Uc[2] = tr[2];
    Uc[3] = tr[3];
    Uc[4] = tr[4] * g.rkm1 + T(0.5) * (tr[1] * P[1] + tr[2] * P[2] + tr[3] * P[3]);
}

// the same with one Newton step: for FP64 values that are rounded to FP32 afterwards (the mixed form's trace
// records and node states), three instructions
__device__ __forceinline__ double rcp64_fast(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// trace record of a conservative FP64 state (ConsToPrim pressure, shifted by pref); FAST for records stored in FP32
template <bool FAST = false>
__device__ __forceinline__ void trace_record(const double* c, double kappa, double pref, double* tr)
{
    double sRho = FAST ? rcp64_fast(c[0]) : rcp64(c[0]);
    double u = c[1] * sRho, v = c[2] * sRho, w = c[3] * sRho;
    double p = (kappa - 1.0) * (c[4] - 0.5 * (c[1] * u + c[2] * v + c[3] * w));
    tr[0] = c[0];
    tr[1] = c[1];
    tr[2] = c[2];
    tr[3] = c[3];
    tr[4] = p - pref;
}

// extended state in the normal frame (eos.h:119-135): DENS, MOM(3), ENER, SRHO, VEL(3), PRES
template <typename T>
struct Ext {
    T dens, mom[3], ener, srho, vel[3], pres;
};

template <typename T>
__device__ __forceinline__ void rotate_state(const T* U, const T* P, const T* nv, const T* t1, const T* t2, Ext<T>& X)
{
    X.dens = U[0];
    X.srho = grcp(X.dens);
    X.ener = U[4];
    X.pres = P[4];
    X.vel[0] = P[1] * nv[0] + P[2] * nv[1] + P[3] * nv[2];
    X.vel[1] = P[1] * t1[0] + P[2] * t1[1] + P[3] * t1[2];
    X.vel[2] = P[1] * t2[0] + P[2] * t2[1] + P[3] * t2[2];
    X.mom[0] = X.dens * X.vel[0];
    X.mom[1] = X.dens * X.vel[1];
    X.mom[2] = X.dens * X.vel[2];
}

// EvalEulerFlux1D_fast (src/equations/navierstokes/flux.f90:1123-1143)
template <typename T>
__device__ __forceinline__ void euler_flux_1d(const Ext<T>& X, T hshift, T* F)
{
    F[0] = X.mom[0];
    F[1] = X.mom[0] * X.vel[0] + X.pres;
    F[2] = X.mom[0] * X.vel[1];
    F[3] = X.mom[0] * X.vel[2];
    F[4] = (X.ener + X.pres + hshift) * X.vel[0];
}

// SplitSurfaceFluxPI (src/equations/navierstokes/splitflux.f90:735-761)
template <typename T>
__device__ __forceinline__ void split_surface_flux_pi(const Ext<T>& L, const Ext<T>& R, T hshift, T* F)
{
    T HL = gdiv(L.ener + L.pres + hshift, L.dens);
    T HR = gdiv(R.ener + R.pres + hshift, R.dens);
    T sr = L.dens + R.dens, su = L.vel[0] + R.vel[0];
    F[0] = T(0.25) * sr * su;
    F[1] = T(0.125) * sr * (su * su) + T(0.5) * (L.pres + R.pres);
    F[2] = T(0.125) * sr * su * (L.vel[1] + R.vel[1]);
    F[3] = T(0.125) * sr * su * (L.vel[2] + R.vel[2]);
    F[4] = T(0.125) * sr * (HL + HR) * su;
}

__device__ __forceinline__ double ext_cons(const Ext<double>& X, int v)
{
    return v == 0 ? X.dens : (v == 4 ? X.ener : X.mom[v - 1]);
}
__device__ __forceinline__ float ext_cons(const Ext<float>& X, int v)
{
    return v == 0 ? X.dens : (v == 4 ? X.ener : X.mom[v - 1]);
}

// Riemann_LF / Riemann_Roe / Riemann_RoeEntropyFix (src/equations/navierstokes/riemann.f90:707-989);
// SPLIT selects the SPLIT_DG branch (PI surface flux as the central part).
template <typename T, bool SPLIT>
__device__ __forceinline__ void riemann_solver(int solver, const Ext<T>& L, const Ext<T>& R, const T* FL, const T* FR,
                                               const Gas<T>& g, T* F)
{
    const T kappa = g.kappa, hshift = g.hshift;
    if (solver == RIEMANN_LF) {
        // max(c_L, c_R) as the root of the larger c^2: the square root is monotone, so the same value, one root
        T lmax = gmax(gabs(R.vel[0]), gabs(L.vel[0])) +
                 gsqrt(gmax(kappa * (L.pres + g.pref) * L.srho, kappa * (R.pres + g.pref) * R.srho));
        if (!SPLIT) {
#pragma unroll
            for (int v = 0; v < 5; v++) F[v] = T(0.5) * ((FL[v] + FR[v]) - lmax * (ext_cons(R, v) - ext_cons(L, v)));
        } else {
            split_surface_flux_pi(L, R, hshift, F);
#pragma unroll
            for (int v = 0; v < 5; v++) F[v] = F[v] - T(0.5) * lmax * (ext_cons(R, v) - ext_cons(L, v));
        }
        return;
    }
    T HL = (L.ener + L.pres + hshift) * L.srho, HR = (R.ener + R.pres + hshift) * R.srho;
    T sqL = gsqrt(L.dens), sqR = gsqrt(R.dens);
    T ss = grcp(sqL + sqR);
    T rv0 = (sqR * R.vel[0] + sqL * L.vel[0]) * ss;
    T rv1 = (sqR * R.vel[1] + sqL * L.vel[1]) * ss;
    T rv2 = (sqR * R.vel[2] + sqL * L.vel[2]) * ss;
    T absVel = rv0 * rv0 + rv1 * rv1 + rv2 * rv2;
    T RoeH = (sqR * HR + sqL * HL) * ss;
    T Roec = gsqrt((kappa - T(1)) * (RoeH - T(0.5) * absVel));
    T a[5] = {rv0 - Roec, rv0, rv0, rv0, rv0 + Roec};
    T al[5], aa[5];
    if (solver == RIEMANN_ROE) {
        T dU[6];
#pragma unroll
        for (int v = 0; v < 5; v++) dU[v] = ext_cons(R, v) - ext_cons(L, v);
        dU[5] = dU[4] - (dU[2] - rv1 * dU[0]) * rv1 - (dU[3] - rv2 * dU[0]) * rv2;
        al[2] = dU[2] - rv1 * dU[0];
        al[3] = dU[3] - rv2 * dU[0];
        al[1] = gdiv(kappa - T(1), Roec * Roec) * (dU[0] * (RoeH - rv0 * rv0) - dU[5] + rv0 * dU[1]);
        al[0] = gdiv(T(0.5), Roec) * (dU[0] * (rv0 + Roec) - dU[1] - Roec * al[1]);
        al[4] = dU[0] - al[0] - al[1];
#pragma unroll
        for (int v = 0; v < 5; v++) aa[v] = gabs(a[v]);
    } else {
        T cL = gsqrt(kappa * (L.pres + g.pref) * L.srho), cR = gsqrt(kappa * (R.pres + g.pref) * R.srho);
        T RoeDens = gsqrt(L.dens * R.dens);
        T d0 = R.dens - L.dens, d1 = R.vel[0] - L.vel[0], d2 = R.vel[1] - L.vel[1], d3 = R.vel[2] - L.vel[2],
          d4 = R.pres - L.pres;
        T tmp = gdiv(T(0.5), Roec * Roec);
        al[0] = tmp * (d4 - RoeDens * Roec * d1);
        al[1] = d0 - d4 * T(2) * tmp;
        al[2] = RoeDens * d2;
        al[3] = RoeDens * d3;
        al[4] = tmp * (d4 + RoeDens * Roec * d1);
        T aL[5] = {L.vel[0] - cL, L.vel[0], L.vel[0], L.vel[0], L.vel[0] + cL};
        T aR[5] = {R.vel[0] - cR, R.vel[0], R.vel[0], R.vel[0], R.vel[0] + cR};
#pragma unroll
        for (int v = 0; v < 5; v++) {
            T da = gmax(T(0), gmax(a[v] - aL[v], aR[v] - a[v]));
            aa[v] = (gabs(a[v]) < da) ? T(0.5) * (gdiv(a[v] * a[v], da) + da) : gabs(a[v]);
        }
    }
    T r1[5] = {T(1), a[0], rv1, rv2, RoeH - rv0 * Roec};
    T r2[5] = {T(1), rv0, rv1, rv2, T(0.5) * absVel};
    T r3[5] = {T(0), T(0), T(1), T(0), rv1};
    T r4[5] = {T(0), T(0), T(0), T(1), rv2};
    T r5[5] = {T(1), a[4], rv1, rv2, RoeH + rv0 * Roec};
    if (!SPLIT) {
#pragma unroll
        for (int v = 0; v < 5; v++)
            F[v] = T(0.5) * ((FL[v] + FR[v]) - al[0] * aa[0] * r1[v] - al[1] * aa[1] * r2[v] - al[2] * aa[2] * r3[v] -
                             al[3] * aa[3] * r4[v] - al[4] * aa[4] * r5[v]);
    } else {
        split_surface_flux_pi(L, R, hshift, F);
#pragma unroll
        for (int v = 0; v < 5; v++)
            F[v] = F[v] - T(0.5) * (al[0] * aa[0] * r1[v] + al[1] * aa[1] * r2[v] + al[2] * aa[2] * r3[v] +
                                    al[3] * aa[3] * r4[v] + al[4] * aa[4] * r5[v]);
    }
}

// Riemann driver (src/equations/navierstokes/riemann.f90:216-289): rotate, solve, rotate back.
template <typename T, bool SPLIT>
__device__ __forceinline__ void riemann_point(int solver, const T* UL, const T* UR, const T* PL, const T* PR, const T* nv,
                                              const T* t1, const T* t2, const Gas<T>& g, T* Fout)
{
    Ext<T> L, R;
    rotate_state(UL, PL, nv, t1, t2, L);
    rotate_state(UR, PR, nv, t1, t2, R);
    T FL[5] = {0, 0, 0, 0, 0}, FR[5] = {0, 0, 0, 0, 0}, F[5];
    if (!SPLIT) {
        euler_flux_1d(L, g.hshift, FL);
        euler_flux_1d(R, g.hshift, FR);
    }
    riemann_solver<T, SPLIT>(solver, L, R, FL, FR, g, F);
    Fout[0] = F[0];
#pragma unroll
    for (int d = 0; d < 3; d++) Fout[1 + d] = nv[d] * F[1] + t1[d] * F[2] + t2[d] * F[3];
    Fout[4] = F[4];
}

// Riemann_LF with the SplitSurfaceFluxPI central part (riemann.f90:707-733, splitflux.f90:735-761) for the
// Cartesian flux F.n, without the normal-frame rotation of the Riemann routine (riemann.f90:216-289): rotating into
// (n,t1,t2) and back leaves the PI momentum flux 1/8 {rho} {u_n} {u} + 1/2 {p} n with {x} = x_L + x_R and the
// dissipation -lmax/2 (U_R - U_L) unchanged, so only u.n is needed (algebraically identical to riemann_point).
// The states come in the caller's order a, b with (L, R) = (a, b) for sgn = +1 and (b, a) for sgn = -1: every
// operation of the central part and of lmax is symmetric in L and R (sums, products, max), and
// U_R - U_L = sgn (U_b - U_a) exactly, so the flux is the one of the (L, R) order bit for bit and the two elements
// of a face form the same value without selecting their operands lane by lane.
template <typename T>
__device__ __forceinline__ void lf_pi_cartesian(const T* Ua, const T* Ub, const T* Pa, const T* Pb, const T* nv, T sgn,
                                                const Gas<T>& g, T* F)
{
    const T vna = Pa[1] * nv[0] + Pa[2] * nv[1] + Pa[3] * nv[2];
    const T vnb = Pb[1] * nv[0] + Pb[2] * nv[1] + Pb[3] * nv[2];
    const T sRhoA = grcp(Ua[0]), sRhoB = grcp(Ub[0]);
    const T Ha = (Ua[4] + Pa[4] + g.hshift) * sRhoA, Hb = (Ub[4] + Pb[4] + g.hshift) * sRhoB;
    // max(c_L, c_R) = sqrt(max(c_L^2, c_R^2)) (monotone root): one square root instead of two
    const T lmax = gmax(gabs(vna), gabs(vnb)) +
                   gsqrt(gmax(g.kappa * (Pa[4] + g.pref) * sRhoA, g.kappa * (Pb[4] + g.pref) * sRhoB));
    const T sr = Ua[0] + Ub[0], su = vna + vnb, q = T(0.125) * sr * su, sp = T(0.5) * (Pa[4] + Pb[4]);
    F[0] = T(2) * q;
#pragma unroll
    for (int d = 0; d < 3; d++) F[1 + d] = q * (Pa[1 + d] + Pb[1 + d]) + sp * nv[d];
    F[4] = q * (Ha + Hb);
    const T hl = T(0.5) * lmax * sgn;
#pragma unroll
    for (int v = 0; v < 5; v++) F[v] = F[v] - hl * (Ub[v] - Ua[v]);
}

// gradients of the lifted variables in the order used here: g[4*dir + c], dir in {x,y,z}, c in {u,v,w,T}
// Shear stress and heat flux terms shared by EvalNormalDiffFlux3D (flux.f90:779-847) and
// EvalTransformed(Euler)DiffFlux3D (flux.f90:509-700).
template <typename T>
struct Stress {
    T txx, tyy, tzz, txy, txz, tyz, qx, qy, qz;
};

template <typename T>
__device__ __forceinline__ void stress(const T* P, const T* g, const Gas<T>& gas, Stress<T>& s)
{
    const T s23 = T(2.0 / 3.0), s43 = T(4.0 / 3.0);
    const T* gx = g;
    const T* gy = g + 4;
    const T* gz = g + 8;
    T mu = gas.mu;
    s.txx = mu * (s43 * gx[0] - s23 * gy[1] - s23 * gz[2]);
    s.tyy = mu * (-s23 * gx[0] + s43 * gy[1] - s23 * gz[2]);
    s.tzz = mu * (-s23 * gx[0] - s23 * gy[1] + s43 * gz[2]);
    s.txy = mu * (gy[0] + gx[1]);
    s.txz = mu * (gz[0] + gx[2]);
    s.tyz = mu * (gz[1] + gy[2]);
    s.qx = s.txx * P[1] + s.txy * P[2] + s.txz * P[3] + gas.lambda * gx[3];
    s.qy = s.txy * P[1] + s.tyy * P[2] + s.tyz * P[3] + gas.lambda * gy[3];
    s.qz = s.txz * P[1] + s.tyz * P[2] + s.tzz * P[3] + gas.lambda * gz[3];
}

// EvalNormalDiffFlux3D (flux.f90:779-847)
template <typename T>
__device__ __forceinline__ void normal_diff_flux(const T* P, const T* g, const T* nv, const Gas<T>& gas, T* f)
{
    const T s23 = T(2.0 / 3.0), s43 = T(4.0 / 3.0);
    const T* gx = g;
    const T* gy = g + 4;
    const T* gz = g + 8;
    T mu = gas.mu;
    T txx = mu * (s43 * gx[0] - s23 * gy[1] - s23 * gz[2]);
    T tyy = mu * (-s23 * gx[0] + s43 * gy[1] - s23 * gz[2]);
    T tzz = mu * (-s23 * gx[0] - s23 * gy[1] + s43 * gz[2]);
    T txy = mu * (gy[0] + gx[1]);
    T txz = mu * (gz[0] + gx[2]);
    T tyz = mu * (gz[1] + gy[2]);
    f[0] = T(0);
    f[1] = -nv[0] * txx - nv[1] * txy - nv[2] * txz;
    f[2] = -nv[0] * txy - nv[1] * tyy - nv[2] * tyz;
    f[3] = -nv[0] * txz - nv[1] * tyz - nv[2] * tzz;
    f[4] = -nv[0] * (txx * P[1] + txy * P[2] + txz * P[3] + gas.lambda * gx[3]) -
           nv[1] * (txy * P[1] + tyy * P[2] + tyz * P[3] + gas.lambda * gy[3]) -
           nv[2] * (txz * P[1] + tyz * P[2] + tzz * P[3] + gas.lambda * gz[3]);
}

// contravariant flux in one direction with metric m: EvalTransformedEulerDiffFlux3D (flux.f90:509-614) when
// EULER, else EvalTransformedDiffFlux3D (flux.f90:619-700)
template <typename T, bool EULER>
__device__ __forceinline__ void transformed_flux(const T* U, const T* P, const Stress<T>& s, const T* m, T hshift, T* r)
{
    if (EULER) {
        T Ep = gdiv(U[4] + P[4] + hshift, U[0]);
        T Mmom = m[0] * U[1] + m[1] * U[2] + m[2] * U[3];
        r[0] = Mmom;
        r[1] = Mmom * P[1] + m[0] * P[4] - m[0] * s.txx - m[1] * s.txy - m[2] * s.txz;
        r[2] = Mmom * P[2] + m[1] * P[4] - m[0] * s.txy - m[1] * s.tyy - m[2] * s.tyz;
        r[3] = Mmom * P[3] + m[2] * P[4] - m[0] * s.txz - m[1] * s.tyz - m[2] * s.tzz;
        r[4] = Mmom * Ep - m[0] * s.qx - m[1] * s.qy - m[2] * s.qz;
    } else {
        r[0] = T(0);
        r[1] = -m[0] * s.txx - m[1] * s.txy - m[2] * s.txz;
        r[2] = -m[0] * s.txy - m[1] * s.tyy - m[2] * s.tyz;
        r[3] = -m[0] * s.txz - m[1] * s.tyz - m[2] * s.tzz;
        r[4] = -m[0] * s.qx - m[1] * s.qy - m[2] * s.qz;
    }
}

// SplitVolumeFluxPI (src/equations/navierstokes/splitflux.f90:669-730) with the node-wise (rho,u,v,w,p,H)
// precomputed (H = (E+p)/rho as in the reference).  The reference forms the three Cartesian two-point fluxes and
// contracts them with 0.5*(M_ref+M); the same sum is regrouped around the contravariant mass flux
//   q = 1/4 {rho}' (0.5*(M_ref+M) . {u}')   with {x}' = x_ref + x:
//   F = [2q, q{u}' + 0.5(M_ref+M)_1 {p}', q{v}' + 0.5(M_ref+M)_2 {p}', q{w}' + 0.5(M_ref+M)_3 {p}', q{H}']
// which is algebraically identical and takes ~25 instead of ~70 operations per pair.  p may be shifted by a
// constant (its contribution to the discrete divergence vanishes with free-stream preserving metrics).
template <typename T>
__device__ __forceinline__ void split_volume_flux_pi(T rR, T uR, T vR, T wR, T pR, T HR, T r, T u, T v, T w, T p, T H,
                                                     const T* MR, const T* M, T* Flux)
{
    T sr = rR + r, su = uR + u, sv = vR + v, sw = wR + w, sp = pR + p, sH = HR + H;
    T m0 = MR[0] + M[0], m1 = MR[1] + M[1], m2 = MR[2] + M[2];
    T q = T(0.125) * sr * (m0 * su + m1 * sv + m2 * sw);
    Flux[0] = T(2) * q;
    Flux[1] = q * su + (T(0.5) * m0) * sp;
    Flux[2] = q * sv + (T(0.5) * m1) * sp;
    Flux[3] = q * sw + (T(0.5) * m2) * sp;
    Flux[4] = q * sH;
}

// The same flux from the halves of both nodes' (rho,u,v,w,p,H) and metrics, returning half of it: the sums of
// halves are the means {x} = (x_ref + x)/2, and
//   G = F/2 = [q, q{u} + {M}_1 {p}, q{v} + {M}_2 {p}, q{w} + {M}_3 {p}, q{H}],  q = {rho} ({M} . {u}).
// Every rescaling against split_volume_flux_pi is by a power of two and so exact: 2 G equals F bit for bit (barring
// subnormals), with five fewer operations per pair.
template <typename T>
__device__ __forceinline__ void split_volume_flux_pi_half(T rR, T uR, T vR, T wR, T pR, T HR, T r, T u, T v, T w,
                                                          T p, T H, const T* MR, const T* M, T* G)
{
    T ar = rR + r, au = uR + u, av = vR + v, aw = wR + w, ap = pR + p, aH = HR + H;
    T m0 = MR[0] + M[0], m1 = MR[1] + M[1], m2 = MR[2] + M[2];
    T q = ar * (m0 * au + m1 * av + m2 * aw);
    G[0] = q;
    G[1] = q * au + m0 * ap;
    G[2] = q * av + m1 * ap;
    G[3] = q * aw + m2 * ap;
    G[4] = q * aH;
}

}  // namespace galx
