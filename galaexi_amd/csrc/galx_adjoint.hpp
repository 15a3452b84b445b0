// galx_adjoint.hpp -- discrete adjoint (vector-Jacobian product) of the split-form DG operator on gfx950.
//
// For Ut = R(U) (split Pirozzoli volume flux, Lax-Friedrichs surface flux with the PI central part, BR1 lifting,
// constant viscosity, FP64) and a cotangent lam = dJ/dUt, the kernels form Ubar = (dR/dU)^T lam exactly (to
// rounding), element per workgroup and without atomics: every element gathers what its own nodes receive.
//   k_adj_pre  : lam = kbar + bdt*Ubar (LSERK stage adjoint, in place), mu = -sJ*lam on the faces scaled by the
//                surface lifting weight L_Hat of the writer's local side -> mu traces
//   k_adj_main : forward recompute (primitives, lifted gradients), adjoint of the weak viscous volume term
//                (D_Hat^T transposed), of the split volume fluxes (two-point flux Jacobians transposed and
//                gathered along the lines), of the face fluxes (transposed LF/PI and viscous Jacobians), of the
//                BR1 lifting (transposed strong-form derivative, own surface term), chained through ConsToPrim;
//                Ubar (+)= result, kbar = a*lam, and the traces L_Hat * n.ghat (ghat = sJ*Gbar) for the
//                neighbours' lifting adjoint
//   k_adj_face : the neighbour's lifting surface term on the face nodes
// The forward operator is the one of k_lift/k_rhs (src/dg/dg.f90:255-425), whose pieces are cited there.
#pragma once
#include "galx_kernels.hpp"

namespace galx {

// adjoint with respect to (rho,u,v,w,p,H,T) and a direct E dependence
struct PrimBar {
    double rho, u, v, w, p, H, T, E;
};

__device__ __forceinline__ void prim_bar_zero(PrimBar& b) { b.rho = b.u = b.v = b.w = b.p = b.H = b.T = b.E = 0.0; }

// transposed Jacobian of ConsToPrim (eos.f90:212-239) with H = (E+p)/rho and T = p/(rho*R)
__device__ __forceinline__ void prim_bar_to_cons(const double* c, const PrimBar& b, double kappa, double R,
                                                 double* Ubar)
{
    double rho = c[0], sRho = 1.0 / rho;
    double u = c[1] * sRho, v = c[2] * sRho, w = c[3] * sRho;
    double q2 = 0.5 * (u * u + v * v + w * w);
    double km1 = kappa - 1.0;
    double p = km1 * (c[4] - rho * q2);
    double H = (c[4] + p) * sRho, T = p * sRho / R;
    double dp_drho = km1 * q2;
    Ubar[0] = b.rho - (b.u * u + b.v * v + b.w * w) * sRho + b.p * dp_drho + b.H * (dp_drho - H) * sRho +
              b.T * (dp_drho * sRho / R - T * sRho);
    double s = -km1 * (b.p + b.H * sRho + b.T * sRho / R);
    Ubar[1] = b.u * sRho + s * u;
    Ubar[2] = b.v * sRho + s * v;
    Ubar[3] = b.w * sRho + s * w;
    Ubar[4] = b.E + b.p * km1 + b.H * kappa * sRho + b.T * km1 * sRho / R;
}

// adjoint of the viscous flux -m.(tau, tau.u + lambda grad T) (EvalNormalDiffFlux3D flux.f90:779-847 for m = n,
// EvalTransformedDiffFlux3D flux.f90:619-700 for m = Ja^d) for the flux cotangent fb(1:4):
// accumulates the velocity cotangent uvw_bar(3) and the gradient cotangent Gbar(12), layout [4*dir + (u,v,w,T)]
__device__ __forceinline__ void visc_flux_adjoint(const double* P, const double* G, const double* m, const double* fb,
                                                  double mu, double lambda, double* uvw_bar, double* Gbar)
{
    const double s23 = 2.0 / 3.0, s43 = 4.0 / 3.0;
    const double *gx = G, *gy = G + 4, *gz = G + 8;
    double txx = mu * (s43 * gx[0] - s23 * gy[1] - s23 * gz[2]);
    double tyy = mu * (-s23 * gx[0] + s43 * gy[1] - s23 * gz[2]);
    double tzz = mu * (-s23 * gx[0] - s23 * gy[1] + s43 * gz[2]);
    double txy = mu * (gy[0] + gx[1]);
    double txz = mu * (gz[0] + gx[2]);
    double tyz = mu * (gz[1] + gy[2]);
    double u = P[1], v = P[2], w = P[3];
    double f1 = fb[1], f2 = fb[2], f3 = fb[3], f4 = fb[4];
    double bxx = -m[0] * (f1 + u * f4);
    double byy = -m[1] * (f2 + v * f4);
    double bzz = -m[2] * (f3 + w * f4);
    double bxy = -m[1] * f1 - m[0] * f2 - (m[0] * v + m[1] * u) * f4;
    double bxz = -m[2] * f1 - m[0] * f3 - (m[0] * w + m[2] * u) * f4;
    double byz = -m[2] * f2 - m[1] * f3 - (m[1] * w + m[2] * v) * f4;
    uvw_bar[0] += -f4 * (m[0] * txx + m[1] * txy + m[2] * txz);
    uvw_bar[1] += -f4 * (m[0] * txy + m[1] * tyy + m[2] * tyz);
    uvw_bar[2] += -f4 * (m[0] * txz + m[1] * tyz + m[2] * tzz);
    Gbar[0] += mu * (s43 * bxx - s23 * byy - s23 * bzz);   // u_x
    Gbar[5] += mu * (-s23 * bxx + s43 * byy - s23 * bzz);  // v_y
    Gbar[10] += mu * (-s23 * bxx - s23 * byy + s43 * bzz); // w_z
    Gbar[4] += mu * bxy;                                   // u_y
    Gbar[1] += mu * bxy;                                   // v_x
    Gbar[8] += mu * bxz;                                   // u_z
    Gbar[2] += mu * bxz;                                   // w_x
    Gbar[9] += mu * byz;                                   // v_z
    Gbar[6] += mu * byz;                                   // w_y
    Gbar[3] += -lambda * m[0] * f4;                        // T_x
    Gbar[7] += -lambda * m[1] * f4;                        // T_y
    Gbar[11] += -lambda * m[2] * f4;                       // T_z
}

// weight of one argument of max(a, b) in the derivative: 1 if it wins, 0 if it loses, 1/2 on a tie (the
// symmetric subgradient, which is what a central difference measures).  Arguments within 1e-12 relative of
// each other count as tied: there the forward's choice depends on rounding alone (e.g. the sound speeds of the
// two traces of a state with continuous density and pressure).
__device__ __forceinline__ double max_weight(double mine, double other)
{
    double d = mine - other, tol = 1e-12 * fmax(fabs(mine), fabs(other));
    return d > tol ? 1.0 : (d < -tol ? 0.0 : 0.5);
}

__device__ __forceinline__ double sign0(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// adjoint of the Lax-Friedrichs flux with the PI central part (Riemann_LF riemann.f90:707-733 with
// SplitSurfaceFluxPI splitflux.f90:735-761, in the normal frame of the driver riemann.f90:216-289) with respect
// to the state of one side (side 0: left/master, 1: right/slave).  Fb: cotangent of the Cartesian flux.
__device__ __forceinline__ void lf_pi_adjoint(const double* UL, const double* UR, const double* nv, const double* t1,
                                              const double* t2, const double* Fb, double kappa, int side,
                                              PrimBar& pb)
{
    const double* Us[2] = {UL, UR};
    double rho[2], vn[2], v1[2], v2[2], p[2], H[2], c[2], E[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double* c5 = Us[s];
        double sr = 1.0 / c5[0];
        double u = c5[1] * sr, v = c5[2] * sr, w = c5[3] * sr;
        rho[s] = c5[0];
        E[s] = c5[4];
        p[s] = (kappa - 1.0) * (c5[4] - 0.5 * (c5[1] * u + c5[2] * v + c5[3] * w));
        H[s] = (c5[4] + p[s]) * sr;
        c[s] = sqrt(kappa * p[s] * sr);
        vn[s] = u * nv[0] + v * nv[1] + w * nv[2];
        v1[s] = u * t1[0] + v * t1[1] + w * t1[2];
        v2[s] = u * t2[0] + v * t2[1] + w * t2[2];
    }
    // cotangent in the normal frame: F = (F'0, F'1 n + F'2 t1 + F'3 t2, F'4)
    double Fp[5] = {Fb[0], Fb[1] * nv[0] + Fb[2] * nv[1] + Fb[3] * nv[2], Fb[1] * t1[0] + Fb[2] * t1[1] + Fb[3] * t1[2],
                    Fb[1] * t2[0] + Fb[2] * t2[1] + Fb[3] * t2[2], Fb[4]};
    double sr = rho[0] + rho[1], su = vn[0] + vn[1], sv = v1[0] + v1[1], sw = v2[0] + v2[1], sH = H[0] + H[1];
    // central part: F'0 = sr su/4, F'1 = sr su^2/8 + sp/2, F'2 = sr su sv/8, F'3 = sr su sw/8, F'4 = sr su sH/8
    double brho = 0.25 * su * Fp[0] + 0.125 * su * (su * Fp[1] + sv * Fp[2] + sw * Fp[3] + sH * Fp[4]);
    double bvn = 0.25 * sr * Fp[0] + 0.125 * sr * (2.0 * su * Fp[1] + sv * Fp[2] + sw * Fp[3] + sH * Fp[4]);
    double bv1 = 0.125 * sr * su * Fp[2], bv2 = 0.125 * sr * su * Fp[3];
    double bp = 0.5 * Fp[1], bH = 0.125 * sr * su * Fp[4], bE = 0.0;
    // dissipation -lmax/2 (U'_R - U'_L), U' = (rho, rho vn, rho v1, rho v2, E)
    double lmax = fmax(fabs(vn[1]), fabs(vn[0])) + fmax(c[0], c[1]);
    double dU[5] = {rho[1] - rho[0], rho[1] * vn[1] - rho[0] * vn[0], rho[1] * v1[1] - rho[0] * v1[0],
                    rho[1] * v2[1] - rho[0] * v2[0], E[1] - E[0]};
    double blmax = 0.0;
#pragma unroll
    for (int k = 0; k < 5; k++) blmax += -0.5 * Fp[k] * dU[k];
    const double sg = side == 1 ? -0.5 * lmax : 0.5 * lmax;
    brho += sg * (Fp[0] + Fp[1] * vn[side] + Fp[2] * v1[side] + Fp[3] * v2[side]);
    bvn += sg * Fp[1] * rho[side];
    bv1 += sg * Fp[2] * rho[side];
    bv2 += sg * Fp[3] * rho[side];
    bE += sg * Fp[4];
    // lmax = max(|vn_R|, |vn_L|) + max(c_L, c_R); c = sqrt(kappa p / rho)
    double wv = max_weight(fabs(vn[side]), fabs(vn[1 - side]));
    double wc = max_weight(c[side], c[1 - side]);
    bvn += blmax * wv * sign0(vn[side]);
    bp += blmax * wc * kappa / (2.0 * c[side] * rho[side]);
    brho += -blmax * wc * c[side] / (2.0 * rho[side]);
    pb.rho += brho;
    pb.u += bvn * nv[0] + bv1 * t1[0] + bv2 * t2[0];
    pb.v += bvn * nv[1] + bv1 * t1[1] + bv2 * t2[1];
    pb.w += bvn * nv[2] + bv1 * t1[2] + bv2 * t2[2];
    pb.p += bp;
    pb.H += bH;
    pb.E += bE;
}

// surface lifting weight of a local side: L_Hat_minus(0) for the xi-, eta-, zeta- sides, L_Hat_plus(N) otherwise
__device__ __forceinline__ double lhat_of(int ls, double LHm, double LHp)
{
    return (ls == 1 || ls == 2 || ls == 5) ? LHm : LHp;
}

// ---------------------------------------------------------------------------------------------------------
// k_adj_pre: lam = kbar + bdt*Ubar (kbar updated in place when Ubar is given), mu traces
//   mutr [side][2 slots][5][n2] = L_Hat(locSide) * (-sJ * lam) at the face nodes of the writer
// ---------------------------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(Dm<N>::block)
    k_adj_pre(double* __restrict__ kbar, const double* __restrict__ Ubar, double bdt, Mesh mesh, double LHm,
              double LHp, double* __restrict__ mutr)
{
    using D = Dm<N>;
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3;
    __shared__ double s_l[5 * n3];
    const int e = blockIdx.x, tid = threadIdx.x;
    if (tid < n3) {
        double* kb = kbar + (long)e * 5 * n3 + tid;
        if (Ubar) {
            const double* ub = Ubar + (long)e * 5 * n3 + tid;
#pragma unroll
            for (int v = 0; v < 5; v++) {
                double l = kb[v * n3] + bdt * ub[v * n3];
                kb[v * n3] = l;
                s_l[v * n3 + tid] = l;
            }
        } else {
#pragma unroll
            for (int v = 0; v < 5; v++) s_l[v * n3 + tid] = kb[v * n3];
        }
    }
    __syncthreads();
    for (int t = tid; t < D::nf; t += D::block) {
        int ls = t / n2 + 1, ab = t - (ls - 1) * n2, a = ab % n, b = ab / n;
        int node = face_vol_node<N>(ls, a, b);
        int c = mesh.conn[e * 6 + ls - 1];
        long side = c >> 3;
        int flip = c & 7;
        int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
        double f = -(lhat_of(ls, LHm, LHp) * mesh.geo[(long)e * 10 * n3 + 9 * n3 + node]);
        double* o = mutr + ((side * 2 + (flip ? 1 : 0)) * 5) * n2 + pq;
#pragma unroll
        for (int v = 0; v < 5; v++) o[v * n2] = f * s_l[v * n3 + node];
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_adj_main (split PI volume flux, LF surface flux, FP64)
//   in : U, U trace records utr (of this U), lam, mu traces; accumulate: Ubar += result, else Ubar = result;
//        kscale: if set, lam = a*lam at the end (the LSERK register adjoint, in place)
//   out: Ubar, lam, nghtr [side][2][4][n2] = L_Hat(locSide) * (n_master . sJ*Gbar) of the writer
// ---------------------------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(Dm<N>::block)
    k_adj_main(const double* __restrict__ U, Mesh mesh, const double* __restrict__ utr, double* __restrict__ lam,
               const double* __restrict__ mutr, const double* __restrict__ mats, GasD gd, double LHm, double LHp,
               double* __restrict__ Ubar, int accumulate, int kscale, double a, double* __restrict__ nghtr)
{
    using D = Dm<N>;
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3;
    constexpr int w_a = 4 * n3 + 6 * 7 * n2;  // forward lifting
    constexpr int w_c = 9 * n3;               // split adjoint partner data
    constexpr int w_d = 6 * 17 * n2;          // face adjoint results
    constexpr int w_e = 4 * n3;               // lifting adjoint contraction
    constexpr int w_ac = w_a > w_c ? w_a : w_c;
    constexpr int w_de = w_d > w_e ? w_d : w_e;
    constexpr int lds_w = w_ac > w_de ? w_ac : w_de;
    __shared__ double s_w[lds_w];
    __shared__ double s_mu[5 * n3];
    __shared__ double s_g[6 * 12 * n2];
    __shared__ Mats<N, double> M;
    const int e = blockIdx.x, tid = threadIdx.x;
    const bool active = tid < n3;
    const int i = tid % n, j = (tid / n) % n, k = tid / n2;
    load_mats<N, double>(M, mats, tid, D::block);
    const Gas<double> gas = make_gas<double>(gd);
    double uc[5], pr[6] = {1, 0, 0, 0, 1, 1}, Hme = 0.0, cU[5] = {1, 0, 0, 0, 1};
    const double* ge = mesh.geo + (long)e * 10 * n3 + tid;
    if (active) {
        load_state<N, double>(e, tid, U, gd, uc, pr, Hme);
        const double* u = U + (long)e * 5 * n3 + tid;
#pragma unroll
        for (int v = 0; v < 5; v++) cU[v] = u[v * n3];
        s_w[tid] = pr[1];
        s_w[n3 + tid] = pr[2];
        s_w[2 * n3 + tid] = pr[3];
        s_w[3 * n3 + tid] = pr[5];
        double msJ = -ge[9 * n3];
        const double* l = lam + (long)e * 5 * n3 + tid;
#pragma unroll
        for (int v = 0; v < 5; v++) s_mu[v * n3 + tid] = msJ * l[v * n3];
    }
    __syncthreads();
    PrimBar pb;
    prim_bar_zero(pb);
    double Gb[12];
#pragma unroll
    for (int q = 0; q < 12; q++) Gb[q] = 0.0;
    {
        double g[12];
        lift_gradient<N, double, double>(e, tid, mesh, utr, M, s_w, s_w + 4 * n3, gas, LHm, LHp, g);
        if (active) {
#pragma unroll
            for (int ls = 1; ls <= 6; ls++) {
                int ab;
                if (!on_face<N>(ls, i, j, k, ab)) continue;
                double* o = s_g + (ls - 1) * 12 * n2 + ab;
#pragma unroll
                for (int q = 0; q < 12; q++) o[q * n2] = g[q];
            }
            // adjoint of the weak viscous volume term (VolInt_weakForm_Visc, volint.f90:60-119):
            // X(i) += sum_l D_Hat_T(l,i) f_d(l)  ->  fbar_d(l) = sum_m D_Hat_T(l,m) mu(m) along the d-line
#pragma unroll 1
            for (int d = 0; d < 3; d++) {
                int idx = d == 0 ? i : (d == 1 ? j : k);
                int stride = d == 0 ? 1 : (d == 1 ? n : n2);
                int base = tid - idx * stride;
                double fb[5] = {0, 0, 0, 0, 0};
                for (int m = 0; m <= N; m++) {
                    double dh = M.DhT[idx + n * m];
#pragma unroll
                    for (int v = 1; v < 5; v++) fb[v] += dh * s_mu[v * n3 + base + m * stride];
                }
                double md[3] = {ge[(3 * d) * n3], ge[(3 * d + 1) * n3], ge[(3 * d + 2) * n3]};
                double uvw[3] = {0, 0, 0};
                visc_flux_adjoint(pr, g, md, fb, gd.mu, gd.lambda, uvw, Gb);
                pb.u += uvw[0];
                pb.v += uvw[1];
                pb.w += uvw[2];
            }
        }
    }
    __syncthreads();
    // ---- adjoint of the split volume flux (VolInt_splitForm_Kernel): X(i) += sum_l DVS(l,i) F#(i,l) with F#
    // symmetric, so U_i receives dF#(i,l)/dU_i^T [DVS(l,i) mu_i + DVS(i,l) mu_l] ----
    {
        double* s_v = s_w;           // rho,u,v,w,p,H
        double* s_m = s_w + 6 * n3;  // metric of the current direction
        if (active) {
            s_v[tid] = pr[0];
            s_v[n3 + tid] = pr[1];
            s_v[2 * n3 + tid] = pr[2];
            s_v[3 * n3 + tid] = pr[3];
            s_v[4 * n3 + tid] = pr[4];
            s_v[5 * n3 + tid] = Hme;
        }
#pragma unroll 1
        for (int d = 0; d < 3; d++) {
            double mme[3] = {0, 0, 0};
            if (active) {
                mme[0] = ge[(3 * d) * n3];
                mme[1] = ge[(3 * d + 1) * n3];
                mme[2] = ge[(3 * d + 2) * n3];
                s_m[tid] = mme[0];
                s_m[n3 + tid] = mme[1];
                s_m[2 * n3 + tid] = mme[2];
            }
            __syncthreads();
            if (active) {
                int idx = d == 0 ? i : (d == 1 ? j : k);
                int stride = d == 0 ? 1 : (d == 1 ? n : n2);
                int base = tid - idx * stride;
                for (int l = 0; l <= N; l++) {
                    if (l == idx) continue;
                    int o = base + l * stride;
                    double dli = M.DVS[l + n * idx], dil = M.DVS[idx + n * l];
                    double wv[5];
#pragma unroll
                    for (int v = 0; v < 5; v++) wv[v] = dli * s_mu[v * n3 + tid] + dil * s_mu[v * n3 + o];
                    double sr = pr[0] + s_v[o], su = pr[1] + s_v[n3 + o], sv = pr[2] + s_v[2 * n3 + o],
                           sw = pr[3] + s_v[3 * n3 + o], sH = Hme + s_v[5 * n3 + o];
                    double m0 = mme[0] + s_m[o], m1 = mme[1] + s_m[n3 + o], m2 = mme[2] + s_m[2 * n3 + o];
                    double ms = m0 * su + m1 * sv + m2 * sw;
                    double q = 0.125 * sr * ms;
                    double A = 2.0 * wv[0] + su * wv[1] + sv * wv[2] + sw * wv[3] + sH * wv[4];
                    double c8 = 0.125 * sr * A;
                    pb.rho += 0.125 * ms * A;
                    pb.u += c8 * m0 + q * wv[1];
                    pb.v += c8 * m1 + q * wv[2];
                    pb.w += c8 * m2 + q * wv[3];
                    pb.p += 0.5 * (m0 * wv[1] + m1 * wv[2] + m2 * wv[3]);
                    pb.H += q * wv[4];
                }
            }
            __syncthreads();
        }
    }
    // ---- adjoint of the face fluxes (FillFlux fillflux.f90:45-172 + SurfInt_Device surfint.t90:352-586):
    // X(master node) += L_Hat F, X(slave node) -= L_Hat F  ->  Fbar = L_Hat mu_master - L_Hat mu_slave ----
    double* s_r = s_w;  // [6][17][n2]: Ubar(5), Gbar(12) of the own face node
    for (int t = tid; t < D::nf; t += D::block) {
        int ls = t / n2 + 1, ab = t - (ls - 1) * n2, a2 = ab % n, b2 = ab / n;
        int node = face_vol_node<N>(ls, a2, b2);
        int c = mesh.conn[e * 6 + ls - 1];
        long side = c >> 3;
        int flip = c & 7;
        int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
        const long nslot = side * 2 + (flip ? 0 : 1);
        double Uo[5], Un[5], Go[12];
        const double* uo = U + (long)e * 5 * n3 + node;
#pragma unroll
        for (int v = 0; v < 5; v++) Uo[v] = uo[v * n3];
        {
            // neighbour state from its trace record (rho, rho u, rho v, rho w, p)
            const double* tr = utr + nslot * 5 * n2 + pq;
            Un[0] = tr[0];
            Un[1] = tr[n2];
            Un[2] = tr[2 * n2];
            Un[3] = tr[3 * n2];
            Un[4] = tr[4 * n2] / (gd.kappa - 1.0) + 0.5 * (Un[1] * Un[1] + Un[2] * Un[2] + Un[3] * Un[3]) / Un[0];
        }
#pragma unroll
        for (int q = 0; q < 12; q++) Go[q] = s_g[(ls - 1) * 12 * n2 + q * n2 + ab];
        const double* fg = mesh.fgeo + side * 10 * n2 + pq;
        double nv[3] = {fg[0], fg[n2], fg[2 * n2]}, t1[3] = {fg[3 * n2], fg[4 * n2], fg[5 * n2]},
               t2[3] = {fg[6 * n2], fg[7 * n2], fg[8 * n2]}, se = fg[9 * n2];
        const double LHo = lhat_of(ls, LHm, LHp);
        const double* mn = mutr + nslot * 5 * n2 + pq;
        double A[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            double mo = LHo * s_mu[v * n3 + node], mnb = mn[v * n2];
            A[v] = se * (flip ? mnb - mo : mo - mnb);
        }
        PrimBar fp;
        prim_bar_zero(fp);
        lf_pi_adjoint(flip ? Un : Uo, flip ? Uo : Un, nv, t1, t2, A, gd.kappa, flip ? 1 : 0, fp);
        // ViscousFlux (riemann.f90:638-700): 0.5*F_v(own).n of the own side
        double Po[4];
        {
            double sr = 1.0 / Uo[0];
            Po[0] = Uo[0];
            Po[1] = Uo[1] * sr;
            Po[2] = Uo[2] * sr;
            Po[3] = Uo[3] * sr;
        }
        double hv[5] = {0.0, 0.5 * A[1], 0.5 * A[2], 0.5 * A[3], 0.5 * A[4]};
        double uvw[3] = {0, 0, 0}, Gbo[12];
#pragma unroll
        for (int q = 0; q < 12; q++) Gbo[q] = 0.0;
        visc_flux_adjoint(Po, Go, nv, hv, gd.mu, gd.lambda, uvw, Gbo);
        fp.u += uvw[0];
        fp.v += uvw[1];
        fp.w += uvw[2];
        double Ub[5];
        prim_bar_to_cons(Uo, fp, gd.kappa, gd.R, Ub);
        double* o = s_r + (ls - 1) * 17 * n2 + ab;
#pragma unroll
        for (int v = 0; v < 5; v++) o[v * n2] = Ub[v];
#pragma unroll
        for (int q = 0; q < 12; q++) o[(5 + q) * n2] = Gbo[q];
    }
    __syncthreads();
    double Ubf[5] = {0, 0, 0, 0, 0};
    double qb[4] = {0, 0, 0, 0};
    if (active) {
#pragma unroll
        for (int ls = 1; ls <= 6; ls++) {
            int ab;
            if (!on_face<N>(ls, i, j, k, ab)) continue;
            const double* o = s_r + (ls - 1) * 17 * n2 + ab;
#pragma unroll
            for (int v = 0; v < 5; v++) Ubf[v] += o[v * n2];
#pragma unroll
            for (int q = 0; q < 12; q++) Gb[q] += o[(5 + q) * n2];
        }
        // ghat = sJ*Gbar: adjoint of g = sJ*(...) in lift_gradient
        double sJ = ge[9 * n3];
#pragma unroll
        for (int q = 0; q < 12; q++) Gb[q] *= sJ;
        // BR1 surface term (Lifting_FillFlux lifting_fillflux.t90:39-93): g_own += sJ L_Hat SurfElem/2 (q_slave -
        // q_master) n_master.  Own part here, the neighbour's part in k_adj_face from the traces written here.
#pragma unroll
        for (int ls = 1; ls <= 6; ls++) {
            int ab;
            if (!on_face<N>(ls, i, j, k, ab)) continue;
            int c = mesh.conn[e * 6 + ls - 1];
            long side = c >> 3;
            int flip = c & 7;
            int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
            const double* fg = mesh.fgeo + side * 10 * n2 + pq;
            double nv[3] = {fg[0], fg[n2], fg[2 * n2]}, se = fg[9 * n2];
            double LH = lhat_of(ls, LHm, LHp);
            double f = flip ? 0.5 * se : -0.5 * se;
            double* o = nghtr + ((side * 2 + (flip ? 1 : 0)) * 4) * n2 + pq;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                double val = LH * (nv[0] * Gb[q] + nv[1] * Gb[4 + q] + nv[2] * Gb[8 + q]);
                o[q * n2] = val;
                qb[q] += f * val;
            }
        }
    }
    __syncthreads();
    // BR1 volume term (lifting_volint.t90:262-328): grad_d q(i) = sum_l D(i,l) q(l) contracted with Ja^d(i)
    //  ->  qbar(l) += sum_i D(i,l) (Ja^d(i) . ghat(i)) along each line
    double* s_a = s_w;  // [4][n3]
#pragma unroll 1
    for (int d = 0; d < 3; d++) {
        if (active) {
            double m0 = ge[(3 * d) * n3], m1 = ge[(3 * d + 1) * n3], m2 = ge[(3 * d + 2) * n3];
#pragma unroll
            for (int q = 0; q < 4; q++) s_a[q * n3 + tid] = m0 * Gb[q] + m1 * Gb[4 + q] + m2 * Gb[8 + q];
        }
        __syncthreads();
        if (active) {
            int idx = d == 0 ? i : (d == 1 ? j : k);
            int stride = d == 0 ? 1 : (d == 1 ? n : n2);
            int base = tid - idx * stride;
            for (int m = 0; m <= N; m++) {
                double dd = M.DT[idx + n * m];  // D(m, idx)
#pragma unroll
                for (int q = 0; q < 4; q++) qb[q] += dd * s_a[q * n3 + base + m * stride];
            }
        }
        __syncthreads();
    }
    if (!active) return;
    pb.u += qb[0];
    pb.v += qb[1];
    pb.w += qb[2];
    pb.T += qb[3];
    double Ubv[5];
    prim_bar_to_cons(cU, pb, gd.kappa, gd.R, Ubv);
    double* ub = Ubar + (long)e * 5 * n3 + tid;
    double* lp = lam + (long)e * 5 * n3 + tid;
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double r = Ubv[v] + Ubf[v];
        ub[v * n3] = accumulate ? ub[v * n3] + r : r;
        if (kscale) lp[v * n3] = a * lp[v * n3];
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_adj_face: the neighbour's BR1 surface term: qbar_own += (+-)SurfElem/2 * (L_Hat n.ghat)_neighbour on the
// face nodes (minus for the master side), chained through ConsToPrim into Ubar
// ---------------------------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(Dm<N>::block)
    k_adj_face(const double* __restrict__ U, Mesh mesh, const double* __restrict__ nghtr, GasD gd,
               double* __restrict__ Ubar)
{
    using D = Dm<N>;
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3;
    __shared__ double s_f[6 * 4 * n2];
    const int e = blockIdx.x, tid = threadIdx.x;
    for (int t = tid; t < D::nf; t += D::block) {
        int ls = t / n2 + 1, ab = t - (ls - 1) * n2;
        int c = mesh.conn[e * 6 + ls - 1];
        long side = c >> 3;
        int flip = c & 7;
        int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
        double se = mesh.fgeo[side * 10 * n2 + 9 * n2 + pq];
        double f = flip ? 0.5 * se : -0.5 * se;
        const double* nb = nghtr + ((side * 2 + (flip ? 0 : 1)) * 4) * n2 + pq;
#pragma unroll
        for (int q = 0; q < 4; q++) s_f[(ls - 1) * 4 * n2 + q * n2 + ab] = f * nb[q * n2];
    }
    __syncthreads();
    if (tid >= n3) return;
    const int i = tid % n, j = (tid / n) % n, k = tid / n2;
    if (i != 0 && i != N && j != 0 && j != N && k != 0 && k != N) return;
    double qb[4] = {0, 0, 0, 0};
#pragma unroll
    for (int ls = 1; ls <= 6; ls++) {
        int ab;
        if (!on_face<N>(ls, i, j, k, ab)) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) qb[q] += s_f[(ls - 1) * 4 * n2 + q * n2 + ab];
    }
    const double* u = U + (long)e * 5 * n3 + tid;
    double c[5];
#pragma unroll
    for (int v = 0; v < 5; v++) c[v] = u[v * n3];
    PrimBar pb;
    prim_bar_zero(pb);
    pb.u = qb[0];
    pb.v = qb[1];
    pb.w = qb[2];
    pb.T = qb[3];
    double Ub[5];
    prim_bar_to_cons(c, pb, gd.kappa, gd.R, Ub);
    double* o = Ubar + (long)e * 5 * n3 + tid;
#pragma unroll
    for (int v = 0; v < 5; v++) o[v * n3] += Ub[v];
}

}  // namespace galx
