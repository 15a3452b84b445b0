// galx_adjoint.hpp -- discrete adjoint (vector-Jacobian product) of the split-form DG operator on gfx950.
//
// For Ut = R(U) (split Pirozzoli volume flux, Lax-Friedrichs surface flux with the PI central part, BR1 lifting,
// constant viscosity, FP64) and a cotangent lam = dJ/dUt, the kernels form Ubar = (dR/dU)^T lam exactly (to
// rounding), element per workgroup and without atomics: every element gathers what its own nodes receive.
//   k_adj_pre  : lam = akbar*kbar + bdt*Ubar (LSERK stage adjoint, in place; akbar = the previous stage's register
//                coefficient, kbar = a*lam of the stage after, applied here instead of by a pass of its own), mu =
//                -sJ*lam on the faces scaled by the
//                surface lifting weight L_Hat of the writer's local side -> mu traces; in a sweep it first adds
//                the pending neighbour lifting term of the previous VJP (so k_adj_face runs once per sweep)
//   k_adj_main : forward recompute (primitives, lifted gradients), adjoint of the weak viscous volume term
//                (D_Hat^T transposed), of the split volume fluxes (two-point flux Jacobians transposed and
//                gathered along the lines), of the face fluxes (transposed LF/PI and viscous Jacobians), of the
//                BR1 lifting (transposed strong-form derivative, own surface term), chained through ConsToPrim;
//                Ubar (+)= result and the traces L_Hat * n.ghat (ghat = sJ*Gbar) for the
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

// transposed Jacobian of ConsToPrim (eos.f90:212-239) with H = (E+p)/rho and T = p/(rho*R); rR = 1/R
__device__ __forceinline__ void prim_bar_to_cons(const double* c, const PrimBar& b, double kappa, double rR,
                                                 double* Ubar)
{
    double rho = c[0], sRho = rcp64(rho);
    double u = c[1] * sRho, v = c[2] * sRho, w = c[3] * sRho;
    double q2 = 0.5 * (u * u + v * v + w * w);
    double km1 = kappa - 1.0;
    double p = km1 * (c[4] - rho * q2);
    double H = (c[4] + p) * sRho, T = p * sRho * rR;
    double dp_drho = km1 * q2;
    Ubar[0] = b.rho - (b.u * u + b.v * v + b.w * w) * sRho + b.p * dp_drho + b.H * (dp_drho - H) * sRho +
              b.T * (dp_drho * sRho * rR - T * sRho);
    double s = -km1 * (b.p + b.H * sRho + b.T * sRho * rR);
    Ubar[1] = b.u * sRho + s * u;
    Ubar[2] = b.v * sRho + s * v;
    Ubar[3] = b.w * sRho + s * w;
    Ubar[4] = b.E + b.p * km1 + b.H * kappa * sRho + b.T * km1 * sRho * rR;
}

// Adjoint of the viscous flux -m.(tau, tau.u + lambda grad T) (EvalNormalDiffFlux3D flux.f90:779-847 for m = n,
// EvalTransformedDiffFlux3D flux.f90:619-700 for m = Ja^d) for the flux cotangent (f1,f2,f3,f4), in two parts:
//   visc_tau_adjoint : the velocity cotangent -f4 * (m.tau), from the stress tau of the forward state
//   visc_grad_adjoint: the gradient cotangent, accumulated in reduced form (b: the symmetric stress cotangent,
//                      t: m*f4 of the heat flux); visc_grad_expand maps the sum to Gbar(12), layout
//                      [4*dir + (u,v,w,T)].  The map is linear, so sums of contributions expand once.
struct ViscGradBar {
    double bxx, byy, bzz, bxy, bxz, byz, tx, ty, tz;
};

__device__ __forceinline__ void visc_grad_zero(ViscGradBar& g)
{
    g.bxx = g.byy = g.bzz = g.bxy = g.bxz = g.byz = g.tx = g.ty = g.tz = 0.0;
}

// tau = (txx, tyy, tzz, txy, txz, tyz)
__device__ __forceinline__ void visc_tau_adjoint(const double* tau, const double* m, double f4, double* uvw_bar)
{
    uvw_bar[0] += -f4 * (m[0] * tau[0] + m[1] * tau[3] + m[2] * tau[4]);
    uvw_bar[1] += -f4 * (m[0] * tau[3] + m[1] * tau[1] + m[2] * tau[5]);
    uvw_bar[2] += -f4 * (m[0] * tau[4] + m[1] * tau[5] + m[2] * tau[2]);
}

__device__ __forceinline__ void visc_grad_adjoint(double u, double v, double w, const double* m, double f1, double f2,
                                                  double f3, double f4, ViscGradBar& g)
{
    g.bxx += -m[0] * (f1 + u * f4);
    g.byy += -m[1] * (f2 + v * f4);
    g.bzz += -m[2] * (f3 + w * f4);
    g.bxy += -m[1] * f1 - m[0] * f2 - (m[0] * v + m[1] * u) * f4;
    g.bxz += -m[2] * f1 - m[0] * f3 - (m[0] * w + m[2] * u) * f4;
    g.byz += -m[2] * f2 - m[1] * f3 - (m[1] * w + m[2] * v) * f4;
    g.tx += m[0] * f4;
    g.ty += m[1] * f4;
    g.tz += m[2] * f4;
}

__device__ __forceinline__ void visc_grad_expand(const ViscGradBar& g, double mu, double lambda, double* Gbar)
{
    const double s23 = 2.0 / 3.0, s43 = 4.0 / 3.0;
    Gbar[0] = mu * (s43 * g.bxx - s23 * g.byy - s23 * g.bzz);   // u_x
    Gbar[5] = mu * (-s23 * g.bxx + s43 * g.byy - s23 * g.bzz);  // v_y
    Gbar[10] = mu * (-s23 * g.bxx - s23 * g.byy + s43 * g.bzz); // w_z
    Gbar[4] = mu * g.bxy;                                       // u_y
    Gbar[1] = mu * g.bxy;                                       // v_x
    Gbar[8] = mu * g.bxz;                                       // u_z
    Gbar[2] = mu * g.bxz;                                       // w_x
    Gbar[9] = mu * g.byz;                                       // v_z
    Gbar[6] = mu * g.byz;                                       // w_y
    Gbar[3] = -lambda * g.tx;                                   // T_x
    Gbar[7] = -lambda * g.ty;                                   // T_y
    Gbar[11] = -lambda * g.tz;                                  // T_z
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
// SplitSurfaceFluxPI splitflux.f90:735-761; the Riemann routine riemann.f90:216-289 rotates into the normal frame and back,
// which leaves the flux unchanged, so the Cartesian form of the forward kernel is differentiated) with respect to
// the state of one side (side 0: left/master, 1: right/slave).  Fb: cotangent of the Cartesian flux:
//   F = (2q, q (u_L + u_R) + (p_L + p_R)/2 n, q (H_L + H_R)) - lmax/2 (U_R - U_L),
//   q = (rho_L + rho_R)(vn_L + vn_R)/8, lmax = max(|vn_R|, |vn_L|) + max(c_L, c_R)
__device__ __forceinline__ void lf_pi_adjoint(const double* UL, const double* UR, const double* nv, const double* Fb,
                                              double kappa, int side, PrimBar& pb)
{
    const double* Us[2] = {UL, UR};
    double rho[2], vn[2], p[2], H[2], c[2], srs[2];
    double us[3];  // velocity of the differentiated side
    double usum[3] = {0, 0, 0};
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double* c5 = Us[s];
        double sr = rcp64(c5[0]);
        double u = c5[1] * sr, v = c5[2] * sr, w = c5[3] * sr;
        rho[s] = c5[0];
        srs[s] = sr;
        p[s] = (kappa - 1.0) * (c5[4] - 0.5 * (c5[1] * u + c5[2] * v + c5[3] * w));
        H[s] = (c5[4] + p[s]) * sr;
        c[s] = sqrt(kappa * p[s] * sr);
        vn[s] = u * nv[0] + v * nv[1] + w * nv[2];
        usum[0] += u;
        usum[1] += v;
        usum[2] += w;
        if (s == side) {
            us[0] = u;
            us[1] = v;
            us[2] = w;
        }
    }
    const double sr = rho[0] + rho[1], su = vn[0] + vn[1], q = 0.125 * sr * su;
    const double Fm = Fb[1] * usum[0] + Fb[2] * usum[1] + Fb[3] * usum[2];
    const double Fn = Fb[1] * nv[0] + Fb[2] * nv[1] + Fb[3] * nv[2];
    // central part
    const double qb = 2.0 * Fb[0] + Fm + (H[0] + H[1]) * Fb[4];
    double brho = 0.125 * su * qb;
    double bvn = 0.125 * sr * qb;
    double bu[3] = {q * Fb[1], q * Fb[2], q * Fb[3]};
    double bp = 0.5 * Fn;
    pb.H += q * Fb[4];
    // dissipation -lmax/2 (U_R - U_L)
    double blmax = 0.0;
#pragma unroll
    for (int v = 0; v < 5; v++) blmax += -0.5 * Fb[v] * (Us[1][v] - Us[0][v]);
    const double lmax = fmax(fabs(vn[1]), fabs(vn[0])) + fmax(c[0], c[1]);
    const double sg = side == 1 ? -0.5 * lmax : 0.5 * lmax;
    brho += sg * (Fb[0] + Fb[1] * us[0] + Fb[2] * us[1] + Fb[3] * us[2]);
#pragma unroll
    for (int d = 0; d < 3; d++) bu[d] += sg * Fb[1 + d] * rho[side];
    pb.E += sg * Fb[4];
    // lmax: the winning |vn| and c = sqrt(kappa p / rho)
    const double wv = max_weight(fabs(vn[side]), fabs(vn[1 - side]));
    const double wc = max_weight(c[side], c[1 - side]);
    bvn += blmax * wv * sign0(vn[side]);
    bp += blmax * wc * kappa * (0.5 * rcp64(c[side])) * srs[side];
    brho += -blmax * wc * c[side] * (0.5 * srs[side]);
    pb.rho += brho;
    pb.u += bu[0] + bvn * nv[0];
    pb.v += bu[1] + bvn * nv[1];
    pb.w += bu[2] + bvn * nv[2];
    pb.p += bp;
}

// surface lifting weight of a local side: L_Hat_minus(0) for the xi-, eta-, zeta- sides, L_Hat_plus(N) otherwise
__device__ __forceinline__ double lhat_of(int ls, double LHm, double LHp)
{
    return (ls == 1 || ls == 2 || ls == 5) ? LHm : LHp;
}

// ---------------------------------------------------------------------------------------------------------
// The neighbour's BR1 surface term of the lifting adjoint (Lifting_FillFlux lifting_fillflux.t90:39-93 transposed):
// qbar_own += (+-)SurfElem/2 * (L_Hat n.ghat)_neighbour on the face nodes (minus for the master side), chained
// through ConsToPrim at the state U of that VJP.  face_tasks fills s_f[6][4][n2]; face_node_ubar returns the
// conservative contribution of a node (zero for interior nodes).
// ---------------------------------------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void lift_face_tasks(int e, int tid, const Mesh& mesh, const double* __restrict__ nghtr,
                                                double* s_f)
{
    using D = Dm<N>;
    static_assert(D::node_per_thread, "one volume node per thread: N <= 9");
    constexpr int n2 = D::n2;
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
}

template <int N>
__device__ __forceinline__ bool on_any_face(int tid)
{
    constexpr int n = N + 1, n2 = n * n;
    const int i = tid % n, j = (tid / n) % n, k = tid / n2;
    return i == 0 || i == N || j == 0 || j == N || k == 0 || k == N;
}

// the conservative contribution at face node tid, from the state c(5) of that node
template <int N>
__device__ __forceinline__ void lift_face_node_ubar_at(int tid, const double* c, const GasD& gd, const double* s_f,
                                                       double* Ub)
{
    constexpr int n = N + 1, n2 = n * n;
    const int i = tid % n, j = (tid / n) % n, k = tid / n2;
    double qb[4] = {0, 0, 0, 0};
#pragma unroll
    for (int ls = 1; ls <= 6; ls++) {
        int ab;
        if (!on_face<N>(ls, i, j, k, ab)) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) qb[q] += s_f[(ls - 1) * 4 * n2 + q * n2 + ab];
    }
    PrimBar pb;
    prim_bar_zero(pb);
    pb.u = qb[0];
    pb.v = qb[1];
    pb.w = qb[2];
    pb.T = qb[3];
    prim_bar_to_cons(c, pb, gd.kappa, gd.rR, Ub);
}

template <int N>
__device__ __forceinline__ bool lift_face_node_ubar(int e, int tid, const double* __restrict__ U, const GasD& gd,
                                                    const double* s_f, double* Ub)
{
    constexpr int n3 = Dm<N>::n3;
    if (!on_any_face<N>(tid)) return false;
    const double* u = U + (long)e * 5 * n3 + tid;
    double c[5];
#pragma unroll
    for (int v = 0; v < 5; v++) c[v] = u[v * n3];
    lift_face_node_ubar_at<N>(tid, c, gd, s_f, Ub);
    return true;
}

// ---------------------------------------------------------------------------------------------------------
// k_adj_pre: lam = akbar*kbar + bdt*Ubar (kbar updated in place when Ubar is given), mu traces
//   mutr [side][2 slots][5][n2] = L_Hat(locSide) * (-sJ * lam) at the face nodes of the writer (the same
//   product, in the same order, as k_adj_main forms for its own side).
//   Uprev (optional, with Ubar): first completes the VJP of the stage before (in reverse order), evaluated at
//   Uprev, with its pending neighbour lifting term from nghtr (the work of k_adj_face, fused).
// ---------------------------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(Dm<N>::block)
    k_adj_pre(double* __restrict__ kbar, double* __restrict__ Ubar, double akbar, double bdt, Mesh mesh, double LHm,
              double LHp, double* __restrict__ mutr, const double* __restrict__ Uprev, const double* __restrict__ nghtr,
              GasD gd)
{
    using D = Dm<N>;
    static_assert(D::node_per_thread, "one volume node per thread: N <= 9");
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3;
    __shared__ double s_l[5 * n3];  // -sJ * lam
    __shared__ double s_f[6 * 4 * n2];
    const int e = blockIdx.x, tid = threadIdx.x;
    const bool active = tid < n3;
    const bool face = active && Uprev && Ubar && on_any_face<N>(tid);
    // the node's own loads first: none of them waits for the face phase
    double kb5[5] = {0, 0, 0, 0, 0}, u5[5] = {0, 0, 0, 0, 0}, up[5] = {1, 0, 0, 0, 1}, msJ = 0.0;
    const long off = (long)e * 5 * n3 + tid;
    if (active) {
#pragma unroll
        for (int v = 0; v < 5; v++) kb5[v] = kbar[off + v * n3];
        if (Ubar) {
#pragma unroll
            for (int v = 0; v < 5; v++) u5[v] = Ubar[off + v * n3];
        }
        if (face) {
#pragma unroll
            for (int v = 0; v < 5; v++) up[v] = Uprev[off + v * n3];
        }
        msJ = -mesh.geo[(long)e * 10 * n3 + 9 * n3 + tid];
    }
    if (Uprev) {
        lift_face_tasks<N>(e, tid, mesh, nghtr, s_f);
        __syncthreads();
    }
    if (active) {
        if (face) {
            double add[5];
            lift_face_node_ubar_at<N>(tid, up, gd, s_f, add);
#pragma unroll
            for (int v = 0; v < 5; v++) {
                u5[v] += add[v];
                Ubar[off + v * n3] = u5[v];
            }
        }
#pragma unroll
        for (int v = 0; v < 5; v++) {
            double l = Ubar ? akbar * kb5[v] + bdt * u5[v] : kb5[v];
            if (Ubar) kbar[off + v * n3] = l;
            s_l[v * n3 + tid] = msJ * l;
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
        const double LH = lhat_of(ls, LHm, LHp);
        double* o = mutr + ((side * 2 + (flip ? 1 : 0)) * 5) * n2 + pq;
#pragma unroll
        for (int v = 0; v < 5; v++) o[v * n2] = LH * s_l[v * n3 + node];
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_adj_main (split PI volume flux, LF surface flux, FP64)
//   in : U, U trace records utr (of this U), lam, mu traces; accumulate: Ubar += result, else Ubar = result
//   out: Ubar, nghtr [side][2][4][n2] = L_Hat(locSide) * (n_master . sJ*Gbar) of the writer
// Phases: split volume adjoint, forward lifting with the viscous volume adjoint, face adjoint, lifting adjoint.
// LDS: one work array reused by the phases, mu = -sJ*lam, the own stress at the face nodes and the matrices
// (76.5 KB at N=7: two workgroups per CU, with AdjWaves capping the registers to match).
// ---------------------------------------------------------------------------------------------------------
template <int N>
struct AdjLds {
    static constexpr int n2 = Dm<N>::n2, n3 = Dm<N>::n3;
    static constexpr int w_a = 9 * n3;               // split adjoint partner data: rho,u,v,w,p,H + metric
    static constexpr int w_b = 4 * Lay<N>::n3 + 6 * 7 * n2;  // forward lifting (padded layout); also >= 4 n3
    static constexpr int w_d = 6 * 12 * n2;          // face adjoint results
    static constexpr int w_ab = w_a > w_b ? w_a : w_b;
    static constexpr int w = w_ab > w_d ? w_ab : w_d;
    static constexpr int bytes = 8 * (w + 5 * n3 + 6 * 6 * n2 + 3 * Lay<N>::np * (N + 1));  // + Mats
};

// waves per SIMD that let two workgroups share a CU when their LDS fits (160 KB per CU); otherwise the
// compiler's choice
template <int N>
struct AdjWaves {
    static constexpr int block = Dm<N>::block;
    static constexpr int value = (block >= 256 && 2 * AdjLds<N>::bytes <= 160 * 1024) ? (2 * block / 64 + 3) / 4 : 1;
};

template <int N>
__global__ void __launch_bounds__(Dm<N>::block, (AdjWaves<N>::value))
    k_adj_main(const double* __restrict__ U, Mesh mesh, const double* __restrict__ utr, const double* __restrict__ lam,
               const double* __restrict__ mutr, const double* __restrict__ mats, GasD gd, double LHm, double LHp,
               double* __restrict__ Ubar, int accumulate, double* __restrict__ nghtr)
{
    using D = Dm<N>;
    static_assert(D::node_per_thread, "one volume node per thread: N <= 9");
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3;
    constexpr int np = Lay<N>::np, n3p = Lay<N>::n3;  // matrices and the lifting input: padded layout
    __shared__ double s_w[AdjLds<N>::w];
    __shared__ double s_mu[5 * n3];
    __shared__ double s_tau[6 * 6 * n2];  // own stress (txx,tyy,tzz,txy,txz,tyz) at the face nodes
    __shared__ Mats<N, double, false> M;  // column-major: the adjoint contractions read (idx, m)
    // the i- and j-lines inside one 64-lane wave (N = 1, 3, 7): shared pairs in the split adjoint, as in k_rhs
    constexpr bool pair_lines = 64 % n == 0 && 64 % n2 == 0;
    const int e = blockIdx.x, tid = threadIdx.x;
    const bool active = tid < n3;
    const int i = tid % n, j = (tid / n) % n, k = tid / n2;
    load_mats<N, double, false>(M, mats, tid, D::block);
    const Gas<double> gas = make_gas<double>(gd);
    double pr[6] = {1, 0, 0, 0, 1, 1};
    const double* ge = mesh.geo + (long)e * 10 * n3 + tid;
    // the node's metrics and sJ are read in each of the three phases that use them: read once into registers
    // instead, they made the VJP 17 % slower at N=7 on MI355X (the kernel, at 128 VGPRs, spilled 172 instead of 76
    // bytes per lane)
    PrimBar pb;
    prim_bar_zero(pb);
    // ---- adjoint of the split volume flux (VolInt_splitForm_Kernel): X(i) += sum_l DVS(l,i) F#(i,l) with F#
    // symmetric, so U_i receives dF#(i,l)/dU_i^T [DVS(l,i) mu_i + DVS(i,l) mu_l] ----
    {
        double* s_v = s_w;           // rho,u,v,w,p,H
        double* s_m = s_w + 6 * n3;  // metric of the current direction
        double Hme = 0.0;
        if (active) {
            double uc[5];
            load_state<N, double>(e, tid, U, gd, uc, pr, Hme);
            s_v[tid] = pr[0];
            s_v[n3 + tid] = pr[1];
            s_v[2 * n3 + tid] = pr[2];
            s_v[3 * n3 + tid] = pr[3];
            s_v[4 * n3 + tid] = pr[4];
            s_v[5 * n3 + tid] = Hme;
            double msJ = -ge[9 * n3];
            const double* l = lam + (long)e * 5 * n3 + tid;
#pragma unroll
            for (int v = 0; v < 5; v++) s_mu[v * n3 + tid] = msJ * l[v * n3];
        }
#pragma unroll 1
        for (int d = 0; d < 3; d++) {
            double mme[3] = {0, 0, 0};
            if (active) {
#pragma unroll
                for (int c = 0; c < 3; c++) mme[c] = ge[(3 * (d) + (c)) * n3];
                s_m[tid] = mme[0];
                s_m[n3 + tid] = mme[1];
                s_m[2 * n3 + tid] = mme[2];
            }
            line_sync<N>(d);  // s_v, s_mu and s_m are read along the lines of d
            if (active) {
                int idx = d == 0 ? i : (d == 1 ? j : k);
                int stride = d == 0 ? 1 : (d == 1 ? n : n2);
                int base = tid - idx * stride;
                // the increment of the primitive adjoint (rho,u,v,w,p,H) of pair (idx, l).  F# depends on the
                // two states only through their sums and the metric sums, so both nodes of the pair receive the
                // same increment (wv, and with it everything below, is symmetric in the pair)
                auto pair_bar = [&](int l, double* P) {
                    const int o = base + l * stride;
                    const double dli = M.dvs(l, idx), dil = M.dvs(idx, l);
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
                    P[0] = 0.125 * ms * A;
                    P[1] = c8 * m0 + q * wv[1];
                    P[2] = c8 * m1 + q * wv[2];
                    P[3] = c8 * m2 + q * wv[3];
                    P[4] = 0.5 * (m0 * wv[1] + m1 * wv[2] + m2 * wv[3]);
                    P[5] = q * wv[4];
                };
                auto add_bar = [&](const double* P) {
                    pb.rho += P[0];
                    pb.u += P[1];
                    pb.v += P[2];
                    pb.w += P[3];
                    pb.p += P[4];
                    pb.H += P[5];
                };
                if (pair_lines && d < 2) {
                    // lines inside one wave, as the forward's shared pairs: the lane at line position idx forms
                    // the pairs (idx, idx+s) for s = 1..(n-1)/2 (mod n) and takes (idx-s, idx) from the lane s
                    // positions back; for even n both ends form (idx, idx+n/2).  The diagonal is left out, as in
                    // the forward's shared pairs.
                    constexpr int msk = n - 1;
                    const int lane0 = (tid & 63) - idx * stride;  // node = thread: lane distance = node stride
#pragma unroll
                    for (int s = 1; s <= n / 2; s++) {
                        double P[6];
                        pair_bar((idx + s) & msk, P);
                        add_bar(P);
                        if (2 * s < n) {
                            const int src = lane0 + ((idx - s) & msk) * stride;
                            double R[6];
#pragma unroll
                            for (int c = 0; c < 6; c++) R[c] = __shfl(P[c], src);
                            add_bar(R);
                        }
                    }
                } else {
                    // l = idx included: F#(U,U) has equal partial derivatives in its two arguments, so the same
                    // formula gives 2 DVS(i,i) dF#/dU_i^T mu_i.  The forward leaves this pair out; DVS(i,i) is a
                    // rounding residue for Gauss-Lobatto, so the term is at rounding level either way, and the loop
                    // over all l spills 8 to 16 bytes per lane less than one that skips it.
                    for (int l = 0; l <= N; l++) {
                        double P[6];
                        pair_bar(l, P);
                        add_bar(P);
                    }
                }
            }
            line_sync<N>(d);
        }
    }
    // ---- forward lifting, then the adjoint of the weak viscous volume term (VolInt_weakForm_Visc,
    // volint.f90:60-119): X(i) += sum_l D_Hat_T(l,i) f_d(l)  ->  fbar_d(l) = sum_m D_Hat_T(l,m) mu(m) ----
    ViscGradBar gb;
    visc_grad_zero(gb);
    {
        if (active) {
            const int pt = Lay<N>::at(tid);  // the lifting reads its input in the padded node layout
            s_w[pt] = pr[1];
            s_w[n3p + pt] = pr[2];
            s_w[2 * n3p + pt] = pr[3];
            s_w[3 * n3p + pt] = pr[5];
        }
        __syncthreads();
        double g[12];
        lift_gradient<N, double, double>(e, tid, mesh, utr, M, s_w, s_w + 4 * n3p, gas, LHm, LHp, g);
        if (active) {
            Stress<double> st;
            stress<double>(pr, g, gas, st);
            const double tau[6] = {st.txx, st.tyy, st.tzz, st.txy, st.txz, st.tyz};
#pragma unroll
            for (int ls = 1; ls <= 6; ls++) {
                int ab;
                if (!on_face<N>(ls, i, j, k, ab)) continue;
                double* o = s_tau + (ls - 1) * 6 * n2 + ab;
#pragma unroll
                for (int q = 0; q < 6; q++) o[q * n2] = tau[q];
            }
            double uvw[3] = {0, 0, 0};
#pragma unroll 1
            for (int d = 0; d < 3; d++) {
                int idx = d == 0 ? i : (d == 1 ? j : k);
                int stride = d == 0 ? 1 : (d == 1 ? n : n2);
                int base = tid - idx * stride;
                double fb[5] = {0, 0, 0, 0, 0};
                for (int m = 0; m <= N; m++) {
                    double dh = M.dht(idx, m);
#pragma unroll
                    for (int v = 1; v < 5; v++) fb[v] += dh * s_mu[v * n3 + base + m * stride];
                }
                double md[3] = {ge[(3 * d + 0) * n3], ge[(3 * d + 1) * n3], ge[(3 * d + 2) * n3]};
                visc_tau_adjoint(tau, md, fb[4], uvw);
                visc_grad_adjoint(pr[1], pr[2], pr[3], md, fb[1], fb[2], fb[3], fb[4], gb);
            }
            pb.u += uvw[0];
            pb.v += uvw[1];
            pb.w += uvw[2];
        }
    }
    __syncthreads();
    // ---- adjoint of the face fluxes (FillFlux fillflux.f90:45-172 + SurfInt_Device surfint.t90:352-586):
    // X(master node) += L_Hat F, X(slave node) -= L_Hat F  ->  Fbar = L_Hat mu_master - L_Hat mu_slave ----
    double* s_r = s_w;  // [6][12][n2]: Ubar(5), the viscous flux cotangent (4) and the normal (3) of a face node
    for (int t = tid; t < D::nf; t += D::block) {
        int ls = t / n2 + 1, ab = t - (ls - 1) * n2, a2 = ab % n, b2 = ab / n;
        int node = face_vol_node<N>(ls, a2, b2);
        int c = mesh.conn[e * 6 + ls - 1];
        long side = c >> 3;
        int flip = c & 7;
        int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
        const long nslot = side * 2 + (flip ? 0 : 1);
        double Uo[5], Un[5];
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
            Un[4] = tr[4 * n2] * gd.rkm1 + 0.5 * (Un[1] * Un[1] + Un[2] * Un[2] + Un[3] * Un[3]) * rcp64(Un[0]);
        }
        const double* fg = mesh.fgeo + side * 10 * n2 + pq;
        double nv[3] = {fg[0], fg[n2], fg[2 * n2]}, se = fg[9 * n2];
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
        lf_pi_adjoint(flip ? Un : Uo, flip ? Uo : Un, nv, A, gd.kappa, flip ? 1 : 0, fp);
        // ViscousFlux (riemann.f90:638-700): 0.5*F_v(own).n of the own side; its velocity part here, its
        // gradient part by the node thread (which holds the velocities)
        double tau[6];
#pragma unroll
        for (int q = 0; q < 6; q++) tau[q] = s_tau[(ls - 1) * 6 * n2 + q * n2 + ab];
        double uvw[3] = {0, 0, 0};
        visc_tau_adjoint(tau, nv, 0.5 * A[4], uvw);
        fp.u += uvw[0];
        fp.v += uvw[1];
        fp.w += uvw[2];
        double Ub[5];
        prim_bar_to_cons(Uo, fp, gd.kappa, gd.rR, Ub);
        double* o = s_r + (ls - 1) * 12 * n2 + ab;
#pragma unroll
        for (int v = 0; v < 5; v++) o[v * n2] = Ub[v];
#pragma unroll
        for (int v = 1; v < 5; v++) o[(4 + v) * n2] = 0.5 * A[v];
#pragma unroll
        for (int d = 0; d < 3; d++) o[(9 + d) * n2] = nv[d];
    }
    __syncthreads();
    double Ubf[5] = {0, 0, 0, 0, 0};
    double qb[4] = {0, 0, 0, 0};
    double Gb[12];
    if (active) {
#pragma unroll
        for (int ls = 1; ls <= 6; ls++) {
            int ab;
            if (!on_face<N>(ls, i, j, k, ab)) continue;
            const double* o = s_r + (ls - 1) * 12 * n2 + ab;
#pragma unroll
            for (int v = 0; v < 5; v++) Ubf[v] += o[v * n2];
            double nv[3] = {o[9 * n2], o[10 * n2], o[11 * n2]};
            visc_grad_adjoint(pr[1], pr[2], pr[3], nv, o[5 * n2], o[6 * n2], o[7 * n2], o[8 * n2], gb);
        }
        // ghat = sJ*Gbar: adjoint of g = sJ*(...) in lift_gradient
        visc_grad_expand(gb, gd.mu, gd.lambda, Gb);
        double sJ = ge[9 * n3];
#pragma unroll
        for (int q = 0; q < 12; q++) Gb[q] *= sJ;
        // BR1 surface term (Lifting_FillFlux lifting_fillflux.t90:39-93): g_own += sJ L_Hat SurfElem/2 (q_slave -
        // q_master) n_master.  Own part here, the neighbour's part from the traces written here (k_adj_face or
        // the next k_adj_pre).
#pragma unroll
        for (int ls = 1; ls <= 6; ls++) {
            int ab;
            if (!on_face<N>(ls, i, j, k, ab)) continue;
            int c = mesh.conn[e * 6 + ls - 1];
            long side = c >> 3;
            int flip = c & 7;
            int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
            const double* o = s_r + (ls - 1) * 12 * n2 + ab;
            double nv[3] = {o[9 * n2], o[10 * n2], o[11 * n2]};
            double se = mesh.fgeo[side * 10 * n2 + 9 * n2 + pq];
            double LH = lhat_of(ls, LHm, LHp);
            double f = flip ? 0.5 * se : -0.5 * se;
            double* ot = nghtr + ((side * 2 + (flip ? 1 : 0)) * 4) * n2 + pq;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                double val = LH * (nv[0] * Gb[q] + nv[1] * Gb[4 + q] + nv[2] * Gb[8 + q]);
                ot[q * n2] = val;
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
            const double m0 = ge[(3 * d + 0) * n3], m1 = ge[(3 * d + 1) * n3], m2 = ge[(3 * d + 2) * n3];
#pragma unroll
            for (int q = 0; q < 4; q++) s_a[q * n3 + tid] = m0 * Gb[q] + m1 * Gb[4 + q] + m2 * Gb[8 + q];
        }
        line_sync<N>(d);
        if (active) {
            int idx = d == 0 ? i : (d == 1 ? j : k);
            int stride = d == 0 ? 1 : (d == 1 ? n : n2);
            int base = tid - idx * stride;
            for (int m = 0; m <= N; m++) {
                double dd = M.dt(idx, m);  // D(m, idx)
#pragma unroll
                for (int q = 0; q < 4; q++) qb[q] += dd * s_a[q * n3 + base + m * stride];
            }
        }
        line_sync<N>(d);
    }
    if (!active) return;
    pb.u += qb[0];
    pb.v += qb[1];
    pb.w += qb[2];
    pb.T += qb[3];
    // U read again: kept in registers since the load (the primitives), it spilled 32 more bytes per lane and
    // measured 3 % slower on MI355X
    double cU[5];
    const double* u = U + (long)e * 5 * n3 + tid;
#pragma unroll
    for (int v = 0; v < 5; v++) cU[v] = u[v * n3];
    double Ubv[5];
    prim_bar_to_cons(cU, pb, gd.kappa, gd.rR, Ubv);
    double* ub = Ubar + (long)e * 5 * n3 + tid;
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double r = Ubv[v] + Ubf[v];
        ub[v * n3] = accumulate ? ub[v * n3] + r : r;
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_adj_face: the neighbour's BR1 surface term of a VJP on its own (the last one of a sweep, or a single VJP)
// ---------------------------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(Dm<N>::block)
    k_adj_face(const double* __restrict__ U, Mesh mesh, const double* __restrict__ nghtr, GasD gd,
               double* __restrict__ Ubar)
{
    using D = Dm<N>;
    static_assert(D::node_per_thread, "one volume node per thread: N <= 9");
    constexpr int n2 = D::n2, n3 = D::n3;
    __shared__ double s_f[6 * 4 * n2];
    const int e = blockIdx.x, tid = threadIdx.x;
    lift_face_tasks<N>(e, tid, mesh, nghtr, s_f);
    __syncthreads();
    if (tid >= n3) return;
    double Ub[5];
    if (!lift_face_node_ubar<N>(e, tid, U, gd, s_f, Ub)) return;
    double* o = Ubar + (long)e * 5 * n3 + tid;
#pragma unroll
    for (int v = 0; v < 5; v++) o[v * n3] += Ub[v];
}

}  // namespace galx
