// galx_kernels_ho.hpp -- the stage kernels for degrees whose elements have more volume nodes than a workgroup has
// threads (N >= 10: (N+1)^3 > 1024).
//
// Same algorithm, operation order and HBM layout as k_lift, k_rhs and k_grad (galx_kernels.hpp), which give every
// volume node a thread of its own.  Here one 1024-thread workgroup per element loops over the element's nodes, and
// what those kernels keep in LDS or in registers across a barrier lives in the element's slice of a workspace in
// HBM (mesh.ws): the line-read arrays of one N = 13 element take 285 KB in FP64, more than the 160 KB of LDS of a
// CU.  Values a node needs again after a barrier (its state, primitive variables, metrics) are recomputed or
// re-read rather than stored.  Each node is handled by one thread in every phase, and every phase that reads
// another node's values follows a barrier.  These degrees serve convergence studies on small meshes
// (regressioncheck/checks/convtest/p_3D runs N = 2..13 on 2^3 elements); the production degrees use the
// one-node-per-thread kernels.
#pragma once
#include "galx_kernels.hpp"

namespace galx {

// per-element workspace, offsets in values of the compute type T
template <int N>
struct HoWs {
    static constexpr int n2 = Dm<N>::n2, n3 = Dm<N>::n3;
    static constexpr int lift = 0;                  // [4][n3]     (u,v,w,T) of every node
    static constexpr int lflux = lift + 4 * n3;     // [6][7][n2]  lifting fluxes and normals of the face nodes
    static constexpr int fgrad = lflux + 42 * n2;   // [6][12][n2] gradients of the face nodes (k_lift)
    static constexpr int nb = fgrad + 72 * n2;      // [6][5][n2]  neighbour trace records
    static constexpr int strs = nb + 30 * n2;       // [9][n3]     stress tensor and heat flux of every node
    static constexpr int v = strs + 9 * n3;         // [6][n3]     halves of (rho,u,v,w,p,H) (split form)
    static constexpr int m = v + 6 * n3;            // [3][n3]     metric of the current direction, halved (split)
    static constexpr int f = m + 3 * n3;            // [5][n3]     contravariant flux of the current direction
    static constexpr int ut = f + 5 * n3;           // [5][n3]     volume integral
    static constexpr int tmp = ut + 5 * n3;         // [5][n3]     split-form flux differencing
    static constexpr int res = tmp + 5 * n3;        // [6][5][n2]  face results
    static constexpr int unew = res + 30 * n2;      // [6][5][n2]  trace records of the new state
    static constexpr int size = unew + 30 * n2;
};

template <int N, typename T>
__device__ __forceinline__ T* ho_slice(const Mesh& mesh, int e)
{
    return static_cast<T*>(mesh.ws) + (long)e * HoWs<N>::size;
}

// the (u,v,w,T) planes of the lifting; split form: the halved (rho,u,v,w,p,H) planes of the volume flux
template <int N, typename T, bool SPLIT>
__device__ __forceinline__ void ho_load_planes(int e, int tid, const double* U, const GasD& gd, T* w)
{
    using W = HoWs<N>;
    constexpr int n3 = Dm<N>::n3;
    for (int node = tid; node < n3; node += Dm<N>::block) {
        T uc[5], pr[6], H;
        load_state<N, T>(e, node, U, gd, uc, pr, H);
        w[W::lift + node] = pr[1];
        w[W::lift + n3 + node] = pr[2];
        w[W::lift + 2 * n3 + node] = pr[3];
        w[W::lift + 3 * n3 + node] = pr[5];
        if (SPLIT) {
            const T hv[6] = {T(0.5) * pr[0], T(0.5) * pr[1], T(0.5) * pr[2], T(0.5) * pr[3], T(0.5) * pr[4],
                             T(0.5) * H};
#pragma unroll
            for (int q = 0; q < 6; q++) w[W::v + q * n3 + node] = hv[q];
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_lift_ho: k_lift for N >= 10
// ---------------------------------------------------------------------------------------------------------
template <int N, typename T, typename TT>
__global__ void __launch_bounds__(Dm<N>::block)
    k_lift_ho(const double* __restrict__ U, Mesh mesh, const TT* __restrict__ utr, TT* __restrict__ vtr,
              const T* __restrict__ mats, GasD gd, T LHm, T LHp)
{
    using D = Dm<N>;
    using W = HoWs<N>;
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3, B = D::block;
    __shared__ MatsD<N, T> M;
    const int e = blockIdx.x, tid = threadIdx.x;
    T* w = ho_slice<N, T>(mesh, e);
    T* s_lift = w + W::lift;
    T* s_lflux = w + W::lflux;
    T* s_g = w + W::fgrad;
    load_mats<N, T>(M, mats, tid, B);
    const Gas<T> gas = make_gas<T>(gd);
    ho_load_planes<N, T, false>(e, tid, U, gd, w);
    __syncthreads();
    lift_face_fluxes<N, T, TT>(e, tid, mesh, utr, s_lift, s_lflux, gas, nullptr);
    __syncthreads();
    // gradients of the surface nodes only, as k_lift
    for (int t = tid; t < Surf<N>::count; t += B) {
        const int gn = surface_node<N>(t);
        T gm[10], g[12];
        load_surf_geo<N, T>(e, t, mesh, gm);
        node_gradient<N, T>(gn, M, s_lift, s_lflux, LHm, LHp, gm, g);
        const int i = gn % n, j = (gn / n) % n, k = gn / n2;
#pragma unroll
        for (int ls = 1; ls <= 6; ls++) {
            int ab;
            if (!on_face<N>(ls, i, j, k, ab)) continue;
            T* o = s_g + (ls - 1) * 12 * n2 + ab;
#pragma unroll
            for (int q = 0; q < 12; q++) o[q * n2] = g[q];
        }
    }
    __syncthreads();
    for (int t = tid; t < D::nf; t += B) {
        int ls = t / n2 + 1, ab = t - (ls - 1) * n2, a = ab % n, b = ab / n;
        int node = face_vol_node<N>(ls, a, b);
        int c = mesh.conn[e * 6 + ls - 1];
        long side = c >> 3;
        int flip = c & 7;
        int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
        const T* fg = GeoOf<T>::fgeo(mesh) + side * 10 * n2 + pq;
        T nv[3] = {fg[0], fg[n2], fg[2 * n2]};
        T P[4] = {T(0), s_lift[node], s_lift[n3 + node], s_lift[2 * n3 + node]};
        T G[12];
        const T* sg = s_g + (ls - 1) * 12 * n2 + ab;
#pragma unroll
        for (int q = 0; q < 12; q++) G[q] = sg[q * n2];
        T f[5];
        normal_diff_flux(P, G, nv, gas, f);
        TT* o = vtr + ((side * 2 + (flip ? 1 : 0)) * 4) * n2 + pq;
#pragma unroll
        for (int q = 0; q < 4; q++) o[q * n2] = TT(f[1 + q]);
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_grad_ho: k_grad for N >= 10 (BR1 volume gradients of (u,v,w,T) for analysis, planes 0..11 of gvol
// [e][15][n3] in FP64)
// ---------------------------------------------------------------------------------------------------------
template <int N, typename T, typename TT>
__global__ void __launch_bounds__(Dm<N>::block)
    k_grad_ho(const double* __restrict__ U, Mesh mesh, const TT* __restrict__ utr, const T* __restrict__ mats,
              GasD gd, T LHm, T LHp, double* __restrict__ gvol)
{
    using D = Dm<N>;
    using W = HoWs<N>;
    constexpr int n3 = D::n3, B = D::block;
    __shared__ Mats<N, T> M;
    const int e = blockIdx.x, tid = threadIdx.x;
    T* w = ho_slice<N, T>(mesh, e);
    load_mats<N, T>(M, mats, tid, B);
    const Gas<T> gas = make_gas<T>(gd);
    ho_load_planes<N, T, false>(e, tid, U, gd, w);
    __syncthreads();
    lift_face_fluxes<N, T, TT>(e, tid, mesh, utr, w + W::lift, w + W::lflux, gas, nullptr);
    __syncthreads();
    for (int node = tid; node < n3; node += B) {
        T g[12];
        node_gradient<N, T, true>(e, node, mesh, M, w + W::lift, w + W::lflux, LHm, LHp, g);
        double* o = gvol + (long)e * 15 * n3 + node;
#pragma unroll
        for (int q = 0; q < 12; q++) o[q * n3] = double(g[q]);
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_rhs_ho: k_rhs for N >= 10.  The split form takes every pair of a line in the loop over all l, the diagonal
// included, as k_rhs does for the degrees whose lines do not tile a wave.
// ---------------------------------------------------------------------------------------------------------
template <int N, typename T, typename TT, bool SPLIT, int RIEMANN>
__global__ void __launch_bounds__(Dm<N>::block)
    k_rhs_ho(const double* U, double* Uout, Mesh mesh, const TT* __restrict__ utr, const TT* __restrict__ vtr,
             TT* __restrict__ utr_next, double* __restrict__ Ut, double* __restrict__ Kreg, const T* __restrict__ mats,
             GasD gd, T LHm, T LHp, Stage st)
{
    using D = Dm<N>;
    using W = HoWs<N>;
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3, B = D::block;
    __shared__ Mats<N, T> M;
    const int e = blockIdx.x, tid = threadIdx.x;
    T* w = ho_slice<N, T>(mesh, e);
    T *s_lift = w + W::lift, *s_lflux = w + W::lflux, *s_nb = w + W::nb, *s_str = w + W::strs;
    T *s_v = w + W::v, *s_m = w + W::m, *s_f = w + W::f, *s_ut = w + W::ut, *s_tmp = w + W::tmp;
    T *s_res = w + W::res, *s_u = w + W::unew;
    load_mats<N, T>(M, mats, tid, B);
    const Gas<T> gas = make_gas<T>(gd);
    const T* geT = GeoOf<T>::geo(mesh) + (long)e * 10 * n3;
    ho_load_planes<N, T, SPLIT>(e, tid, U, gd, w);
    for (int node = tid; node < n3; node += B)
#pragma unroll
        for (int v = 0; v < 5; v++) {
            s_ut[v * n3 + node] = T(0);
            s_tmp[v * n3 + node] = T(0);
        }
    __syncthreads();
    lift_face_fluxes<N, T, TT>(e, tid, mesh, utr, s_lift, s_lflux, gas, s_nb);
    __syncthreads();
    // BR1 gradients -> stress tensor and heat flux of every node
    for (int node = tid; node < n3; node += B) {
        T uc[5], pr[6], H, gm[10], g[12];
        load_state<N, T>(e, node, U, gd, uc, pr, H);
        load_node_geo<N, T>(e, node, mesh, gm);
        node_gradient<N, T>(node, M, s_lift, s_lflux, LHm, LHp, gm, g);
        Stress<T> s;
        stress(pr, g, gas, s);
        const T sv[9] = {s.txx, s.tyy, s.tzz, s.txy, s.txz, s.tyz, s.qx, s.qy, s.qz};
#pragma unroll
        for (int c = 0; c < 9; c++) s_str[c * n3 + node] = sv[c];
    }
    // ---- volume terms, per direction (VolInt_weakForm, or VolInt_weakForm_Visc + VolInt_splitForm_Kernel) ----
#pragma unroll 1
    for (int d = 0; d < 3; d++) {
        __syncthreads();
        for (int node = tid; node < n3; node += B) {
            T uc[5], pr[6], H;
            load_state<N, T>(e, node, U, gd, uc, pr, H);
            const T* sp = s_str + node;
            const Stress<T> s{sp[0], sp[n3], sp[2 * n3], sp[3 * n3], sp[4 * n3], sp[5 * n3], sp[6 * n3], sp[7 * n3],
                              sp[8 * n3]};
            const T mme[3] = {geT[(3 * d) * n3 + node], geT[(3 * d + 1) * n3 + node], geT[(3 * d + 2) * n3 + node]};
            T f[5];
            if (!SPLIT) {
                transformed_flux<T, true>(uc, pr, s, mme, gas.hshift, f);
#pragma unroll
                for (int v = 0; v < 5; v++) s_f[v * n3 + node] = f[v];
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) s_m[c * n3 + node] = T(0.5) * mme[c];
                transformed_flux<T, false>(uc, pr, s, mme, gas.hshift, f);
#pragma unroll
                for (int v = 1; v < 5; v++) s_f[(v - 1) * n3 + node] = f[v];
            }
        }
        __syncthreads();
        for (int node = tid; node < n3; node += B) {
            const int i = node % n, j = (node / n) % n, k = node / n2;
            const int idx = d == 0 ? i : (d == 1 ? j : k);
            const auto ln = Lay<N>::line(d, i, j, k);
            T ut[5];
#pragma unroll
            for (int v = 0; v < 5; v++) ut[v] = s_ut[v * n3 + node];
            if (!SPLIT) {
                for (int l = 0; l <= N; l++) {
                    const T dh = M.dht(l, idx);
                    const int o = ln(l);
#pragma unroll
                    for (int v = 0; v < 5; v++) ut[v] = ut[v] + dh * s_f[v * n3 + o];
                }
            } else {
                for (int l = 0; l <= N; l++) {
                    const T dh = M.dht(l, idx);
                    const int o = ln(l);
#pragma unroll
                    for (int v = 1; v < 5; v++) ut[v] = ut[v] + dh * s_f[(v - 1) * n3 + o];
                }
                T hv[6], hm[3], tmp[5];
#pragma unroll
                for (int q = 0; q < 6; q++) hv[q] = s_v[q * n3 + node];
#pragma unroll
                for (int c = 0; c < 3; c++) hm[c] = s_m[c * n3 + node];
#pragma unroll
                for (int v = 0; v < 5; v++) tmp[v] = s_tmp[v * n3 + node];
                for (int l = 0; l <= N; l++) {
                    const int o = ln(l);
                    T Mo[3] = {s_m[o], s_m[n3 + o], s_m[2 * n3 + o]};
                    T G[5];
                    split_volume_flux_pi_half<T>(hv[0], hv[1], hv[2], hv[3], hv[4], hv[5], s_v[o], s_v[n3 + o],
                                                 s_v[2 * n3 + o], s_v[3 * n3 + o], s_v[4 * n3 + o], s_v[5 * n3 + o],
                                                 hm, Mo, G);
                    const T dv = M.dvs(l, idx);
#pragma unroll
                    for (int v = 0; v < 5; v++) tmp[v] = tmp[v] + dv * G[v];
                }
#pragma unroll
                for (int v = 0; v < 5; v++) s_tmp[v * n3 + node] = tmp[v];
            }
#pragma unroll
            for (int v = 0; v < 5; v++) s_ut[v * n3 + node] = ut[v];
        }
    }
    // ---- face fluxes (FillFlux) and surface integral (SurfInt_Device), as k_rhs ----
    for (int t = tid; t < D::nf; t += B) {
        int ls = t / n2 + 1, ab = t - (ls - 1) * n2, a = ab % n, b = ab / n;
        int node = face_vol_node<N>(ls, a, b);
        int c = mesh.conn[e * 6 + ls - 1];
        long side = c >> 3;
        int flip = c & 7;
        int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
        T Uo[5], fvo[4], Un[5], fvn[4];
        {
            const double* uo = U + (long)e * 5 * n3 + node;
            double cv[5], tr[5];
#pragma unroll
            for (int v = 0; v < 5; v++) cv[v] = uo[v * n3];
            trace_record<sizeof(TT) == 4>(cv, gd.kappa, gd.pref, tr);
#pragma unroll
            for (int v = 0; v < 5; v++) Uo[v] = T(TT(tr[v]));
        }
        const TT* vo = vtr + ((side * 2 + (flip ? 1 : 0)) * 4) * n2 + pq;
#pragma unroll
        for (int q = 0; q < 4; q++) fvo[q] = T(vo[q * n2]);
        const long nslot = side * 2 + (flip ? 0 : 1);
        const T* un = s_nb + (ls - 1) * 5 * n2 + ab;
        const TT* vn = vtr + nslot * 4 * n2 + pq;
#pragma unroll
        for (int v = 0; v < 5; v++) Un[v] = un[v * n2];
#pragma unroll
        for (int q = 0; q < 4; q++) fvn[q] = T(vn[q * n2]);
        T F[5];
        face_flux<N, T, TT, SPLIT, RIEMANN>(Uo, fvo, Un, fvn, GeoOf<T>::fgeo(mesh) + side * 10 * n2 + pq, n2, flip,
                                            gas, F);
        T sg = flip ? T(-1) : T(1);
#pragma unroll
        for (int v = 0; v < 5; v++) s_res[(ls - 1) * 5 * n2 + v * n2 + ab] = sg * F[v];
    }
    __syncthreads();
    // ---- gather, -sJ, LSERK register update; trace records of the new state ----
    for (int node = tid; node < n3; node += B) {
        const int i = node % n, j = (node / n) % n, k = node / n2;
        T loc[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int ls = 1; ls <= 6; ls++) {
            int ab;
            if (!on_face<N>(ls, i, j, k, ab)) continue;
            T LH = (ls == 1 || ls == 2 || ls == 5) ? LHm : LHp;
#pragma unroll
            for (int v = 0; v < 5; v++) loc[v] = loc[v] + s_res[(ls - 1) * 5 * n2 + v * n2 + ab] * LH;
        }
        double* ut_out = (st.mode == 0 ? Ut : Kreg) + (long)e * 5 * n3 + node;
        double* uo = Uout + (long)e * 5 * n3 + node;
        const double* ui = U + (long)e * 5 * n3 + node;
        const double sJd = mesh.geo[(long)e * 10 * n3 + 9 * n3 + node];
        double unew[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            T utv = s_ut[v * n3 + node];
            if (SPLIT) utv = utv + T(2) * s_tmp[v * n3 + node];
            // the enthalpy offset of the split flux (mixed form), as in k_rhs
            if (SPLIT && sizeof(T) == 4 && v == 4) utv = utv + T(gd.href) * (T(2) * s_tmp[node]);
            T tot = utv + loc[v];
            double utd = double(tot * T(-1)) * sJd;
            if (st.mode == 0) {
                ut_out[v * n3] = utd;
            } else {
                double kk = ut_out[v * n3] * st.mRKA + utd;
                ut_out[v * n3] = kk;
                unew[v] = ui[v * n3] + kk * st.bdt;
                uo[v * n3] = unew[v];
            }
        }
        if (st.mode != 0) {
            double tr[5];
            trace_record<sizeof(TT) == 4>(unew, gd.kappa, gd.pref, tr);
#pragma unroll
            for (int ls = 1; ls <= 6; ls++) {
                int ab;
                if (!on_face<N>(ls, i, j, k, ab)) continue;
#pragma unroll
                for (int v = 0; v < 5; v++) s_u[(ls - 1) * 5 * n2 + v * n2 + ab] = T(TT(tr[v]));
            }
        }
    }
    if (st.mode == 0) return;
    __syncthreads();
    for (int t = tid; t < D::nf; t += B) {
        int ls = t / n2 + 1, ab = t - (ls - 1) * n2;
        int c = mesh.conn[e * 6 + ls - 1];
        long side = c >> 3;
        int flip = c & 7;
        int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
        TT* o = utr_next + ((side * 2 + (flip ? 1 : 0)) * 5) * n2 + pq;
#pragma unroll
        for (int v = 0; v < 5; v++) o[v * n2] = TT(s_u[(ls - 1) * 5 * n2 + v * n2 + ab]);
    }
}

}  // namespace galx
