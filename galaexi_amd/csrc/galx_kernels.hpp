// galx_kernels.hpp -- element-per-workgroup DGSEM kernels for gfx950.
//
// Data layout in HBM (one rank's element share on one GPU):
//   U, Ut, Ut_tmp : double [nElems][5][n^3]          (element-major, variable planes, i fastest)
//   geo           : double [nElems][10][n^3]         (Ja^1(3), Ja^2(3), Ja^3(3), sJ)
//   sgeo          : double [nElems][10][surface_count(N)] geo at the surface nodes only (k_lift)
//   fgeo          : double [nSides][10][n^2]         (NormVec(3), TangVec1(3), TangVec2(3), SurfElem), master (p,q)
//   utr[2]        : TT [nSides][2][5][n^2]           U traces, slot 0 master / 1 slave, master (p,q); double buffered
//   vtr           : TT [nSides][2][4][n^2]           viscous normal-flux traces F_v(U,grad U).n_master, components
//                                                    (rho u, rho v, rho w, E) of each side (the mass component is 0)
//   geo32, sgeo32, fgeo32 : float, as geo, sgeo, fgeo  (mixed form only: their FP32 roundings)
//   conn          : int [nElems][6]                  (SideID-1) << 3 | flip
//   inv           : int [5 flips][6 locSides][n^2]   S2V2_inv as p + n*q
// TT (trace type) is double in the FP64 form and float in the mixed form.
//
// Per RK stage: k_lift (BR1 gradients -> viscous normal-flux traces) then k_rhs (gradients recomputed, viscous
// weak-form volume term, split-form flux differencing or weak Euler term, Riemann + viscous face fluxes, surface
// integral, -sJ scaling and the LSERK register update; writes U, Ut_tmp and the next U traces).  The BR1 face flux
// 0.5*(F_v(master) + F_v(slave)).n (riemann.f90:638-700) needs only the normal viscous flux of each side, so the
// sides exchange those 4 values per face node instead of the 12 gradient components.
#pragma once
#include "galx_physics.hpp"

#ifndef GALX_UNROLL_L
#define GALX_UNROLL_L 2
#endif
#ifndef GALX_UNROLL_SPLIT
#define GALX_UNROLL_SPLIT 1
#endif
#ifndef GALX_MIXED_WAVES
#define GALX_MIXED_WAVES 6
#endif

namespace galx {

// one workgroup per element: n^3 threads rounded up to whole waves, at most 1024 (N <= 9 gives every volume node a
// thread of its own; the high-order kernels of N >= 10 loop their 1024 threads over the nodes)
template <int N>
struct Dm {
    static constexpr int n = N + 1, n2 = n * n, n3 = n2 * n, nf = 6 * n2;
    static constexpr int block = n3 <= 64 ? 64 : (n3 <= 1024 ? ((n3 + 63) / 64) * 64 : 1024);
    static constexpr bool node_per_thread = n3 <= block;
    static constexpr int ft = (nf + block - 1) / block;
};

// nodes on the element surface, n^3 - (n-2)^3 of them
__host__ __device__ constexpr int surface_count(int N) { return 2 * (N + 1) * (N + 1) + (N - 1) * 4 * N; }

// LDS layout of the per-node arrays read along element lines (lifting, split and viscous volume terms) and of
// the D-matrices: node (i,j,k) at i + n*(j + n*k), the order of the element planes, row stride np = n.  Measured
// at N=7 and not used: a padded row stride n+1 (k_rhs FP64 1.99 vs 1.84 ms, mixed 1.31 vs 1.26 ms) and an XOR
// swizzle of i by j (FP64 1.86 vs 1.82 ms, mixed 1.29 vs 1.26 ms), both meant to spread an i-line read over more
// LDS banks.
template <int N>
struct Lay {
    static constexpr int n = N + 1, np = n, n3 = n * n * n;
    __device__ static __forceinline__ int idx(int i, int j, int k) { return i + n * (j + n * k); }
    __device__ static __forceinline__ int at(int node) { return node; }
    // node l of the direction-d line through (i,j,k): C + s*l
    struct Line {
        int C, s;
        __device__ __forceinline__ int operator()(int l) const { return C + s * l; }
    };
    __device__ static __forceinline__ Line line(int d, int i, int j, int k)
    {
        if (d == 0) return Line{n * (j + n * k), 1};
        if (d == 1) return Line{i + n * n * k, n};
        return Line{i + n * j, n * n};
    }
};

// volume node of face node (a,b) of local side ls (GL trace); ls numbering of src/flexi.h:98-103
template <int N>
__device__ __forceinline__ int face_vol_node(int ls, int a, int b)
{
    constexpr int n = N + 1;
    switch (ls) {
    case 5: return n * (a + n * b);
    case 3: return N + n * (a + n * b);
    case 2: return a + n * n * b;
    case 4: return a + n * (N + n * b);
    case 1: return a + n * b;
    default: return a + n * (b + n * N);
    }
}

// is node (i,j,k) on local side ls, and its face coordinates (a,b) there
template <int N>
__device__ __forceinline__ bool on_face(int ls, int i, int j, int k, int& ab)
{
    constexpr int n = N + 1;
    switch (ls) {
    case 1: ab = i + n * j; return k == 0;
    case 2: ab = i + n * k; return j == 0;
    case 3: ab = j + n * k; return i == N;
    case 4: ab = i + n * k; return j == N;
    case 5: ab = j + n * k; return i == 0;
    default: ab = i + n * j; return k == N;
    }
}

// do the lines of direction d of an element lie inside one 64-lane wave (threads = nodes i + n (j + n k))?  i-lines
// if n divides 64, j-lines if n^2 does, k-lines if the element fits one wave.  The LDS exchanges along such lines
// need only the wave's own ordering, no workgroup barrier.
template <int N>
__device__ __forceinline__ bool lines_in_wave(int d)
{
    constexpr int n = N + 1;
    return d == 0 ? 64 % n == 0 : (d == 1 ? 64 % (n * n) == 0 : n * n * n <= 64);
}

// orders the wave's LDS accesses before and after this point (a wave's LDS instructions execute in order; this
// keeps the compiler from moving them across)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the barrier of an exchange along the lines of direction d
template <int N>
__device__ __forceinline__ void line_sync(int d)
{
    if (lines_in_wave<N>(d))
        wave_lds_sync();
    else
        __syncthreads();
}

// gas constants and the pressure offset in FP64, with the derived constants the kernels use (make_gasd fills
// them); each kernel derives its Gas<T>
struct GasD {
    double kappa, R, mu, lambda, pref;
    double rR;      // 1/R
    double rkm1;    // 1/(kappa-1)
    double eshift;  // pref/(kappa-1): the energy offset of a state that carries p' = p - pref
    double hshift;  // kappa*pref/(kappa-1): the enthalpy offset
    double href;    // the constant the split volume flux subtracts from the node enthalpy (0: none)
};

__host__ __device__ inline GasD make_gasd(double kappa, double R, double mu, double lambda, double pref,
                                          double href = 0.0)
{
    return GasD{kappa, R, mu, lambda, pref, 1.0 / R, 1.0 / (kappa - 1.0), pref / (kappa - 1.0),
                kappa * pref / (kappa - 1.0), href};
}

template <typename T>
__device__ __forceinline__ Gas<T> make_gas(const GasD& d)
{
    return Gas<T>{T(d.kappa), T(d.R), T(d.mu), T(d.lambda), T(d.pref), T(d.hshift), T(d.rR), T(d.rkm1)};
}

struct Mesh {
    const double* geo;
    const double* fgeo;
    const int* conn;
    const int* inv;
    const double* sgeo;  // [nElems][10][surface_count(N)]: geo at the surface nodes, in surface_node order
    // the FP32 roundings of geo, fgeo and sgeo (mixed form only, else null).  The mixed kernels use every metric,
    // face normal, tangent and SurfElem value as an FP32 operand; reading these copies gives them the same
    // operands, rounded once by the host-side set-up instead of on every load, at half the bytes.  The FP64
    // arrays stay the master copy: the Jacobian of the update, the time step and the adjoint read them.
    const float* geo32;
    const float* fgeo32;
    const float* sgeo32;
    void* ws;  // per-element workspace of the high-order kernels (N >= 10), else null
};

// the geometry arrays a kernel of compute type T reads its metric and face operands from
template <typename T>
struct GeoOf;
template <>
struct GeoOf<double> {
    __device__ static __forceinline__ const double* geo(const Mesh& m) { return m.geo; }
    __device__ static __forceinline__ const double* fgeo(const Mesh& m) { return m.fgeo; }
    __device__ static __forceinline__ const double* sgeo(const Mesh& m) { return m.sgeo; }
};
template <>
struct GeoOf<float> {
    __device__ static __forceinline__ const float* geo(const Mesh& m) { return m.geo32; }
    __device__ static __forceinline__ const float* fgeo(const Mesh& m) { return m.fgeo32; }
    __device__ static __forceinline__ const float* sgeo(const Mesh& m) { return m.sgeo32; }
};

struct Stage {
    double mRKA, bdt;
    int mode;  // 0: write Ut; 1: LSERK register update
};

// matrices staged in LDS: D_T (transpose of D), D_Hat_T, DVolSurf, all (0:N,0:N); dt(a,b), dht(a,b), dvs(a,b)
// return entry (a,b) of the column-major matrix.  The forward kernels read (l, idx), idx the lane's own node
// index along a line and l common to the wave: stored transposed (TR, idx fastest), the lanes of a wave read
// consecutive entries instead of entries n apart, and k_rhs/k_lift run 2.6 %/1.5 % faster in FP64.  The adjoint
// reads (idx, m) in its contractions and keeps the column-major storage (TR false).  Leading dimension
// Lay<N>::np.
template <int N, typename T, bool TR = true>
struct Mats {
    static constexpr int ld = Lay<N>::np;
    T DT[ld * (N + 1)];
    T DhT[ld * (N + 1)];
    T DVS[ld * (N + 1)];
    __device__ static __forceinline__ int off(int a, int b) { return TR ? b + ld * a : a + ld * b; }
    __device__ __forceinline__ T dt(int a, int b) const { return DT[off(a, b)]; }
    __device__ __forceinline__ T dht(int a, int b) const { return DhT[off(a, b)]; }
    __device__ __forceinline__ T dvs(int a, int b) const { return DVS[off(a, b)]; }
};

// D_T alone, for a kernel that only differentiates (k_lift): 1 KB less LDS at N=7 in FP64
template <int N, typename T, bool TR = true>
struct MatsD {
    static constexpr int ld = Lay<N>::np;
    T DT[ld * (N + 1)];
    __device__ static __forceinline__ int off(int a, int b) { return TR ? b + ld * a : a + ld * b; }
    __device__ __forceinline__ T dt(int a, int b) const { return DT[off(a, b)]; }
};

template <int N, typename T, bool TR>
__device__ __forceinline__ void load_mats(MatsD<N, T, TR>& M, const T* __restrict__ g, int tid, int nthreads)
{
    constexpr int n = N + 1, nn = n * n;
    for (int x = tid; x < nn; x += nthreads) M.DT[MatsD<N, T, TR>::off(x % n, x / n)] = g[x];
}

template <int N, typename T, bool TR>
__device__ __forceinline__ void load_mats(Mats<N, T, TR>& M, const T* __restrict__ g, int tid, int nthreads)
{
    constexpr int n = N + 1, nn = n * n;
    for (int x = tid; x < 3 * nn; x += nthreads) {
        T v = g[x];
        int r = x % nn, o = Mats<N, T, TR>::off(r % n, r / n);
        if (x < nn) M.DT[o] = v;
        else if (x < 2 * nn) M.DhT[o] = v;
        else M.DVS[o] = v;
    }
}

// ---------------------------------------------------------------------------------------------------------
// BR1 lifting (src/dg/lifting/lifting_br1.t90:47-167, strong form, non-conservative volume integral
// lifting_volint.t90:262-328, Lifting_FillFlux lifting_fillflux.t90:39-93, SurfInt_Device_Single_Flux with
// doApplyJacobian surfint.t90:591-725).  Lifted variables (u,v,w,T); the density gradient of the reference's
// nVarLifting=5 never enters the fluxes and is not formed.
//   s_lift  : T[4][Lay<N>::n3]   own (u,v,w,T) in the padded node layout, filled by the caller before the call
//             (synchronised)
//   s_lflux : T[6][7][n2] scratch
//   s_nb    : T[6][5][n2] (optional) receives the neighbour trace records of the six faces, for later face tasks
// returns g[12] = grad(u,v,w,T) in x, then y, then z for node tid (valid if tid < n3).
// ---------------------------------------------------------------------------------------------------------
// one face task: face node ab of local side ls of element e, with its side, flip and master (p,q) index (from the
// connectivity and the inverse S2V2 table: two dependent loads)
struct FaceTask {
    long side;
    int ls, ab, flip, pq;
};

template <int N>
__device__ __forceinline__ FaceTask face_task(int e, int t, const Mesh& mesh)
{
    constexpr int n2 = (N + 1) * (N + 1);
    FaceTask f;
    f.ls = t / n2 + 1;
    f.ab = t - (f.ls - 1) * n2;
    const int c = mesh.conn[e * 6 + f.ls - 1];
    f.side = c >> 3;
    f.flip = c & 7;
    f.pq = mesh.inv[(f.flip * 6 + f.ls - 1) * n2 + f.ab];
    return f;
}

// the memory operands of a lifting face task: the neighbour's trace record, the normal and SurfElem
template <typename T>
struct LiftOperands {
    T tr[5], nv[3], se;
};

template <int N, typename T, typename TT>
__device__ __forceinline__ void load_lift_operands(const FaceTask& f, const Mesh& mesh, const TT* __restrict__ utr,
                                                   LiftOperands<T>& p)
{
    constexpr int n2 = (N + 1) * (N + 1);
    const TT* nb = utr + ((f.side * 2 + (f.flip ? 0 : 1)) * 5) * n2 + f.pq;
#pragma unroll
    for (int v = 0; v < 5; v++) p.tr[v] = T(nb[v * n2]);
    const T* fg = GeoOf<T>::fgeo(mesh) + f.side * 10 * n2 + f.pq;
    p.nv[0] = fg[0];
    p.nv[1] = fg[n2];
    p.nv[2] = fg[2 * n2];
    p.se = fg[9 * n2];
}

// BR1 lifting flux 0.5*SurfElem*(-L_master + L_slave) and the normal of one face node (s_lflux [6][7][n2]); the
// neighbour's trace record also to s_nb [6][5][n2] when given
template <int N, typename T>
__device__ __forceinline__ void lift_face_task(const FaceTask& f, const LiftOperands<T>& p, const T* s_lift,
                                               T* s_lflux, const Gas<T>& gas, T* s_nb)
{
    constexpr int n = N + 1, n2 = n * n;
    const int node = face_vol_node<N>(f.ls, f.ab % n, f.ab / n);
    if (s_nb) {
#pragma unroll
        for (int v = 0; v < 5; v++) s_nb[(f.ls - 1) * 5 * n2 + v * n2 + f.ab] = p.tr[v];
    }
    T uc[5], pr[6];
    trace_state(p.tr, gas, uc, pr);
    T Lnb[4] = {pr[1], pr[2], pr[3], pr[5]};
    T* o = s_lflux + (f.ls - 1) * 7 * n2 + f.ab;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        T own = s_lift[q * Lay<N>::n3 + Lay<N>::at(node)];
        T Lm = f.flip ? Lnb[q] : own, Lsl = f.flip ? own : Lnb[q];
        o[q * n2] = T(0.5) * p.se * (T(-1) * Lm + Lsl);
    }
    o[4 * n2] = p.nv[0];
    o[5 * n2] = p.nv[1];
    o[6 * n2] = p.nv[2];
}

// the lifting face tasks of every face node of element e
template <int N, typename T, typename TT>
__device__ __forceinline__ void lift_face_fluxes(int e, int tid, const Mesh& mesh, const TT* __restrict__ utr,
                                                 const T* s_lift, T* s_lflux, const Gas<T>& gas, T* s_nb)
{
    for (int t = tid; t < Dm<N>::nf; t += Dm<N>::block) {
        const FaceTask f = face_task<N>(e, t, mesh);
        LiftOperands<T> p;
        load_lift_operands<N, T, TT>(f, mesh, utr, p);
        lift_face_task<N, T>(f, p, s_lift, s_lflux, gas, s_nb);
    }
}

// BR1 gradient g[12] = grad(u,v,w,T) (x, then y, then z) at volume node `node` of element e, from s_lift and the
// lifting fluxes of lift_face_fluxes (the node half of lift_gradient; any thread may evaluate any node)
template <int N, typename T>
__device__ __forceinline__ void load_node_geo(int e, int node, const Mesh& mesh, T* gm)
{
    constexpr int n3 = Dm<N>::n3;
    const T* ge = GeoOf<T>::geo(mesh) + (long)e * 10 * n3 + node;
#pragma unroll
    for (int q = 0; q < 10; q++) gm[q] = ge[q * n3];
}

// the same ten values of surface node t (node surface_node<N>(t)) from the surface-node copy: 10 * surface_count
// contiguous doubles per element instead of every 64-byte row of the element's volume planes
template <int N, typename T>
__device__ __forceinline__ void load_surf_geo(int e, int t, const Mesh& mesh, T* gm)
{
    constexpr int ns = surface_count(N);
    const T* ge = GeoOf<T>::sgeo(mesh) + (long)e * 10 * ns + t;
#pragma unroll
    for (int q = 0; q < 10; q++) gm[q] = ge[q * ns];
}

// the BR1 gradient of volume node `node` from its reference-direction derivatives gxi, get, gze of (u,v,w,T), the
// metrics gm (Ja^1(3), Ja^2(3), Ja^3(3), sJ) and the surface terms of the lifting fluxes (lifting_volint.t90:262-328,
// surfint.t90:591-725)
template <int N, typename T>
__device__ __forceinline__ void gradient_combine(int node, const T* s_lflux, T LHm, T LHp, const T* gm, const T* gxi,
                                                 const T* get, const T* gze, T* g)
{
    constexpr int n = N + 1, n2 = n * n;
    const int i = node % n, j = (node / n) % n, k = node / n2;
    const T mf[3] = {gm[0], gm[1], gm[2]}, mg[3] = {gm[3], gm[4], gm[5]}, mh[3] = {gm[6], gm[7], gm[8]};
    const T sJ = gm[9];
    T loc[12];
#pragma unroll
    for (int q = 0; q < 12; q++) loc[q] = T(0);
#pragma unroll
    for (int ls = 1; ls <= 6; ls++) {
        int ab;
        if (!on_face<N>(ls, i, j, k, ab)) continue;
        const T* o = s_lflux + (ls - 1) * 7 * n2 + ab;
        T LH = (ls == 1 || ls == 2 || ls == 5) ? LHm : LHp;
        T n0 = o[4 * n2], n1 = o[5 * n2], n2v = o[6 * n2];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            T lf = o[q * n2];
            loc[q] = loc[q] + T(1) * (lf * n0) * LH;
            loc[4 + q] = loc[4 + q] + T(1) * (lf * n1) * LH;
            loc[8 + q] = loc[8 + q] + T(1) * (lf * n2v) * LH;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        T vx = mf[0] * gxi[q] + mg[0] * get[q] + mh[0] * gze[q];
        T vy = mf[1] * gxi[q] + mg[1] * get[q] + mh[1] * gze[q];
        T vz = mf[2] * gxi[q] + mg[2] * get[q] + mh[2] * gze[q];
        g[q] = sJ * (vx + loc[q]);
        g[4 + q] = sJ * (vy + loc[4 + q]);
        g[8 + q] = sJ * (vz + loc[8 + q]);
    }
}

// the same with fewer live registers (k_lift): each direction's derivatives become the Cartesian sums as they are
// formed, and the surface terms are added to those sums face by face, with no separate surface accumulator.  The
// sums round in another order than gradient_combine's, within the FP64 parity bar.
template <int N, typename T>
__device__ __forceinline__ void gradient_combine_lean(int node, const T* s_lflux, T LHm, T LHp, const T* gm, T* gxi,
                                                      T* get, T* gze, T* g)
{
    constexpr int n = N + 1, n2 = n * n;
    const int i = node % n, j = (node / n) % n, k = node / n2;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const T a = gxi[q], b = get[q], c = gze[q];
        g[q] = gm[0] * a + gm[3] * b + gm[6] * c;
        g[4 + q] = gm[1] * a + gm[4] * b + gm[7] * c;
        g[8 + q] = gm[2] * a + gm[5] * b + gm[8] * c;
    }
#pragma unroll
    for (int ls = 1; ls <= 6; ls++) {
        int ab;
        if (!on_face<N>(ls, i, j, k, ab)) continue;
        const T* o = s_lflux + (ls - 1) * 7 * n2 + ab;
        T LH = (ls == 1 || ls == 2 || ls == 5) ? LHm : LHp;
        T n0 = o[4 * n2], n1 = o[5 * n2], n2v = o[6 * n2];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            T lf = o[q * n2];
            g[q] = g[q] + (lf * n0) * LH;
            g[4 + q] = g[4 + q] + (lf * n1) * LH;
            g[8 + q] = g[8 + q] + (lf * n2v) * LH;
        }
    }
    const T sJ = gm[9];
#pragma unroll
    for (int q = 0; q < 12; q++) g[q] = sJ * g[q];
}

// gm: the node's Ja^1(3), Ja^2(3), Ja^3(3), sJ (load_node_geo), loaded by the caller ahead of the barrier that
// precedes this call so that their latency overlaps the face tasks.  LEAN: the line loop one step at a time and
// gradient_combine_lean, so that k_lift fits 80 registers without spilling (three FP64 workgroups per CU)
template <int N, typename T, class MT, bool LEAN = false>
__device__ __forceinline__ void node_gradient(int node, const MT& M, const T* s_lift, const T* s_lflux,
                                              T LHm, T LHp, const T* gm, T* g)
{
    constexpr int n = N + 1, n2 = n * n;
    constexpr int n3p = Lay<N>::n3;
    const int i = node % n, j = (node / n) % n, k = node / n2;
    T gxi[4] = {0, 0, 0, 0}, get[4] = {0, 0, 0, 0}, gze[4] = {0, 0, 0, 0};
    auto line_step = [&](int l) {
        T di = M.dt(l, i), dj = M.dt(l, j), dk = M.dt(l, k);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            gxi[q] = gxi[q] + di * s_lift[q * n3p + Lay<N>::idx(l, j, k)];
            get[q] = get[q] + dj * s_lift[q * n3p + Lay<N>::idx(i, l, k)];
            gze[q] = gze[q] + dk * s_lift[q * n3p + Lay<N>::idx(i, j, l)];
        }
    };
    if (LEAN) {
#pragma unroll 1
        for (int l = 0; l <= N; l++) line_step(l);
    } else {
#pragma unroll GALX_UNROLL_L
        for (int l = 0; l <= N; l++) line_step(l);
    }
    if (LEAN)
        gradient_combine_lean<N, T>(node, s_lflux, LHm, LHp, gm, gxi, get, gze, g);
    else
        gradient_combine<N, T>(node, s_lflux, LHm, LHp, gm, gxi, get, gze, g);
}

// 16x16x4 matrix-core step of type T: D = A B + C with A 16x4, B 4x16; lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16], and register r of lane l holds D[row(l / 16, r)][l % 16] (gfx950, checked on the device: rows
// 4 (l / 16) + r in FP32, (l / 16) + 4 r in FP64)
template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static __forceinline__ v4 step(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int row(int kq, int r) { return kq + 4 * r; }
};
template <>
struct Mfma16<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static __forceinline__ v4 step(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int row(int kq, int r) { return 4 * kq + r; }
};

// BR1 gradient of the thread's own node at N = 7 (512 threads, node = thread), the xi- and eta-line derivatives on
// the matrix cores.  The 8 x 8 nodes of a k-plane are one wave, and so are their i- and j-lines: for each of the two
// directions one chain of four 16x16x4 steps forms D F for the 8 lines x 4 variables of the wave's plane, with
// A = blockdiag(D, D) (two column groups of 16 line-variables) and B the lines' values from LDS.  The results reach
// the node lanes through the wave's own scratch (s_scr, 4 x 64 values; no workgroup barrier, the wave's LDS
// operations are ordered).  The zeta-lines cross the waves and stay on the VALU, where D(k, l) is uniform in the
// wave.
template <typename T, class MT>
__device__ __forceinline__ void node_gradient_mfma(int tid, const MT& M, const T* s_lift, const T* s_lflux, T* s_scr,
                                                   T LHm, T LHp, const T* gm, T* g)
{
    constexpr int N = 7, n3p = Lay<N>::n3;
    const int lane = tid & 63, w = tid >> 6;  // the wave's plane k = w
    const int r16 = lane & 15, kq = lane >> 4;
    T a[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int c = 4 * s + kq;  // column of blockdiag(D, D); D(a, l) = dt(l, a)
        a[s] = ((r16 >> 3) == (c >> 3)) ? M.dt(c & 7, r16 & 7) : T(0);
    }
    T gd2[2][4];
#pragma unroll
    for (int dir = 0; dir < 2; dir++) {
        typename Mfma16<T>::v4 acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int pos = 4 * (s & 1) + kq;           // line position of B's row
            const int col = 16 * (s >> 1) + r16;        // line-variable: line t, variable q
            const int t = col & 7, q = col >> 3;
            const int node = dir == 0 ? pos + 8 * t + 64 * w : t + 8 * pos + 64 * w;
            acc = Mfma16<T>::step(a[s], s_lift[q * n3p + node], acc);
        }
        T* scr = s_scr + w * 256;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = Mfma16<T>::row(kq, r), ap = row & 7;  // output line position
            const int col = 16 * (row >> 3) + r16, t = col & 7, q = col >> 3;
            scr[q * 64 + (dir == 0 ? ap + 8 * t : t + 8 * ap)] = acc[r];
        }
        wave_lds_sync();
#pragma unroll
        for (int q = 0; q < 4; q++) gd2[dir][q] = scr[q * 64 + lane];
        wave_lds_sync();
    }
    const int i = tid & 7, j = (tid >> 3) & 7;
    T gze[4] = {0, 0, 0, 0};
#pragma unroll GALX_UNROLL_L
    for (int l = 0; l <= N; l++) {
        const T dk = M.dt(l, w);
#pragma unroll
        for (int q = 0; q < 4; q++) gze[q] = gze[q] + dk * s_lift[q * n3p + Lay<N>::idx(i, j, l)];
    }
    gradient_combine<N, T>(tid, s_lflux, LHm, LHp, gm, gd2[0], gd2[1], gze, g);
}

template <int N, typename T, bool TR>
__device__ __forceinline__ void node_gradient(int e, int node, const Mesh& mesh, const Mats<N, T, TR>& M,
                                              const T* s_lift, const T* s_lflux, T LHm, T LHp, T* g)
{
    T gm[10];
    load_node_geo<N, T>(e, node, mesh, gm);
    node_gradient<N, T>(node, M, s_lift, s_lflux, LHm, LHp, gm, g);
}

// the k=0 plane, the rings of the inner k planes, the k=N plane
template <int N>
struct Surf {
    static constexpr int n = N + 1, ring = 4 * N, count = surface_count(N);
};

// volume node (i + n*(j + n*k)) of surface node t, 0 <= t < surface_count(N); the host builds the surface-node
// metric array in this order
__host__ __device__ __forceinline__ int surface_node_of(int N, int t)
{
    const int n = N + 1, n2 = n * n, ring = 4 * N;
    if (t < n2) return t;
    t -= n2;
    if (t < (n - 2) * ring) {
        const int k = 1 + t / ring, q = t % ring;
        int i, j;
        if (q < n) {
            i = q;
            j = 0;
        } else if (q < 2 * n) {
            i = q - n;
            j = N;
        } else if (q < 3 * n - 2) {
            i = 0;
            j = 1 + q - 2 * n;
        } else {
            i = N;
            j = 1 + q - (3 * n - 2);
        }
        return i + n * (j + n * k);
    }
    return t - (n - 2) * ring + n2 * N;
}

template <int N>
__host__ __device__ __forceinline__ int surface_node(int t)
{
    return surface_node_of(N, t);
}

template <int N, typename T, typename TT, bool TR>
__device__ __forceinline__ void lift_gradient(int e, int tid, const Mesh& mesh, const TT* __restrict__ utr,
                                              const Mats<N, T, TR>& M, const T* s_lift, T* s_lflux,
                                              const Gas<T>& gas, T LHm, T LHp, T* g, T* s_nb = nullptr,
                                              const T* gm = nullptr)
{
    lift_face_fluxes<N, T, TT>(e, tid, mesh, utr, s_lift, s_lflux, gas, s_nb);
    __syncthreads();
    if (tid < Dm<N>::n3) {
        if (gm)
            node_gradient<N, T>(tid, M, s_lift, s_lflux, LHm, LHp, gm, g);
        else
            node_gradient<N, T, TR>(e, tid, mesh, M, s_lift, s_lflux, LHm, LHp, g);
    }
}

// volume state of node tid: ConsToPrim (eos.f90:212-239) in FP64, then shifted by the pressure offset and
// rounded to T; H = (E+p)/rho - href also from the FP64 values (the split volume flux's enthalpy, see k_rhs)
template <int N, typename T>
__device__ __forceinline__ void load_state(int e, int tid, const double* __restrict__ U, const GasD& gd, T* uc, T* pr,
                                           T& H)
{
    constexpr int n3 = Dm<N>::n3;
    const double* u = U + (long)e * 5 * n3 + tid;
    double c[5];
#pragma unroll
    for (int v = 0; v < 5; v++) c[v] = u[v * n3];
    // the mixed form rounds these to FP32: its 1/rho takes one Newton step fewer
    double sRho = sizeof(T) == 4 ? rcp64_fast(c[0]) : rcp64(c[0]);
    double vx = c[1] * sRho, vy = c[2] * sRho, vz = c[3] * sRho;
    double p = (gd.kappa - 1.0) * (c[4] - 0.5 * (c[1] * vx + c[2] * vy + c[3] * vz));
    uc[0] = T(c[0]);
    uc[1] = T(c[1]);
    uc[2] = T(c[2]);
    uc[3] = T(c[3]);
    uc[4] = T(c[4] - gd.eshift);
    pr[0] = T(c[0]);
    pr[1] = T(vx);
    pr[2] = T(vy);
    pr[3] = T(vz);
    pr[4] = T(p - gd.pref);
    pr[5] = T(p * sRho * gd.rR);
    H = T((c[4] + p) * sRho - gd.href);
}

template <typename TT>
__device__ __forceinline__ void store_trace(TT* o, int n2, const double* c, const GasD& gd)
{
    double tr[5];
    trace_record<sizeof(TT) == 4>(c, gd.kappa, gd.pref, tr);
#pragma unroll
    for (int v = 0; v < 5; v++) o[v * n2] = TT(tr[v]);
}

// ---------------------------------------------------------------------------------------------------------
// k_trace: U traces of all faces (ProlongToFaceCons, EvalElemFaceGL_Device prolongtoface.t90:268-344, in the
// element-wise form: each element writes its own slot of each of its 6 sides).
// ---------------------------------------------------------------------------------------------------------
template <int N, typename TT>
__global__ void __launch_bounds__(Dm<N>::block)
    k_trace(const double* __restrict__ U, Mesh mesh, TT* __restrict__ utr, GasD gd)
{
    using D = Dm<N>;
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3;
    const int e = blockIdx.x, tid = threadIdx.x;
    for (int t = tid; t < D::nf; t += D::block) {
        int ls = t / n2 + 1, ab = t - (ls - 1) * n2, a = ab % n, b = ab / n;
        int node = face_vol_node<N>(ls, a, b);
        int c = mesh.conn[e * 6 + ls - 1];
        long side = c >> 3;
        int flip = c & 7;
        int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
        TT* o = utr + ((side * 2 + (flip ? 1 : 0)) * 5) * n2 + pq;
        const double* u = U + (long)e * 5 * n3 + node;
        double cv[5];
#pragma unroll
        for (int v = 0; v < 5; v++) cv[v] = u[v * n3];
        store_trace(o, n2, cv, gd);
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_lift: BR1 gradients and, on the element's own face nodes, the normal viscous flux
// F_v(U, grad U).n_master (EvalNormalDiffFlux3D flux.f90:779-847 as ViscousFlux riemann.f90:638-700 applies it)
// written to the element's slot of the viscous-flux traces; one face task per face node forms it.  Every node
// thread forms its gradient (eight waves in parallel), and the face-node gradients reach the face tasks through
// LDS: measured faster than face tasks that each evaluate their node's gradient (fewer, longer waves).
// ---------------------------------------------------------------------------------------------------------
// minimum waves per SIMD for k_lift: the mixed form's LDS (27 KB at N=7) lets four 512-thread workgroups share a
// CU at 8 waves per SIMD (at most 64 registers, no spills): 2 % faster on MI355X at N=7.  FP64 at 6 waves: three
// 44.5 KB workgroups per CU at N=7, at most 80 registers, which the lean gradient fits without spilling (74 at
// N=7): FP64 k_lift 9 % faster than at the compiler's choice of 120 registers and two workgroups.  With the
// plain gradient (88 bytes of spills per lane) it measured 63 % slower.  0 = compiler's choice
#ifndef GALX_LIFT_MIXED_WAVES
#define GALX_LIFT_MIXED_WAVES 8
#endif
#ifndef GALX_LIFT_FP64_WAVES
#define GALX_LIFT_FP64_WAVES 6
#endif
template <int N, typename T>
struct LiftWaves {
    static constexpr int w = sizeof(T) == 4 ? GALX_LIFT_MIXED_WAVES : GALX_LIFT_FP64_WAVES;
    static constexpr int value = (w > 0 && Dm<N>::block >= 256) ? w : 1;
};

template <int N, typename T, typename TT>
__global__ void __launch_bounds__(Dm<N>::block, (LiftWaves<N, T>::value))
    k_lift(const double* __restrict__ U, Mesh mesh, const TT* __restrict__ utr, TT* __restrict__ vtr,
           const T* __restrict__ mats, GasD gd, T LHm, T LHp)
{
    using D = Dm<N>;
    static_assert(D::node_per_thread, "one volume node per thread: N <= 9");
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3;
    constexpr int n3p = Lay<N>::n3;
    // own (u,v,w,T) | lifting fluxes, then the own face nodes' stress and heat flux (9 values, not the 12 gradients)
    constexpr int w_f = 6 * 9 * n2 > 6 * 7 * n2 ? 6 * 9 * n2 : 6 * 7 * n2;
    __shared__ T s_buf[4 * n3p + w_f];
    __shared__ MatsD<N, T> M;
    const int e = blockIdx.x, tid = threadIdx.x;
    T* s_lift = s_buf;
    T* s_lflux = s_buf + 4 * n3p;
    load_mats<N, T>(M, mats, tid, D::block);
    const Gas<T> gas = make_gas<T>(gd);
    T uc[5], pr[6], H;
    // only the surface nodes' gradients are used (for F_v.n on the faces): the first Surf<N>::count threads
    // evaluate them, one surface node each, with the metrics from the surface-node copy of the geometry, and the
    // interior nodes' gradients and metrics are skipped (measured on MI355X at N=7 against every node's gradient:
    // 8 % faster in FP64, 12 % in the mixed form; with the metrics from the volume planes, whose 64-byte rows all
    // hold surface nodes, 1.8 % faster in FP64 and 3 % slower in the mixed form)
    const bool gnode = tid < Surf<N>::count;
    const int gn = gnode ? surface_node<N>(tid) : 0;
    T gm[10];
    if (gnode) load_surf_geo<N, T>(e, tid, mesh, gm);
    if (tid < n3) {
        load_state<N, T>(e, tid, U, gd, uc, pr, H);
        const int pt = Lay<N>::at(tid);
        s_lift[pt] = pr[1];
        s_lift[n3p + pt] = pr[2];
        s_lift[2 * n3p + pt] = pr[3];
        s_lift[3 * n3p + pt] = pr[5];
    }
    __syncthreads();
    T g[12];
    lift_face_fluxes<N, T, TT>(e, tid, mesh, utr, s_lift, s_lflux, gas, nullptr);
    __syncthreads();
    if (gnode) node_gradient<N, T, MatsD<N, T>, true>(gn, M, s_lift, s_lflux, LHm, LHp, gm, g);
    __syncthreads();
    // the stress tensor and heat flux of each surface node (the terms of EvalNormalDiffFlux3D, flux.f90:779-847,
    // before the normal) to its face slots: 9 values per face node instead of the 12 gradients (9 KB less LDS at
    // N=7 in FP64, and FP64 k_lift 4 % faster on MI355X)
    T* s_q = s_buf + 4 * n3p;  // [6][9][n2]: txx, tyy, tzz, txy, txz, tyz, qx, qy, qz
    if (gnode) {
        const int pn = Lay<N>::at(gn);
        const T P[4] = {T(0), s_lift[pn], s_lift[n3p + pn], s_lift[2 * n3p + pn]};
        Stress<T> st;
        stress(P, g, gas, st);
        const T q9[9] = {st.txx, st.tyy, st.tzz, st.txy, st.txz, st.tyz, st.qx, st.qy, st.qz};
        int i = gn % n, j = (gn / n) % n, k = gn / n2;
#pragma unroll
        for (int ls = 1; ls <= 6; ls++) {
            int ab;
            if (!on_face<N>(ls, i, j, k, ab)) continue;
            T* o = s_q + (ls - 1) * 9 * n2 + ab;
#pragma unroll
            for (int q = 0; q < 9; q++) o[q * n2] = q9[q];
        }
    }
    __syncthreads();
    // F_v.n_master of each own face node (the index and normal loaded again: cache hits, and no registers held
    // through the gradients)
    for (int t = tid; t < D::nf; t += D::block) {
        const FaceTask f = face_task<N>(e, t, mesh);
        const T* fg = GeoOf<T>::fgeo(mesh) + f.side * 10 * n2 + f.pq;
        const T nv[3] = {fg[0], fg[n2], fg[2 * n2]};
        const T* sq = s_q + (f.ls - 1) * 9 * n2 + f.ab;
        T t9[9];
#pragma unroll
        for (int q = 0; q < 9; q++) t9[q] = sq[q * n2];
        // EvalNormalDiffFlux3D: -n.tau and -n.(tau.u + lambda grad T)
        T fl[4];
        fl[0] = -nv[0] * t9[0] - nv[1] * t9[3] - nv[2] * t9[4];
        fl[1] = -nv[0] * t9[3] - nv[1] * t9[1] - nv[2] * t9[5];
        fl[2] = -nv[0] * t9[4] - nv[1] * t9[5] - nv[2] * t9[2];
        fl[3] = -nv[0] * t9[6] - nv[1] * t9[7] - nv[2] * t9[8];
        TT* o = vtr + ((f.side * 2 + (f.flip ? 1 : 0)) * 4) * n2 + f.pq;
#pragma unroll
        for (int q = 0; q < 4; q++) o[q * n2] = TT(fl[q]);
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_grad: the BR1 volume gradients of (u,v,w,T) of the current state, as the reference keeps them in
// gradUx/y/z (lifting_br1.t90:47-167 with PP_OPTLIFT), for analysis: planes 0..11 of gvol [e][15][n3] in FP64
// ---------------------------------------------------------------------------------------------------------
template <int N, typename T, typename TT>
__global__ void __launch_bounds__(Dm<N>::block)
    k_grad(const double* __restrict__ U, Mesh mesh, const TT* __restrict__ utr, const T* __restrict__ mats, GasD gd,
           T LHm, T LHp, double* __restrict__ gvol)
{
    using D = Dm<N>;
    static_assert(D::node_per_thread, "one volume node per thread: N <= 9");
    constexpr int n2 = D::n2, n3 = D::n3;
    constexpr int n3p = Lay<N>::n3;
    __shared__ T s_buf[4 * n3p + 6 * 7 * n2];
    __shared__ Mats<N, T> M;
    const int e = blockIdx.x, tid = threadIdx.x;
    load_mats<N, T>(M, mats, tid, D::block);
    const Gas<T> gas = make_gas<T>(gd);
    T uc[5], pr[6], H;
    if (tid < n3) {
        load_state<N, T>(e, tid, U, gd, uc, pr, H);
        const int pt = Lay<N>::at(tid);
        s_buf[pt] = pr[1];
        s_buf[n3p + pt] = pr[2];
        s_buf[2 * n3p + pt] = pr[3];
        s_buf[3 * n3p + pt] = pr[5];
    }
    __syncthreads();
    T g[12];
    lift_gradient<N, T, TT>(e, tid, mesh, utr, M, s_buf, s_buf + 4 * n3p, gas, LHm, LHp, g);
    if (tid < n3) {
        double* o = gvol + (long)e * 15 * n3 + tid;
#pragma unroll
        for (int q = 0; q < 12; q++) o[q * n3] = double(g[q]);
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_grad_rho: the BR1 gradient of the density, the lifting variable LIFT_DENS that the reference lifts by default
// (PP_OPTLIFT = 0, src/equations/navierstokes/idealgas/eos.h:133-142) and that no flux here needs.  The same
// strong-form lifting as k_grad (lifting_br1.t90:47-167: sJ [sum_d Ja^d d/dxi_d rho + L_hat 0.5 (rho_slave -
// rho_master) SurfElem n]), in FP64 from the FP64 state and the neighbours' trace records; analysis only, so it
// reads its operands from memory directly.  Writes planes 12..14 (x, y, z) of gvol [e][15][n3].
//   DT: D_T (0:N,0:N) column-major in FP64
// ---------------------------------------------------------------------------------------------------------
template <int N, typename TT>
__global__ void __launch_bounds__(Dm<N>::block)
    k_grad_rho(const double* __restrict__ U, Mesh mesh, const TT* __restrict__ utr, const double* __restrict__ DT,
               double LHm, double LHp, double* __restrict__ gvol)
{
    using D = Dm<N>;
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3;
    const int e = blockIdx.x;
    const double* rho = U + (long)e * 5 * n3;  // density plane of element e
    const double* ge = mesh.geo + (long)e * 10 * n3;
    for (int node = threadIdx.x; node < n3; node += D::block) {
        const int i = node % n, j = (node / n) % n, k = node / n2;
        double dxi = 0.0, det = 0.0, dze = 0.0;
        for (int l = 0; l <= N; l++) {
            dxi += DT[l + n * i] * rho[Lay<N>::idx(l, j, k)];
            det += DT[l + n * j] * rho[Lay<N>::idx(i, l, k)];
            dze += DT[l + n * k] * rho[Lay<N>::idx(i, j, l)];
        }
        double loc[3] = {0.0, 0.0, 0.0};
        for (int ls = 1; ls <= 6; ls++) {
            int ab;
            if (!on_face<N>(ls, i, j, k, ab)) continue;
            const int c = mesh.conn[e * 6 + ls - 1];
            const long side = c >> 3;
            const int flip = c & 7;
            const int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
            const double nb = double(utr[((side * 2 + (flip ? 0 : 1)) * 5) * n2 + pq]);  // neighbour's rho
            const double own = rho[node];
            const double Lm = flip ? nb : own, Lsl = flip ? own : nb;
            const double* fg = mesh.fgeo + side * 10 * n2 + pq;
            const double lf = 0.5 * fg[9 * n2] * (-1.0 * Lm + Lsl);
            const double LH = (ls == 1 || ls == 2 || ls == 5) ? LHm : LHp;
            for (int d = 0; d < 3; d++) loc[d] = loc[d] + (lf * fg[d * n2]) * LH;
        }
        const double sJ = ge[9 * n3 + node];
        double* o = gvol + (long)e * 15 * n3 + node;
        for (int d = 0; d < 3; d++) {
            const double v = ge[d * n3 + node] * dxi + ge[(3 + d) * n3 + node] * det + ge[(6 + d) * n3 + node] * dze;
            o[(12 + d) * n3] = sJ * (v + loc[d]);
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// k_rhs: DGTimeDerivative_weakForm (src/dg/dg.f90:255-425) after the lifting, fused with the LSERK register
// update VAXPB_OUT_VAXPB_IN (src/globals/vector.f90:163-183, src/timedisc/timestep.f90:85-119).
//   SPLIT: VolInt_weakForm_Visc + VolInt_splitForm_Kernel (src/dg/volint.f90:60-119,265-353)
//   else : VolInt_weakForm with EvalTransformedEulerDiffFlux3D (src/dg/volint.f90:129-188)
//   FillFlux: Riemann + ViscousFlux + SurfElem (src/dg/fillflux.f90:45-172), SurfInt_Device (surfint.t90:352-586)
//   Ut = -sJ*Ut (src/dg/dg.f90:413,423)
// LDS: one work region (lifting, volume fluxes, split-form partner data, face results, new traces).  The face
// tasks take the normal viscous flux of both sides from the viscous-flux traces (the own slot written by this
// element's k_lift), so the gradient registers die after the stress tensor is formed.
// U is read and the new state written to Uout, which may be U itself (in-place LSERK update) or another
// buffer (the adjoint's recompute keeps every stage input).
// ---------------------------------------------------------------------------------------------------------
template <int N, typename T, typename TT, bool SPLIT, int RIEMANN>
__device__ __forceinline__ void face_flux(const T* tro, const T* fvo, const T* trn, const T* fvn,
                                          const T* __restrict__ fg, int n2, int flip, const Gas<T>& gas, T* F)
{
    T Uo[5], Un[5], Po[6], Pn[6];
    trace_state(tro, gas, Uo, Po);
    trace_state(trn, gas, Un, Pn);
    T nv[3] = {fg[0], fg[n2], fg[2 * n2]};
    if (SPLIT && RIEMANN == RIEMANN_LF) {
        // own side first; the master is the own side unless flipped
        lf_pi_cartesian<T>(Uo, Un, Po, Pn, nv, flip ? T(-1) : T(1), gas, F);
    } else {
        const T *UL = flip ? Un : Uo, *UR = flip ? Uo : Un, *PL = flip ? Pn : Po, *PR = flip ? Po : Pn;
        T t1[3] = {fg[3 * n2], fg[4 * n2], fg[5 * n2]};
        T t2[3] = {fg[6 * n2], fg[7 * n2], fg[8 * n2]};
        riemann_point<T, SPLIT>(RIEMANN, UL, UR, PL, PR, nv, t1, t2, gas, F);
    }
    // ViscousFlux (riemann.f90:638-700): 0.5*(F_v(master) + F_v(slave)).n, summed in an order both sides share
    T se = fg[9 * n2];
    F[0] = F[0] * se;
#pragma unroll
    for (int v = 1; v < 5; v++) F[v] = (F[v] + (T(0.5) * fvo[v - 1] + T(0.5) * fvn[v - 1])) * se;
}

// minimum waves per SIMD for k_rhs: 4 lets two 512-thread workgroups share a CU at N=7 in FP64 (registers capped
// at 128), 6 lets three share it in the mixed form (capped at 80, which its LDS share allows); the heavier
// entropy-fixed Roe solver and the small blocks of N<5 keep the compiler's choice.
template <int N, int RIEMANN, typename T>
struct RhsWaves {
    static constexpr int value =
        (RIEMANN == RIEMANN_ROEEF || Dm<N>::block < 256) ? 1 : (sizeof(T) == 4 ? GALX_MIXED_WAVES : 4);
};

template <int N, typename T, typename TT, bool SPLIT, int RIEMANN>
__global__ void __launch_bounds__(Dm<N>::block, (RhsWaves<N, RIEMANN, T>::value))
    k_rhs(const double* U, double* Uout, Mesh mesh, const TT* __restrict__ utr, const TT* __restrict__ vtr,
          TT* __restrict__ utr_next, double* __restrict__ Ut, double* __restrict__ Kreg, const T* __restrict__ mats,
          GasD gd, T LHm, T LHp, Stage st)
{
    using D = Dm<N>;
    static_assert(D::node_per_thread, "one volume node per thread: N <= 9");
    constexpr int n = D::n, n2 = D::n2, n3 = D::n3;
    constexpr int np = Lay<N>::np, n3p = Lay<N>::n3;  // padded node layout of the line-read arrays
    // FP64 at N = 7: the lifting's in-plane line derivatives on the matrix cores (node_gradient_mfma), with a 4 x 64
    // scratch per wave after the lifting arrays.  The mixed form keeps them on the VALU: with the scratch (6.6 KB more
    // LDS) it measured 12 % slower on MI355X.
    constexpr bool mfma_lift = N == 7 && sizeof(T) == 8;
    constexpr int w_a0 = 4 * n3p + 6 * 7 * n2;
    constexpr int w_a = w_a0 + (mfma_lift ? 8 * 4 * n2 : 0);  // lifting
    // mixed split form: the nine metric terms of every node stay in LDS (FP32, halves) in s_geo for the whole
    // kernel, so the volume terms read them there instead of from memory a second time (7.8 % faster at N=7)
    constexpr bool geo_lds = SPLIT && sizeof(T) == 4;
    constexpr int w_c = SPLIT ? (geo_lds ? 10 : 13) * n3p : 5 * n3p;  // volume fluxes / split-form partner data
    constexpr int w_d = 6 * 5 * n2;  // face results, new traces
    // split form: the i- and j-lines lie inside one 64-lane wave (n and n^2 divide 64: N = 1, 3, 7), so the pairs of
    // those lines are shared between lanes instead of formed twice
    constexpr bool pair_lines = SPLIT && 64 % n == 0 && 64 % n2 == 0;
    constexpr int w_ac = w_a > w_c ? w_a : w_c;
    constexpr int lds_w = w_ac > w_d ? w_ac : w_d;
    __shared__ T s_w[lds_w];
    __shared__ T s_geo[geo_lds ? 9 * n3p : 1];  // [3 directions][3][n3p] metrics Ja^1, Ja^2, Ja^3
    __shared__ T s_nb[6 * 5 * n2];  // neighbour trace records, from the lifting to the face tasks
    __shared__ Mats<N, T> M;
    __shared__ T s_dp[pair_lines ? n2 : 1];  // DVolSurf by cyclic shift for the shared pairs
    const int e = blockIdx.x, tid = threadIdx.x;
    const bool active = tid < n3;
    const int i = tid % n, j = (tid / n) % n, k = tid / n2;
    const int pt = Lay<N>::at(tid);
    load_mats<N, T>(M, mats, tid, D::block);
    if (pair_lines && tid < n2) {
        const int sh = tid / n, a = tid % n;
        s_dp[tid] = mats[2 * n2 + ((a + sh) & (n - 1)) + n * a];
    }
    const Gas<T> gas = make_gas<T>(gd);
    T uc[5], pr[6], Hme = T(0);
    // the node's metrics, issued with the state so that the lifting does not wait for them (measured on MI355X at
    // N=7: 8 % faster in the mixed form; 0.8 % slower in FP64 as a prefetch alone, so the FP64 weak form reads them
    // where it needs them).  The FP64 split form keeps them in registers for the volume terms and the update
    // (`geo_reg`) instead of reading them a second time: 51 B/DOF fewer HBM reads and 7 % faster at N=7.
    constexpr bool pgeo = sizeof(T) == 4 || SPLIT;
    constexpr bool geo_reg = SPLIT && sizeof(T) == 8;
    T gm[10];
    if (active) {
        if (pgeo) load_node_geo<N, T>(e, tid, mesh, gm);
        load_state<N, T>(e, tid, U, gd, uc, pr, Hme);
        if (geo_lds) {
#pragma unroll
            for (int c = 0; c < 9; c++) s_geo[c * n3p + pt] = T(0.5) * gm[c];  // halves: see the split loop
        }
        s_w[pt] = pr[1];
        s_w[n3p + pt] = pr[2];
        s_w[2 * n3p + pt] = pr[3];
        s_w[3 * n3p + pt] = pr[5];
    }
    __syncthreads();
    Stress<T> s;
    {
        T g[12];
        if (mfma_lift) {
            lift_face_fluxes<N, T, TT>(e, tid, mesh, utr, s_w, s_w + 4 * n3p, gas, s_nb);
            if (!pgeo) load_node_geo<N, T>(e, tid, mesh, gm);
            __syncthreads();
            node_gradient_mfma<T>(tid, M, s_w, s_w + 4 * n3p, s_w + w_a0, LHm, LHp, gm, g);
        } else {
            lift_gradient<N, T, TT>(e, tid, mesh, utr, M, s_w, s_w + 4 * n3p, gas, LHm, LHp, g, s_nb,
                                    pgeo ? gm : nullptr);
        }
        if (active) stress(pr, g, gas, s);
    }
    __syncthreads();

    T ut[5] = {0, 0, 0, 0, 0};
    const double* ge = mesh.geo + (long)e * 10 * n3 + tid;  // FP64 sJ of the update
    const T* geT = GeoOf<T>::geo(mesh) + (long)e * 10 * n3 + tid;
    if (!SPLIT) {
        // ---- full transformed fluxes, weak-form D_Hat_T contraction (VolInt_weakForm) ----
        T* s_f = s_w;  // [5][n3p]
#pragma unroll 1
        for (int d = 0; d < 3; d++) {
            if (active) {
                T m[3] = {geT[(3 * d) * n3], geT[(3 * d + 1) * n3], geT[(3 * d + 2) * n3]};
                T f[5];
                transformed_flux<T, true>(uc, pr, s, m, gas.hshift, f);
#pragma unroll
                for (int v = 0; v < 5; v++) s_f[v * n3p + pt] = f[v];
            }
            line_sync<N>(d);
            if (active) {
                int idx = d == 0 ? i : (d == 1 ? j : k);
                const auto ln = Lay<N>::line(d, i, j, k);
#pragma unroll GALX_UNROLL_L
                for (int l = 0; l <= N; l++) {
                    T dh = M.dht(l, idx);
                    const int o = ln(l);
#pragma unroll
                    for (int v = 0; v < 5; v++) ut[v] = ut[v] + dh * s_f[v * n3p + o];
                }
            }
            line_sync<N>(d);
        }
    } else {
        // ---- per direction: viscous contravariant flux with the weak-form D_Hat_T contraction
        // (VolInt_weakForm_Visc) and the split-form flux differencing (VolInt_splitForm_Kernel); the direction's
        // metric is read once for both ----
        // the split flux runs on halves (split_volume_flux_pi_half): s_v and s_m hold half of each node's
        // (rho,u,v,w,p,H) and metric, and the sum is doubled once at the end (exact)
        T* s_v = s_w;                                // [6][n3p] rho,u,v,w,p,H (halves)
        T* s_m = s_w + 6 * n3p;                      // [3][n3p] metric of the current direction (halves)
        T* s_f = s_w + (geo_lds ? 6 : 9) * n3p;      // [4][n3p] viscous contravariant flux (momentum, energy) of
                                                     // the current direction
        const T hv[6] = {T(0.5) * pr[0], T(0.5) * pr[1], T(0.5) * pr[2], T(0.5) * pr[3], T(0.5) * pr[4],
                         T(0.5) * Hme};
        if (active) {
#pragma unroll
            for (int q = 0; q < 6; q++) s_v[q * n3p + pt] = hv[q];
        }
        T tmp[5] = {0, 0, 0, 0, 0};
#pragma unroll 1
        for (int d = 0; d < 3; d++) {
            T mme[3] = {0, 0, 0}, hm[3] = {0, 0, 0};
            if (geo_lds) s_m = s_geo + 3 * d * n3p;
            if (active) {
                if (geo_lds) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        hm[c] = s_m[c * n3p + pt];
                        mme[c] = T(2) * hm[c];
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        // the prefetched values, kept in registers since the load (geo_reg)
                        mme[c] = d == 0 ? gm[c] : (d == 1 ? gm[3 + c] : gm[6 + c]);
                        hm[c] = T(0.5) * mme[c];
                        s_m[c * n3p + pt] = hm[c];
                    }
                }
                T f[5];
                transformed_flux<T, false>(uc, pr, s, mme, gas.hshift, f);
#pragma unroll
                for (int v = 1; v < 5; v++) s_f[(v - 1) * n3p + pt] = f[v];
            }
            // s_v (written after the lifting barrier) is read along the lines of d too: the k-lines' barrier below
            // comes before any read across waves
            line_sync<N>(d);
            if (active) {
                int idx = d == 0 ? i : (d == 1 ? j : k);
                const auto ln = Lay<N>::line(d, i, j, k);
#pragma unroll GALX_UNROLL_L
                for (int l = 0; l <= N; l++) {
                    T dh = M.dht(l, idx);
                    const int o = ln(l);
#pragma unroll
                    for (int v = 1; v < 5; v++) ut[v] = ut[v] + dh * s_f[(v - 1) * n3p + o];
                }
                if (pair_lines && d < 2) {
                    // the two-point flux is symmetric, G(a,b) = G(b,a) bit for bit (sums and products of the
                    // same operands), so each pair of a line is formed once: the lane at line position a forms
                    // the pairs (a, a+s) for s = 1..(n-1)/2 (mod n) and reads the pair (a-s, a) from the lane s
                    // positions back with a cross-lane permute; for even n both ends of a pair (a, a+n/2) form
                    // it.  The diagonal, DVolSurf(a,a) = 0 up to rounding for Gauss-Lobatto (dg.f90:229-233),
                    // is left out.  About half the flux evaluations of the loop over all l below.
                    constexpr int msk = n - 1;                  // n is a power of two here
                    const int st = d == 0 ? 1 : n;              // lane distance between neighbours on the line
                    const int lane0 = (tid & 63) - idx * st;    // lane of line position 0
                    const T* dp = s_dp + idx;                   // s_dp[s*n + a] = DVolSurf((a+s) mod n, a)
#pragma unroll
                    for (int s = 1; s <= n / 2; s++) {
                        const int b = (idx + s) & msk;          // partner formed here
                        const int o = ln(b);
                        T Mo[3] = {s_m[o], s_m[n3p + o], s_m[2 * n3p + o]};
                        T G[5];
                        split_volume_flux_pi_half<T>(hv[0], hv[1], hv[2], hv[3], hv[4], hv[5], s_v[o],
                                                     s_v[n3p + o], s_v[2 * n3p + o], s_v[3 * n3p + o],
                                                     s_v[4 * n3p + o], s_v[5 * n3p + o], hm, Mo, G);
                        const T db = dp[s * n];
                        if (2 * s < n) {
                            const int src = lane0 + ((idx - s) & msk) * st;  // the lane that formed (a-s, a)
                            const T da = dp[(n - s) * n];
#pragma unroll
                            for (int v = 0; v < 5; v++) {
                                const T r = __shfl(G[v], src);
                                tmp[v] = tmp[v] + db * G[v];
                                tmp[v] = tmp[v] + da * r;
                            }
                        } else {
#pragma unroll
                            for (int v = 0; v < 5; v++) tmp[v] = tmp[v] + db * G[v];
                        }
                    }
                } else {
                // every partner l other than the node itself, l = idx + s (mod n) for s = 1..N: no per-lane
                // branch.  VolInt_splitForm_Kernel (volint.f90:307-317) also adds the diagonal pair, whose
                // DVolSurf(i,i) vanishes for Gauss-Lobatto up to rounding (dg.f90:229-233); the shared pairs above
                // leave it out too (one flux of eight fewer on the k-lines at N=7: FP64 k_rhs 1.2 % faster)
#pragma unroll GALX_UNROLL_SPLIT
                for (int s = 1; s <= N; s++) {
                    const int l = idx + s <= N ? idx + s : idx + s - n;
                    const int o = ln(l);
                    T Mo[3] = {s_m[o], s_m[n3p + o], s_m[2 * n3p + o]};
                    T G[5];
                    split_volume_flux_pi_half<T>(hv[0], hv[1], hv[2], hv[3], hv[4], hv[5], s_v[o], s_v[n3p + o],
                                                 s_v[2 * n3p + o], s_v[3 * n3p + o], s_v[4 * n3p + o],
                                                 s_v[5 * n3p + o], hm, Mo, G);
                    T dv = M.dvs(l, idx);
#pragma unroll
                    for (int v = 0; v < 5; v++) tmp[v] = tmp[v] + dv * G[v];
                }
                }
            }
            line_sync<N>(d);
        }
#pragma unroll
        for (int v = 0; v < 5; v++) ut[v] = ut[v] + T(2) * tmp[v];
        // enthalpy offset (mixed form): the pair fluxes carried H - href, and the energy flux q {H} is
        // q {H - href} + href q, so the energy sum takes href times the mass sum back once per node.  The FP32
        // products then hold the enthalpy fluctuation at full precision instead of rounding q {H}, whose mean part
        // (about 250 for a Mach 0.1 flow) dominates it, at every pair.
        if (sizeof(T) == 4) ut[4] = ut[4] + T(gd.href) * (T(2) * tmp[0]);
    }

    // ---- face fluxes (FillFlux) and surface integral (SurfInt_Device) ----
    T* s_res = s_w;  // [6][5][n2]
    for (int t = tid; t < D::nf; t += D::block) {
        int ls = t / n2 + 1, ab = t - (ls - 1) * n2, a = ab % n, b = ab / n;
        int node = face_vol_node<N>(ls, a, b);
        int c = mesh.conn[e * 6 + ls - 1];
        long side = c >> 3;
        int flip = c & 7;
        int pq = mesh.inv[(flip * 6 + ls - 1) * n2 + ab];
        T Uo[5], fvo[4], Un[5], fvn[4];
        {
            // the record the neighbour reads.  Mixed form: from the own slot of the FP32 trace buffer (0.5 % faster
            // at N=7 on MI355X, though it reads 15 B/DOF more from HBM than the gather of the FP64 state, which
            // the load phase had brought into the caches).  FP64: from U at the face node (the own slot measured
            // 1.5 % slower at N=7).  Formed by the node
            // threads from the state they hold and passed through LDS instead, it measured 6 % slower in FP64 at
            // N=7: the state then stays live through the volume terms and spills.
            if (sizeof(TT) == 4) {
                const TT* ot = utr + ((side * 2 + (flip ? 1 : 0)) * 5) * n2 + pq;
#pragma unroll
                for (int v = 0; v < 5; v++) Uo[v] = T(ot[v * n2]);
            } else {
                const double* uo = U + (long)e * 5 * n3 + node;
                double c[5], tr[5];
#pragma unroll
                for (int v = 0; v < 5; v++) c[v] = uo[v * n3];
                trace_record<false>(c, gd.kappa, gd.pref, tr);
#pragma unroll
                for (int v = 0; v < 5; v++) Uo[v] = T(TT(tr[v]));
            }
        }
        const TT* vo = vtr + ((side * 2 + (flip ? 1 : 0)) * 4) * n2 + pq;  // written by k_lift of this element
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
    double unew[5];
    if (active) {
        T loc[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int ls = 1; ls <= 6; ls++) {
            int ab;
            if (!on_face<N>(ls, i, j, k, ab)) continue;
            T LH = (ls == 1 || ls == 2 || ls == 5) ? LHm : LHp;
#pragma unroll
            for (int v = 0; v < 5; v++) loc[v] = loc[v] + s_res[(ls - 1) * 5 * n2 + v * n2 + ab] * LH;
        }
        double* ut_out = (st.mode == 0 ? Ut : Kreg) + (long)e * 5 * n3 + tid;
        double* un = Uout + (long)e * 5 * n3 + tid;
        double kv[5] = {0, 0, 0, 0, 0}, uv[5] = {0, 0, 0, 0, 0};
        const double sJd = geo_reg ? double(gm[9]) : ge[9 * n3];
        if (st.mode != 0) {
            const double* uo = U + (long)e * 5 * n3 + tid;
#pragma unroll
            for (int v = 0; v < 5; v++) {
                kv[v] = ut_out[v * n3];
                uv[v] = uo[v * n3];
            }
        }
#pragma unroll
        for (int v = 0; v < 5; v++) {
            T tot = ut[v] + loc[v];
            double utd = double(tot * T(-1)) * sJd;
            if (st.mode == 0) {
                ut_out[v * n3] = utd;
            } else {
                double kk = kv[v] * st.mRKA + utd;
                ut_out[v * n3] = kk;
                unew[v] = uv[v] + kk * st.bdt;
                un[v * n3] = unew[v];
            }
        }
    }
    if (st.mode == 0) return;
    // ---- next U traces ----
    __syncthreads();
    T* s_u = s_w;  // [6][5][n2] trace records of the new U, rounded to the trace type
    if (active) {
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
    __syncthreads();
    for (int t = tid; t < D::nf; t += D::block) {
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

// ---------------------------------------------------------------------------------------------------------
// time step: CalcMaxEigenvalue + CALCTIMESTEP (src/equations/navierstokes/calctimestep.f90:51-296), FP64
// ---------------------------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(Dm<N>::block)
    k_dt_elem(const double* __restrict__ U, const double* __restrict__ geo, double kappa, double mu0, double KPmax,
              double CFLScale, double DFLScale, double* __restrict__ dtc, double* __restrict__ dtv,
              int* __restrict__ err)
{
    using D = Dm<N>;
    constexpr int n3 = D::n3;
    constexpr int nw = D::block / 64;
    __shared__ double s_red[6][nw];
    const int e = blockIdx.x, tid = threadIdx.x;
    double lam[6] = {0, 0, 0, 0, 0, 0};
    int bad = 0;
    for (int node = tid; node < n3; node += D::block) {
        const double* u = U + (long)e * 5 * n3 + node;
        const double* ge = geo + (long)e * 10 * n3 + node;
        double rho = u[0], srho = 1.0 / rho;
        double vel[3] = {u[n3] * srho, u[2 * n3] * srho, u[3 * n3] * srho};
        double p = (kappa - 1.0) * (u[4 * n3] - 0.5 * (vel[0] * u[n3] + vel[1] * u[2 * n3] + vel[2] * u[3 * n3]));
        double c = sqrt(kappa * p * srho);
        double sj = ge[9 * n3];
        double vsJ[3] = {vel[0] * sj, vel[1] * sj, vel[2] * sj};
        if (!(rho > 0.0)) bad = 1;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double m0 = ge[(3 * d) * n3], m1 = ge[(3 * d + 1) * n3], m2 = ge[(3 * d + 2) * n3];
            double adv = sj * sqrt(m0 * m0 + m1 * m1 + m2 * m2);
            double lc = fabs(m0 * vsJ[0] + m1 * vsJ[1] + m2 * vsJ[2]) + c * adv;
            double a0 = m0 * sj, a1 = m1 * sj, a2 = m2 * sj;
            double lv = mu0 * srho * (KPmax * (a0 * a0 + a1 * a1 + a2 * a2));
            // running maximum over the thread's nodes, NaN-propagating (one node per thread for N <= 9)
            lam[d] = (lc > lam[d] || lc != lc || node == tid) ? lc : lam[d];
            lam[3 + d] = (lv > lam[3 + d] || lv != lv || node == tid) ? lv : lam[3 + d];
        }
    }
    // wave reduction (max, NaN-propagating)
#pragma unroll
    for (int q = 0; q < 6; q++) {
        double v = lam[q];
        for (int off = 32; off > 0; off >>= 1) {
            double o = __shfl_xor(v, off, 64);
            v = (o > v || o != o) ? o : v;
        }
        if ((tid & 63) == 0) s_red[q][tid >> 6] = v;
    }
    int anybad = __syncthreads_or(bad);
    if (tid == 0) {
        double m[6];
        for (int q = 0; q < 6; q++) {
            double v = s_red[q][0];
            for (int w = 1; w < nw; w++) v = (s_red[q][w] > v || s_red[q][w] != s_red[q][w]) ? s_red[q][w] : v;
            m[q] = v;
        }
        dtc[e] = CFLScale * 2.0 / (m[0] + m[1] + m[2]);
        dtv[e] = DFLScale * 4.0 / (m[3] + m[4] + m[5]);
        if (anybad) atomicOr(err, 1);
    }
}

__global__ void __launch_bounds__(1024) k_dt_reduce(int nElems, const double* __restrict__ dtc,
                                                    const double* __restrict__ dtv, int* __restrict__ err,
                                                    double* __restrict__ out)
{
    __shared__ double s[2][16];
    __shared__ int s_nan[2];
    const int tid = threadIdx.x;
    if (tid < 2) s_nan[tid] = 0;
    __syncthreads();
    double a = 1e300, b = 1e300;
    int na = 0, nb = 0;
    for (int x = tid; x < nElems; x += 1024) {
        double va = dtc[x], vb = dtv[x];
        if (va != va) na = 1; else a = fmin(a, va);
        if (vb != vb) nb = 1; else b = fmin(b, vb);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a = fmin(a, __shfl_xor(a, off, 64));
        b = fmin(b, __shfl_xor(b, off, 64));
    }
    if (na) atomicOr(&s_nan[0], 1);
    if (nb) atomicOr(&s_nan[1], 1);
    if ((tid & 63) == 0) {
        s[0][tid >> 6] = a;
        s[1][tid >> 6] = b;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < 16; w++) {
            a = fmin(a, s[0][w]);
            b = fmin(b, s[1][w]);
        }
        out[0] = a;
        out[1] = b;
        out[2] = a < b ? a : b;
        int ev = err[0] ? 1 : (s_nan[0] ? 2 : (s_nan[1] ? 3 : 0));
        out[3] = double(ev);
    }
}

}  // namespace galx
