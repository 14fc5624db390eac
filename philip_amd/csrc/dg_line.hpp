// Line-structured fused stage kernel for the split (two-point) form of the
// Euler equations on collocated GLL nodes, and of the Navier-Stokes equations
// on the mixed path (V: the viscous volume and surface terms added to the same
// items, navier_stokes.cpp / viscous_numerical_flux.cpp).
//
// The skew-symmetric Hadamard volume term (strong_dg.cpp:1017-1579, Chan 2019
// eq. 15) couples node a only with the nodes b on its own grid line in each
// direction, so one work item per (element, direction K, line) evaluates every
// two-point flux F#(u_a, u_b), a < b, of its line once, keeps the N nodal
// sums of the line in registers (node a gets -Sk[a][b] F, node b gets
// -Sk[b][a] F, the flux being symmetric), adds the two surface numerical
// fluxes at the face nodes (strong_dg.cpp:2368-3576, lift 1/w_0) as items of
// their own, and writes one partial result per node to LDS.  The direction is
// a template parameter of the item's code, so no register array is indexed
// per lane.
//
// Per block: EPB whole elements.
//   A  node-parallel  load the stage state (FP64, coalesced), per-node cache of
//                     derived quantities into LDS; in the same phase the
//                     neighbours' face-node states (one item per face node)
//                     are loaded into LDS, so the line phase makes no global
//                     loads
//   L  item-parallel  pair-line items (a line's pairs, sums kept in registers;
//                     a long line's pairs shared by LS items) and face-node
//                     items (the surface term, written at once over the face
//                     node's trace slot); after a barrier the line sums go to
//                     LDS over the (now dead) node cache, one direction at a
//                     time
//   E  node-parallel  line partials plus surface terms, FP64 epilogue
// T = float: mixed path (float fluxes, FP64 state and RK registers); T = double:
// FP64 path.  Both evaluate the cached-node restatement of the point physics
// (dg_fast.hpp).  The reference's entropy projection u(v(u)) is the identity
// for collocated GLL nodes (the projection M^-1 chi^T W is I) and is not formed.
#pragma once
#include <type_traits>

#include "dg_fast.hpp"
#include "dg_stage2.hpp"

namespace dgx {

// Issue the loads of the low-storage RK registers S2, S3 at the start of the
// kernel (held in registers through the line phase) instead of in the epilogue:
// on for the mixed path (its register budget keeps 4 waves/SIMD), off for FP64,
// which issues them once the line phase is over instead, so that they are in
// flight while the line sums go to LDS (TGV p=3 32^3: 220 -> 208 us per stage;
// p=2 -3 %, p=6 -6 %, p=4 unchanged; at 3D p=5 slower, 366 -> 399 us, so not there).
constexpr bool LINE_PREFETCH_RK_F = true;
constexpr bool LINE_PREFETCH_RK_D = false;
template <int D, int N> constexpr bool line_late_prefetch_d() { return !(D == 3 && N == 6); }
// ES: elements per block relative to the 256-node default (ES > 0: ES times,
// ES < 0: 1/|ES|); per-element strides do not depend on it.  XJ: the block has a
// thread for every item of the line phase (one item pass, the extra threads idle in
// the node phases) instead of one thread per node.
template <int D, int N, typename T, int ES = 1, bool V = false, bool XJ = false> struct LineGeo {
    using G = Geo<D, N>;
    static constexpr int NP = G::NP, NS = D + 2, NL = G::NL;
    static constexpr int EPB = ES > 0 ? G::EPB * ES : (G::EPB / -ES > 0 ? G::EPB / -ES : 1);
    static constexpr bool FAST = std::is_same<T, float>::value;
    // node cache fields: rho, E, vel[D], p, b, |v|^2, sqrt(rho), c, h
    static constexpr int NQ = D + 8;
    // element strides padded to 1 mod 32 words so the stride-N lines of
    // neighbouring elements fall in different LDS banks
    static constexpr int pad1(int x) { return x + (33 - x % 32) % 32; }
    // 3D, N = 4: node cache in the skewed order of qpos (the 16 x-lines of an element would
    // otherwise read 4 words apart, 4 lanes per bank pair); its float element stride is 16 mod 32
    // words so the two elements a 32-lane access spans take complementary bank halves.  At p=3
    // the stage kernel ran 3 % (mixed) and 1.5 % (FP64) faster; at N = 6 the skew was slower.
    static constexpr bool QSKEW = (D == 3 && N == 4);
    static constexpr int pad16(int x) { return x + (48 - x % 32) % 32; }
    // 2D, float: the items of a wave are consecutive lines of consecutive elements, so a line
    // phase along y reads node line + N pos of elements e, e+1, ..: with a stride of 1 mod 32
    // words five to seven lanes met in one bank.  The element stride's residue mod 64 is chosen
    // per N (a search over the 32- and 64-bank layouts and the three elements-per-block tiles)
    // so that no more than two lanes share a bank along x or y.
    static constexpr int res2d(int n) { return n == 3 ? 41 : n == 4 ? 41 : n == 5 ? 25 : n == 6 ? 53 : n == 7 ? 8 : 1; }
    static constexpr int pad2d(int x) { return x + ((res2d(N) - x % 64) % 64 + 64) % 64; }
    static constexpr bool P2D = (D == 2 && FAST && N >= 3);
    static constexpr int ESQ = (QSKEW && FAST) ? pad16(NQ * NP) : (P2D ? pad2d(NQ * NP) : pad1(NQ * NP));
    // Long lines are shared by LS items (the pairs and the viscous columns split evenly), so
    // that the line phase keeps most of the block busy: from p = 4 in 2D and for Navier-Stokes;
    // for 3D Euler at p = 5, and at p = 6 on the FP64 path only (at p = 4 the extra partials
    // cost more than the balance gains; at p = 6 the mixed path gained nothing and its batched
    // kernel's item loop was no longer unrolled, spilling the line sums to scratch)
    static constexpr int LS = (V || D == 2) ? (N >= 7 ? 3 : (N >= 5 ? 2 : 1))
                                            : (N >= 7 ? (FAST ? 1 : 3) : (N >= 6 ? 2 : 1));
    static constexpr int ESV = P2D ? pad2d(LS * NS * NP) : pad1(LS * NS * NP);   // line partials [part][state][node]
    static constexpr int NFN = 2 * D * NL;           // face nodes per element
    // neighbour traces' states [state][face][face node]; each slot then takes that face node's
    // surface term
    static constexpr int ESN = pad1(NS * NFN);
    static constexpr int ITEMS = EPB * D * NL * LS;
    static constexpr int FITEMS = EPB * NFN;
    static constexpr int round64(int x) { return ((x + 63) / 64) * 64; }
    static constexpr int BLOCK = XJ ? round64(EPB * NP > ITEMS + FITEMS ? EPB * NP : ITEMS + FITEMS)
                                    : round64(EPB * NP);
    // Navier-Stokes (V, mixed path only): own viscous flux in every direction
    // [direction][state][node] and the neighbours' viscous flux along the face direction at
    // their face nodes [state][face][face node]
    static_assert(!V || FAST, "the Navier-Stokes line kernel is the mixed-precision path");
    static constexpr int ESFV = V ? pad1(D * NS * NP) : 0;
    static constexpr int ESNV = V ? pad1(NS * NFN) : 0;
    static constexpr int EIN = ESQ + ESN + ESFV + ESNV;
    // [traces / surface terms | node cache | V: viscous fluxes]; the line partials take the node
    // cache's place (and what follows it) once every item has run
    static constexpr int SBUF = EPB * ESN + ((ESQ + ESFV + ESNV > ESV) ? EPB * (ESQ + ESFV + ESNV) : EPB * ESV);
};

// Elements-per-block variant of the single-solve line kernel: 1 for Euler; for
// Navier-Stokes (V) the 256-node default block halved until its LDS fits 64 KB
// (down to one element).
template <int D, int N, typename T, bool V> constexpr int line_es()
{
    if constexpr (V) {
        using L1 = LineGeo<D, N, T, 1, true>;
        constexpr int W = (L1::SBUF / L1::EPB) * (int)sizeof(T);
        int div = 1;
        while (Geo<D, N>::EPB / div > 1 && (Geo<D, N>::EPB / div) * W > 65536) div *= 2;
        return div == 1 ? 1 : -div;
    } else {
        return 1;
    }
}
template <int D, int N, typename T, bool V> using LineGeoK = LineGeo<D, N, T, line_es<D, N, T, V>(), V>;

// ---- node cache in LDS ----------------------------------------------------
// Position of node i (x fastest) in its element's cache.  3D, N = 4: the skewed order
// ((x + y) mod N) + N ((y + z) mod N) + N^2 z, in which every grid line in every direction meets
// each bank group once: with one coordinate fixed the other two map one-to-one onto
// (x + y, y + z), so the N^2 lines of a direction read N^2 different positions mod N^2
// (16 lanes of an element: the 16 distinct 8-byte bank pairs).  Otherwise the identity.
template <int D, int N> __device__ __forceinline__ int qpos(int i)
{
    if constexpr (D == 3 && N == 4) {
        const int x = i % N, y = (i / N) % N, z = i / (N * N);
        return (x + y) % N + N * ((y + z) % N) + N * N * z;
    } else {
        return i;
    }
}

// A node's cache holds rho, E, the velocity, p, beta, |v|^2, sqrt(rho), c and H (D + 8 values);
// the momentum is not stored but rebuilt as rho v on load.  Every flux reads nodes in this
// canonical form (node_canon), whether they come from the cache or from a neighbour trace whose
// cache is built at the face, so a face evaluates to the same bits in every block layout.

template <typename T, int D>
__device__ __forceinline__ void nodef_store(T *q, int st, const NodeC<T, D> &n)
{
    q[0] = n.u[0];
    q[st] = n.u[D + 1];
    for (int k = 0; k < D; ++k) q[(2 + k) * st] = n.vel[k];
    q[(D + 2) * st] = n.p;
    q[(D + 3) * st] = n.b;
    q[(D + 4) * st] = n.v2;
    q[(D + 5) * st] = n.sr;
    q[(D + 6) * st] = n.c;
    q[(D + 7) * st] = n.h;
}

template <typename T, int D> __device__ __forceinline__ void node_canon(NodeC<T, D> &n)
{
    for (int k = 0; k < D; ++k) n.u[1 + k] = n.u[0] * n.vel[k];
}

template <typename T, int D>
__device__ __forceinline__ void nodef_load(const T *q, int st, NodeC<T, D> &n)
{
    n.u[0] = q[0];
    n.u[D + 1] = q[st];
    for (int k = 0; k < D; ++k) n.vel[k] = q[(2 + k) * st];
    n.p = q[(D + 2) * st];
    n.b = q[(D + 3) * st];
    n.v2 = q[(D + 4) * st];
    n.sr = q[(D + 5) * st];
    n.c = q[(D + 6) * st];
    n.h = q[(D + 7) * st];
    node_canon<T, D>(n);
}

// A neighbour's face node from its conservative state (field-major, stride st).
template <typename T, int D>
__device__ __forceinline__ void trace_node(const T *q, int st, const FastConst<T> &fc, NodeC<T, D> &n)
{
    for (int s = 0; s < D + 2; ++s) n.u[s] = q[s * st];
    node_fill<T, D>(n, fc);
    node_canon<T, D>(n);
}

template <typename T, int D>
__device__ __forceinline__ void nodef_select(bool first, const NodeC<T, D> &x, const NodeC<T, D> &y, NodeC<T, D> &o)
{
    for (int s = 0; s < D + 2; ++s) o.u[s] = first ? x.u[s] : y.u[s];
    for (int k = 0; k < D; ++k) o.vel[k] = first ? x.vel[k] : y.vel[k];
    o.p = first ? x.p : y.p;
    o.b = first ? x.b : y.b;
    o.v2 = first ? x.v2 : y.v2;
    o.sr = first ? x.sr : y.sr;
    o.c = first ? x.c : y.c;
    o.h = first ? x.h : y.h;
}

// operator tables in the flux arithmetic type
template <typename T> __device__ __forceinline__ T op_sk(const Ops1D &o, int a, int b)
{
    if constexpr (std::is_same<T, float>::value) return o.Skf[a][b];
    else return o.Sk[a][b];
}
template <typename T> __device__ __forceinline__ T op_d(const Ops1D &o, int a, int b)
{
    if constexpr (std::is_same<T, float>::value) return o.Df[a][b];
    else return o.D[a][b];
}
template <typename T> __device__ __forceinline__ T op_inv_wb(const Ops1D &o)
{
    if constexpr (std::is_same<T, float>::value) return o.inv_wbf;
    else return o.inv_wb;
}

// Pairs of part H of LS: the N(N-1)/2 pairs a < b of a line, in the order (0,1), (0,2), ..,
// (1,2), .., split into LS contiguous ranges (the line's work shared by LS items).
template <int N, int LS, int H> __device__ __forceinline__ constexpr bool pair_in_part(int p0, int p1)
{
    constexpr int NPAIR = N * (N - 1) / 2;
    const int idx = p0 * (2 * N - p0 - 1) / 2 + (p1 - p0 - 1);
    return idx >= H * NPAIR / LS && idx < (H + 1) * NPAIR / LS;
}

template <typename T, int D, int N, int K, int LS = 1, int H = 0>
__device__ __forceinline__ void pairs_f(const Ops1D &ops, const PhysConst &pc, const FastConst<T> &fc, const T *sq,
                                        int line, T (&acc)[N][D + 2])
{
    constexpr int NP = Geo<D, N>::NP, NS = D + 2;
#pragma unroll
    for (int p = 0; p < N; ++p)
        for (int s = 0; s < NS; ++s) acc[p][s] = T(0);
#pragma unroll
    for (int p0 = 0; p0 < N; ++p0) {
#pragma unroll
        for (int p1 = p0 + 1; p1 < N; ++p1) {
            if (!pair_in_part<N, LS, H>(p0, p1)) continue;
            NodeC<T, D> A, B;
            nodef_load<T, D>(sq + qpos<D, N>(line_node<D, N>(K, line, p0)), NP, A);
            nodef_load<T, D>(sq + qpos<D, N>(line_node<D, N>(K, line, p1)), NP, B);
            T F[NS];
            pair_flux_fast<T, D, K>(A, B, fc, pc, F);
            const T s01 = op_sk<T>(ops, p0, p1), s10 = op_sk<T>(ops, p1, p0);
            for (int s = 0; s < NS; ++s) {
                acc[p0][s] -= s01 * F[s];
                acc[p1][s] -= s10 * F[s];
            }
        }
    }
}

// ---- Navier-Stokes terms (either precision) --------------------------------

// Viscous part of the volume term along one line: node a gets -sum_b D[a][b] F_v,K(u_b, q_b)
// (the conservative D-divergence of strong_dg.cpp:1386-1433 for the dissipative flux); sfv
// holds the element's viscous fluxes [direction][state][node].
template <int D, int N, int K, typename T, int LS = 1, int H = 0>
__device__ __forceinline__ void visc_line(const Ops1D &ops, const T *sfv, int line, T (&acc)[N][D + 2])
{
    constexpr int NP = Geo<D, N>::NP, NS = D + 2;
#pragma unroll
    for (int b = H * N / LS; b < (H + 1) * N / LS; ++b) {
        const int nb = line_node<D, N>(K, line, b);
        T fv[NS];
        for (int s = 0; s < NS; ++s) fv[s] = sfv[(K * NS + s) * NP + nb];
#pragma unroll
        for (int a2 = 0; a2 < N; ++a2) {
            const T dab = op_d<T>(ops, a2, b);
            for (int s = 0; s < NS; ++s) acc[a2][s] -= dab * fv[s];
        }
    }
}

// Viscous numerical flux of a face node along sgn e_K (viscous_numerical_flux.cpp:144-328)
// from the one-sided fluxes fa = F_v(ui, qi).e_K, fb = F_v(un, qn).e_K:
// BR2  fvs = {{F_v}} . n;  SIPG also - penalty {{F_v(u, [[u]])}} . n.
template <int D, int K, typename T>
__device__ __forceinline__ void face_visc(const PhysConst &pc, T pen, const T *ui, const T *un, const T *fa,
                                          const T *fb, T sgn, T (&fvs)[D + 2])
{
    constexpr int NS = D + 2;
    for (int s = 0; s < NS; ++s) fvs[s] = T(0.5) * (fa[s] * sgn + fb[s] * sgn);
    if (pc.diss_flux == 0) {
        T jump[NS * D], ja[NS], jb[NS];
        for (int s = 0; s < NS; ++s)
            for (int d = 0; d < D; ++d) jump[s * D + d] = (d == K) ? (ui[s] - un[s]) * sgn : (ui[s] - un[s]) * T(0);
        viscous_flux_dir<T, D>(ui, jump, K, pc, ja);
        viscous_flux_dir<T, D>(un, jump, K, pc, jb);
        for (int s = 0; s < NS; ++s) fvs[s] = fvs[s] - pen * (T(0.5) * (ja[s] * sgn + jb[s] * sgn));
    }
}

// Block data a face item reads (LDS pointers; element strides from LineGeo).
template <typename T> struct FaceIn {
    const T *sq;    // node caches of the block
    const T *sfv;   // viscous flux of the block's nodes, every direction (Navier-Stokes)
    const T *tr;    // this face node's neighbour trace (conservative state, stride tst)
    const T *trv;   // the neighbour's viscous flux along the face direction there (Navier-Stokes)
    int tst;
};

// Surface term of element e at face node fn of its face (K, side): the
// numerical flux with the reference orientation (the lower cell index is the
// interior side), times -1/w_0 for the interior and +1/w_0 for the exterior
// side, so the neighbour's term is the negative.  nbl >= 0: the neighbour is
// element nbl of this block (node cache in LDS); otherwise its trace state is
// tr (field-major, stride tst) and its cache is built here.  Both ways give
// the same canonical node, so a face evaluated in two blocks (once from each
// side) sees the same bits on both and the flux is the same whichever block
// holds the pair.
template <typename T, int D, int N, int K, bool V, bool BND = false>
__device__ __forceinline__ void face_f(const Ops1D &ops, const PhysConst &pc, const FastConst<T> &fc,
                                       const FaceIn<T> &in, int le, int side, int fn, int e, int nb, int nbl,
                                       T pen, T (&out)[D + 2], T (&out_nb)[D + 2])
{
    using LG = LineGeo<D, N, T, 1, V>;
    constexpr int NP = LG::NP, NS = D + 2;
    const int fi = line_node<D, N>(K, fn, side ? N - 1 : 0);
    const int fj = line_node<D, N>(K, fn, side ? 0 : N - 1);
    NodeC<T, D> qi, qn, L, R;
    nodef_load<T, D>(in.sq + le * LG::ESQ + qpos<D, N>(fi), NP, qi);
    bool own = e < nb;
    if constexpr (BND) {
        if (nb < 0) {
            // boundary face of type -nb: the exterior state from the interior node
            T ub[NS];
            boundary_state<T, D>(-nb, qi.u, K, side ? T(1) : T(-1), pc, ub);
            for (int s = 0; s < NS; ++s) qn.u[s] = ub[s];
            node_fill<T, D>(qn, fc);
            node_canon<T, D>(qn);
            own = true;
        } else if (nbl >= 0) {
            nodef_load<T, D>(in.sq + nbl * LG::ESQ + qpos<D, N>(fj), NP, qn);
        } else {
            trace_node<T, D>(in.tr, in.tst, fc, qn);
        }
    } else {
        if (nbl >= 0) nodef_load<T, D>(in.sq + nbl * LG::ESQ + qpos<D, N>(fj), NP, qn);
        else trace_node<T, D>(in.tr, in.tst, fc, qn);
    }
    nodef_select<T, D>(own, qi, qn, L);
    nodef_select<T, D>(own, qn, qi, R);
    const T sgn = side ? T(1) : T(-1);
    T fs[NS];
    num_flux_fast<T, D, K>(L, R, own ? sgn : -sgn, fc, pc, fs);
    const T iwb = op_inv_wb<T>(ops);
    if constexpr (!V) {
        const T w = own ? -iwb : iwb;
        for (int s = 0; s < NS; ++s) {
            out[s] = w * fs[s];
            out_nb[s] = -out[s];
        }
    } else {
        // the viscous surface term: the average is antisymmetric between the two sides, the
        // interior-flux part (strong form) is not, so both sides' values are returned
        T fa[NS], fb[NS], fvs[NS];
        for (int s = 0; s < NS; ++s) {
            fa[s] = in.sfv[le * LG::ESFV + (K * NS + s) * NP + fi];
            fb[s] = (nbl >= 0) ? in.sfv[nbl * LG::ESFV + (K * NS + s) * NP + fj] : in.trv[s * in.tst];
        }
        face_visc<D, K, T>(pc, pen, qi.u, qn.u, fa, fb, sgn, fvs);
        const T fsgn = own ? T(1) : T(-1);
        for (int s = 0; s < NS; ++s) {
            const T A = -(fs[s] * fsgn) - fvs[s];
            out[s] = iwb * (A + sgn * fa[s]);
            out_nb[s] = iwb * ((-sgn * fb[s]) - A);
        }
    }
}

// One block's work: elements blk0 .. blk0 + EPB - 1 of the solve described by a.
// V: Navier-Stokes (split convective form plus the viscous volume and surface
// terms, with the lifted gradient a.aux computed beforehand).
template <int D, int N, typename T, int ES, bool V = false, bool XJ = false, bool BND = false>
__device__ __forceinline__ void line_block(const StageArgs &a, const Ops1D &ops, const PhysConst &pc, const int blk0)
{
    using LG = LineGeo<D, N, T, ES, V, XJ>;
    constexpr int NP = LG::NP, NS = LG::NS, NL = LG::NL, EPB = LG::EPB, BLOCK = LG::BLOCK, LS = LG::LS;
    constexpr bool FAST = LG::FAST;
    constexpr int NFI = (LG::FITEMS + BLOCK - 1) / BLOCK;  // face-trace loads per thread
    constexpr int NFS = EPB * 2 * D;                       // face sides of the block
    constexpr int NWAVE = BLOCK / 64;
    // A face item writes its surface term over its own trace slot as soon as it has it; the
    // node cache (and the viscous fluxes) are dead once every item has run, and the line
    // partials take their storage.
    __shared__ T sBuf[LG::SBUF];
    T *const sN = sBuf;                       // traces, then surface terms [element][state][face][face node]
    T *const sQ = sN + EPB * LG::ESN;         // per-node cache [element][field][node]
    T *const sFv = sQ + EPB * LG::ESQ;        // V: viscous flux [element][direction][state][node]
    T *const sNv = sFv + EPB * LG::ESFV;      // V: neighbours' viscous flux, as the traces
    T *const sV = sQ;                         // line partials [element][part][state][node]
    __shared__ int sNb[EPB][2 * D];           // face neighbours
    __shared__ short sList[NWAVE][NFS];       // per-wave list of the faces evaluated in this block

    const int tid = threadIdx.x;
    const FastConst<T> fc = fast_const<T>(pc);
    // a neighbour is read from the block's own node cache when it is another element of the block
    auto in_block = [&](int nb, int fe) { return nb >= blk0 && nb < blk0 + EPB && nb != fe; };

    // ---------------- phase A: loads first, then the node cache ----------
    const int le = tid / NP, i = tid % NP;
    const int e = blk0 + le;
    const bool active = (le < EPB) && (e < a.ncells);
    // neighbour indices first: the trace loads depend on them
    int nbv[NFI];
#pragma unroll
    for (int j = 0; j < NFI; ++j) {
        const int t = tid + j * BLOCK;
        const int fe = blk0 + t / LG::NFN;
        nbv[j] = (t < LG::FITEMS && fe < a.ncells) ? a.nbr[fe * 2 * D + (t % LG::NFN) / NL] : 0;
    }
    double ud[NS];
    double un[NFI][NS];
    int fdst[NFI];
    if (active)
        for (int s = 0; s < NS; ++s) ud[s] = a.u[((long)e * NS + s) * NP + i];
    double pf2[NS], pf3[NS];
    constexpr bool PREFETCH = FAST ? LINE_PREFETCH_RK_F : LINE_PREFETCH_RK_D;
    const bool pf = PREFETCH && active && a.mode == MODE_LSRK && !a.lsrk_first;
    if (pf)
        for (int s = 0; s < NS; ++s) {
            const long idx = ((long)e * NS + s) * NP + i;
            pf2[s] = ld_stream(a.s2 + idx);  // streamed once: keep L2 for neighbour traces
            pf3[s] = ld_stream(a.s3 + idx);
        }
    // neighbour face traces (one item per face node), loaded together with the own state
#pragma unroll
    for (int j = 0; j < NFI; ++j) {
        const int t = tid + j * BLOCK;
        fdst[j] = -1;
        if (t < LG::FITEMS) {
            const int fle = t / LG::NFN, r = t % LG::NFN;
            const int fe = blk0 + fle;
            if (fe < a.ncells) {
                const int f = r / NL, fn = r % NL;
                const int nb = nbv[j];
                if (fn == 0) sNb[fle][f] = nb;
                if (nb >= 0 && !in_block(nb, fe)) {
                    const int fj = line_node<D, N>(f >> 1, fn, (f & 1) ? 0 : N - 1);
                    for (int s = 0; s < NS; ++s) un[j][s] = a.u[((long)nb * NS + s) * NP + fj];
                    fdst[j] = t;  // = fle * NFN + r
                }
            }
        }
    }
    if (active) {
        NodeC<T, D> n;
        for (int s = 0; s < NS; ++s) n.u[s] = T(ud[s]);
        node_fill<T, D>(n, fc);
        flag_nonphysical<T>(a, n.u[0], n.p);
        nodef_store<T, D>(sQ + le * LG::ESQ + qpos<D, N>(i), NP, n);
    }
#pragma unroll
    for (int j = 0; j < NFI; ++j) {
        if (fdst[j] < 0) continue;
        T *q = sN + (fdst[j] / LG::NFN) * LG::ESN + fdst[j] % LG::NFN;
        for (int s = 0; s < NS; ++s) q[s * LG::NFN] = T(un[j][s]);
    }
    if constexpr (V) {
        // viscous fluxes from the lifted gradient: the own nodes' in every direction, the
        // neighbours' face nodes' along the face direction
        if (active) {
            T g[NS * D], ut[NS], fv[D][NS];
#pragma unroll
            for (int x = 0; x < NS * D; ++x) g[x] = aux_at<T>(a, x % D, ((long)e * NS + x / D) * NP + i);
            for (int s = 0; s < NS; ++s) ut[s] = T(ud[s]);
            viscous_flux_all<T, D>(ut, g, pc, fv);
#pragma unroll
            for (int k = 0; k < D; ++k)
                for (int s = 0; s < NS; ++s) sFv[le * LG::ESFV + (k * NS + s) * NP + i] = fv[k][s];
        }
#pragma unroll
        for (int j = 0; j < NFI; ++j) {
            if (fdst[j] < 0) continue;
            const int t = tid + j * BLOCK;
            const int f = (t % LG::NFN) / NL, fn = t % NL;
            const int fj = line_node<D, N>(f >> 1, fn, (f & 1) ? 0 : N - 1);
            const long nb = nbv[j];
            T g[NS * D], ur[NS], fv[D][NS];
#pragma unroll
            for (int x = 0; x < NS * D; ++x) g[x] = aux_at<T>(a, x % D, (nb * NS + x / D) * NP + fj);
            for (int s = 0; s < NS; ++s) ur[s] = T(un[j][s]);
            viscous_flux_all<T, D>(ur, g, pc, fv);
            T fk[NS];
            for (int s = 0; s < NS; ++s) fk[s] = fv[0][s];
#pragma unroll
            for (int k = 1; k < D; ++k)
                for (int s = 0; s < NS; ++s) fk[s] = (f >> 1) == k ? fv[k][s] : fk[s];
            for (int s = 0; s < NS; ++s)
                sNv[(fdst[j] / LG::NFN) * LG::ESNV + fdst[j] % LG::NFN + s * LG::NFN] = fk[s];
        }
    }
    __syncthreads();

    // ---------------- faces evaluated in this block ----------------------
    // Face side (le, f) is evaluated here unless its neighbour is an element
    // of this block with a lower index (that element evaluates the shared face
    // once for both).  Each wave builds the same list in its own LDS row,
    // ordered by direction so that a wave's face items mostly share one.
    const int lane = tid & 63, wv = tid >> 6;
    int nface = 0;
    for (int c0 = 0; c0 < NFS; c0 += 64) {
        const int id = c0 + lane;  // id = (k * EPB + le) * 2 + side
        bool want = false;
        if (id < NFS) {
            const int kk = id / (2 * EPB), r = id % (2 * EPB);
            const int fle = r >> 1, f = 2 * kk + (r & 1);
            if (blk0 + fle < a.ncells) {
                const int nb = sNb[fle][f];
                want = !(in_block(nb, blk0 + fle) && nb < blk0 + fle);
            }
        }
        const unsigned long long m = __ballot(want);
        if (want) sList[wv][nface + __popcll(m & ((1ull << lane) - 1ull))] = (short)id;
        nface += __popcll(m);
    }

    // ---------------- phase L: pair-line items, then face items ----------
    constexpr int PITEMS = LG::ITEMS;                          // (direction, part, element, line)
    constexpr int JP = (PITEMS + BLOCK - 1) / BLOCK;           // slots that can hold a line item
    constexpr int JT = (PITEMS + NFS * NL + BLOCK - 1) / BLOCK;
    // V with in-block neighbours: the neighbour's side of a shared face is not the negated own
    // side (the interior viscous flux is one-sided), so it is held separately
    constexpr bool NBT = V && EPB > 1;
    const int items = PITEMS + nface * NL;
    const T pen = T(a.penalty);
    T rp[JP][N][NS];
#pragma unroll
    for (int j = 0; j < JT; ++j) {
        const int t = tid + j * BLOCK;
        if (j < JP && t < PITEMS) {
            // t = ((k * LS + h) * EPB + element) * NL + line: direction k, part h of the line
            const int kh = t / (EPB * NL), r = t % (EPB * NL);
            const int k = kh / LS, h = kh % LS;
            const int lle = r / NL, line = r % NL;
            if (blk0 + lle >= a.ncells) continue;
            const T *q = sQ + lle * LG::ESQ;
            auto pitem = [&](auto kc, auto hc) {
                constexpr int K = decltype(kc)::value, H = decltype(hc)::value;
                pairs_f<T, D, N, K, LS, H>(ops, pc, fc, q, line, rp[j < JP ? j : 0]);
                if constexpr (V) visc_line<D, N, K, T, LS, H>(ops, sFv + lle * LG::ESFV, line, rp[j < JP ? j : 0]);
            };
            auto pdir = [&](auto kc) {
                if (h == 0) pitem(kc, std::integral_constant<int, 0>{});
                else if (LS > 1 && h == 1) pitem(kc, std::integral_constant<int, (LS > 1 ? 1 : 0)>{});
                else if (LS > 2) pitem(kc, std::integral_constant<int, (LS > 2 ? 2 : 0)>{});
            };
            if (k == 0) pdir(std::integral_constant<int, 0>{});
            else if (k == 1) pdir(std::integral_constant<int, 1>{});
            else if (D == 3) pdir(std::integral_constant<int, (D == 3 ? 2 : 1)>{});
        } else if (t >= PITEMS && t < items) {
            const int u = (t - PITEMS) / NL, fn = (t - PITEMS) % NL;
            const int id = sList[wv][u];
            const int k = id / (2 * EPB), r = id % (2 * EPB);
            const int fle = r >> 1, side = r & 1;
            const int nb = sNb[fle][2 * k + side];
            const int nbl = in_block(nb, blk0 + fle) ? nb - blk0 : -1;
            const int fo = (2 * k + side) * NL + fn;
            T *const slot = sN + fle * LG::ESN + fo;
            const FaceIn<T> in{sQ, sFv, slot, sNv + fle * LG::ESNV + fo, LG::NFN};
            T rf[NS], rfn[NS];
            auto fitem = [&](auto kc) {
                constexpr int K = decltype(kc)::value;
                face_f<T, D, N, K, V, BND>(ops, pc, fc, in, fle, side, fn, blk0 + fle, nb, nbl, pen, rf, rfn);
            };
            if (k == 0) fitem(std::integral_constant<int, 0>{});
            else if (k == 1) fitem(std::integral_constant<int, 1>{});
            else if (D == 3) fitem(std::integral_constant<int, (D == 3 ? 2 : 1)>{});
            // the trace slot was read only by this item; a neighbour in the block has no trace on
            // the shared face (and no item of its own for it), so its slot takes its term
            for (int s = 0; s < NS; ++s) slot[s * LG::NFN] = rf[s];
            if (nbl >= 0) {
                T *const w2 = sN + nbl * LG::ESN + ((2 * k + side) ^ 1) * NL + fn;
                for (int s = 0; s < NS; ++s) w2[s * LG::NFN] = NBT ? rfn[s] : -rf[s];
            }
        }
    }
    __syncthreads();  // node cache and traces consumed; their storage now takes the partials
    constexpr bool LATE_PF = !PREFETCH && !FAST && line_late_prefetch_d<D, N>();
    const bool late_pf = LATE_PF && active && a.mode == MODE_LSRK && !a.lsrk_first;
    // FP64 at N >= 4: the stage state is read again here for the epilogue (an L2 hit: the block
    // read it in phase A) instead of being held in registers through the line phase (TGV 32^3
    // p=3: 177 -> 170 us per stage, 64 -> 16 B/lane of spills; p=4 24^3 294 -> 244 us; at p=2
    // slower)
    constexpr bool REREAD = LATE_PF && N >= 4;
    double ue[NS];
    if (REREAD && active && a.mode == MODE_LSRK)
        for (int s = 0; s < NS; ++s) ue[s] = a.u[((long)e * NS + s) * NP + i];
    if (late_pf)
        for (int s = 0; s < NS; ++s) {
            const long idx = ((long)e * NS + s) * NP + i;
            pf2[s] = ld_stream(a.s2 + idx);
            pf3[s] = ld_stream(a.s3 + idx);
        }
    // Line partials of one direction at a time into one [part][state][node] buffer: direction 0
    // stores, each later direction adds after a barrier, so every part's sum is ((d0 + d1) + d2)
    // in this order in every block layout.
    auto line_out = [&](auto kc, bool add) {
        constexpr int KD = decltype(kc)::value;
#pragma unroll
        for (int j = 0; j < JP; ++j) {
            const int t = tid + j * BLOCK;
            if (t >= PITEMS) continue;
            const int kh = t / (EPB * NL), r = t % (EPB * NL);
            if (kh / LS != KD) continue;
            const int lle = r / NL, line = r % NL;
            if (blk0 + lle >= a.ncells) continue;
            T *v = sV + lle * LG::ESV + (kh % LS) * NS * NP;
#pragma unroll
            for (int p = 0; p < N; ++p) {
                const int n = line_node<D, N>(KD, line, p);
                for (int s = 0; s < NS; ++s) {
                    if (add) v[s * NP + n] += rp[j][p][s];
                    else v[s * NP + n] = rp[j][p][s];
                }
            }
        }
    };
    line_out(std::integral_constant<int, 0>{}, false);
    __syncthreads();
    line_out(std::integral_constant<int, 1>{}, true);
    if constexpr (D == 3) {
        __syncthreads();
        line_out(std::integral_constant<int, (D == 3 ? 2 : 1)>{}, true);
    }
    __syncthreads();
    if (!active) return;

    // ---------------- phase E: volume + surface terms, FP64 epilogue ------
    int c[3] = {0, 0, 0};
    node_coords<D, N>(i, c);
    T acc[NS];
    const T *v = sV + le * LG::ESV + i;
    for (int s = 0; s < NS; ++s) {
        T xs = v[s * NP];
        for (int h = 1; h < LS; ++h) xs += v[(h * NS + s) * NP];
        acc[s] = xs;
    }
    const T *fb = sN + le * LG::ESN;
    for (int k = 0; k < D; ++k) {
        if (c[k] != 0 && c[k] != N - 1) continue;
        const int f = 2 * k + (c[k] == N - 1 ? 1 : 0);
        const T *w = fb + f * NL + node_line<D, N>(k, i);
        for (int s = 0; s < NS; ++s) acc[s] += w[s * LG::NFN];
    }
    if constexpr (REREAD) stage_epilogue<D, N, T>(a, ops, e, i, c, ue, acc, late_pf, pf2, pf3);
    else stage_epilogue<D, N, T>(a, ops, e, i, c, ud, acc, pf || late_pf, pf2, pf3);
}

// Elements-per-block variant (ES of LineGeo) of the single-solve Euler line kernel: the mixed path
// in 3D takes twice the default (at p=3 a 2x2x2 Morton brick, so that more faces stay inside the
// block): p=3 32^3 110 -> 102.5 us per stage, p=4 24^3 125 -> 122 us, p=6 16^3 277 -> 181 us; at
// p=2 it was slower (90 -> 95 us) and at p=5 no faster; FP64 at p=3 was slower (179 -> 216 us).
template <int D, int N, typename T> constexpr int line_es_single()
{
    return (D == 3 && (N == 4 || N == 5 || N == 7) && std::is_same<T, float>::value) ? 2 : 1;
}
template <int D, int N, typename T> using LineGeoS = LineGeo<D, N, T, line_es_single<D, N, T>()>;

// One solve: blocks in XCD-aware order over the element groups.  BND: the mesh has boundary
// faces (a separate instantiation, so the periodic kernels carry no boundary code).
template <int D, int N, typename T, bool BND = false>
__global__ void __launch_bounds__((LineGeoS<D, N, T>::BLOCK))
line_kernel(StageArgs a, Ops1D ops, PhysConst pc)
{
    line_block<D, N, T, line_es_single<D, N, T>(), false, false, BND>(a, ops, pc,
                                                                       xcd_block() * LineGeoS<D, N, T>::EPB);
}
// Waves per SIMD the single-solve 3D line kernel is held to (amdgpu_waves_per_eu; 0: the
// compiler's choice).  The mixed path keeps the compiler's choice (a 4-wave bound at p=6 was
// faster with an earlier, larger LDS layout; it now measures 296 against 278 us per stage at 16^3).
// FP64 path at p=3..5: 3 (168 VGPRs, ~64 B/lane spilled at p=3,
// two -> three blocks per CU): TGV p=3 32^3 258 -> 222 us per stage, p=4 24^3 327 -> 305 us,
// p=5 20^3 398 -> 370 us (4 waves at p=3: 395 us); at p=2 no change and at p=6 much slower
// (448 -> 724 us), so those keep the compiler's choice.
template <int D, int N, typename T> constexpr int line_waves()
{
    if constexpr (D == 3 && N >= 4 && N <= 6 && std::is_same<T, double>::value) return 3;
    return 0;
}
template <int D, int N, typename T, int W, bool BND = false>
__global__ void __launch_bounds__((LineGeoS<D, N, T>::BLOCK)) __attribute__((amdgpu_waves_per_eu(W)))
line_kernel_wb(StageArgs a, Ops1D ops, PhysConst pc)
{
    line_block<D, N, T, line_es_single<D, N, T>(), false, false, BND>(a, ops, pc,
                                                                       xcd_block() * LineGeoS<D, N, T>::EPB);
}

// Navier-Stokes, mixed path.  line_kernel_ns_w4 is held to 4 waves/SIMD (128 VGPRs; a few
// values spill): used where that measured faster, p=5 (142 -> 122 us per stage at 16^3); at
// p=2..4 and p=6 it was not (p=3: 93 -> 98 us), so those keep the compiler's choice.
template <int D, int N>
__global__ void __launch_bounds__((LineGeoK<D, N, float, true>::BLOCK))
line_kernel_ns(StageArgs a, Ops1D ops, PhysConst pc)
{
    line_block<D, N, float, line_es<D, N, float, true>(), true>(a, ops, pc,
                                                                xcd_block() * LineGeoK<D, N, float, true>::EPB);
}
template <int D, int N>
__global__ void __launch_bounds__((LineGeoK<D, N, float, true>::BLOCK)) __attribute__((amdgpu_waves_per_eu(4)))
line_kernel_ns_w4(StageArgs a, Ops1D ops, PhysConst pc)
{
    line_block<D, N, float, line_es<D, N, float, true>(), true>(a, ops, pc,
                                                                xcd_block() * LineGeoK<D, N, float, true>::EPB);
}

// Many independent solves of one (dim, p, precision) in one launch: block b
// (in XCD-aware order) takes element group blocks[b] of solve blocks[b].inst,
// whose stage arguments and physical constants are read from device tables.
struct BatchBlock {
    int inst, blk0;
};
template <int D, int N, typename T, int ES, bool XJ>
__global__ void __launch_bounds__((LineGeo<D, N, T, ES, false, XJ>::BLOCK))
line_kernel_batch(const BatchBlock *__restrict__ blocks, const StageArgs *__restrict__ args,
                  const PhysConst *__restrict__ pcs, Ops1D ops)
{
    const BatchBlock bb = blocks[xcd_block()];
    line_block<D, N, T, ES, false, XJ>(args[bb.inst], ops, pcs[bb.inst], bb.blk0);
}
// The same held to W waves/SIMD (amdgpu_waves_per_eu): a candidate tile of the batched
// autotuner.
template <int D, int N, typename T, int ES, int W, bool XJ>
__global__ void __launch_bounds__((LineGeo<D, N, T, ES, false, XJ>::BLOCK)) __attribute__((amdgpu_waves_per_eu(W)))
line_kernel_batch_wb(const BatchBlock *__restrict__ blocks, const StageArgs *__restrict__ args,
                     const PhysConst *__restrict__ pcs, Ops1D ops)
{
    const BatchBlock bb = blocks[xcd_block()];
    line_block<D, N, T, ES, false, XJ>(args[bb.inst], ops, pcs[bb.inst], bb.blk0);
}

// Tiles the batched autotuner times per (dim, nodes per direction, precision): elements per
// block (ES: 1 = the 256-node default, 2 = twice as many, -2 = half as many) x a waves/SIMD bound
// (0 = the compiler's register allocation; else amdgpu_waves_per_eu(W), trading spills for
// occupancy) x threads per block (one per node, or one per line-phase item: XJ).  None changes
// the arithmetic of an element (the line split LS is fixed per shape for that reason: it
// reorders the pair sums), so every tile gives the single solve's bits.
struct BatchTile {
    int es, waves;
    bool xj;
};
constexpr int BATCH_TILES_2D = 12, BATCH_TILES_3D = 2;
constexpr int batch_tile_count(int dim) { return dim == 2 ? BATCH_TILES_2D : BATCH_TILES_3D; }
constexpr int batch_waves(int dim, bool fast) { return dim == 2 ? (fast ? 5 : 4) : 0; }
constexpr BatchTile batch_tile(int dim, bool fast, int v)
{
    const int es = (v % 3 == 0) ? 1 : (v % 3 == 1 ? 2 : -2);
    return BatchTile{es, (v / 3) % 2 ? batch_waves(dim, fast) : 0, v >= 6};
}

} // namespace dgx
