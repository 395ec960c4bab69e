// Plain fused element kernel: one pass per element block computes Flou's rhs! for the
// element and applies the mass matrix and, optionally, an RK stage combination.
//
//   Hyperbolic.jl:31-69  fill! -> project2faces! -> volume -> applyBCs! ->
//                        interface_fluxes! -> surface -> apply_massmatrix!
//
// Arithmetic follows the reference term by term (see physics_plain.hpp); the FP64
// translation units are built with -ffp-contract=off so that every value is rounded where
// Julia rounds it.  Per element the kernel:
//   1. stages the element's nodal block (all variables) in LDS, coalesced per variable;
//   2. caches per-node primitives (Euler) and, for the strong form, the contravariant
//      volume fluxes in LDS;
//   3. evaluates the numerical flux on every node of its 2d faces from its own trace
//      (LDS) and the neighbour's trace (read from HBM/L2), on the master side's frame,
//      so both elements of a face obtain the same bits Flou's face loop stores in
//      Fn.sides[1] / Fn.sides[2] (Interfaces.jl:111-136);
//   4. accumulates volume terms (directions 1..d, columns in order), the FR lift
//      (OpDivergence.jl:42-100) and divides by J, one thread per node.
// Compile with -DFLOUHIP_PREC_F64 or -DFLOUHIP_PREC_F32 and -DFLOUHIP_ND=1|2|3.
#include "common.hpp"
#include <type_traits>
#include "physics_plain.hpp"

namespace flouhip {

// HybridDivOperator on Gauss nodes (_surf_hybrid_nb_tensorproduct!, OpDivergence.jl:732-779),
// one direction, after the two-point fluxes with the projected end states (sNf) and the row
// sums (D♯ F♯)_i ω_i (sHs) are in LDS:
//   (a) Fl_i -= l·Fl + Fn_l, Fr_i -= r·Fr - Fn_r, in place once every node has read the line;
//   (b) one thread per line runs the subcell recurrence from F̄[0] = -Fn_l:
//         F̄[p+1] = ((F̄[p] + (D♯F♯)_p ω_p) - l[p] Fl_p) + r[p] Fr_p, blended with the FV flux
//       of positions p, p+1; F̄[p+1] replaces the consumed row sum of position p;
//   (c) dQ[i] += (F̄[ii] - F̄[ii+1]) / ω[ii] with F̄[N] = Fn_r.
// Every thread of the block calls it (it holds barriers); the caller synchronises after it.
template <typename T, int ND, int N, int NV, int EPB, int TPB>
__device__ void hybrid_nb_line_phases(const RhsArgs<T>& a, const T* sOps, T (&sQ)[EPB][NV][Cfg<ND, N>::NPE],
                                      T (&sNf)[EPB][2][NV][Cfg<ND, N>::NPE], T (&sHs)[EPB][NV][Cfg<ND, N>::NPE],
                                      T (&sFn)[EPB][2 * ND][NV][Cfg<ND, N>::NPF], int dir, int e0, int le, int node,
                                      bool valid, int tid, T (&acc)[NV]) {
    constexpr int NPF = Cfg<ND, N>::NPF;
    const T* lm = sOps + ops_lm(N);
    const T* lp = sOps + ops_lp(N);
    const T* w = sOps + ops_w(N);
    int f = 0, ii = 0;
    T fl[NV], fr[NV];
    if (valid) {
        node_line<ND, N>(node, dir, f, ii);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int n0 = line_node<ND, N>(dir, f, 0);
            T lF = lm[0] * sNf[le][0][v][n0], rF = lp[0] * sNf[le][1][v][n0];
            for (int k = 1; k < N; ++k) {
                const int nk = line_node<ND, N>(dir, f, k);
                lF = lF + lm[k] * sNf[le][0][v][nk];
                rF = rF + lp[k] * sNf[le][1][v][nk];
            }
            fl[v] = sNf[le][0][v][node] - (lF + sFn[le][2 * dir][v][f]);
            fr[v] = sNf[le][1][v][node] - (rF - sFn[le][2 * dir + 1][v][f]);
        }
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            sNf[le][0][v][node] = fl[v];
            sNf[le][1][v][node] = fr[v];
        }
    }
    __syncthreads();
    T gs[3 * ND + 1];
    plain::subgrid_frame<T, ND>(dir, a.sjac[dir], gs);
    for (int it = tid; it < EPB * NPF; it += TPB) {
        const int le2 = it / NPF, f2 = it % NPF;
        if (e0 + le2 >= a.ne) continue;
        T Fc[NV], Wl[NV], q0[NV];
        const int n0 = line_node<ND, N>(dir, f2, 0);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            Fc[v] = -sFn[le2][2 * dir][v][f2];
            q0[v] = sQ[le2][v][n0];
        }
        plain::cons2entropy<T, ND>(q0, a.gamma, a.errflag, Wl);
        for (int p = 0; p + 1 < N; ++p) {
            const int ni = line_node<ND, N>(dir, f2, p), nr = line_node<ND, N>(dir, f2, p + 1);
            T Fx[NV], Ql[NV], Qr[NV], Rl[NV], Rr[NV], Fn[NV], Fv[NV], Wr[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                Fx[v] = ((Fc[v] + sHs[le2][v][ni]) - lm[p] * sNf[le2][0][v][ni]) + lp[p] * sNf[le2][1][v][ni];
                Ql[v] = sQ[le2][v][ni];
                Qr[v] = sQ[le2][v][nr];
            }
            plain::rotate2face<T, ND>(Ql, gs, Rl);
            plain::rotate2face<T, ND>(Qr, gs, Rr);
            plain::euler_numericalflux<T, ND>(Rl, Rr, a.savg, a.sdiss, a.typo, a.intensity, a.gamma, a.errflag,
                                              Fn);
            plain::rotate2phys<T, ND>(Fn, gs, Fv);
#pragma unroll
            for (int v = 0; v < NV; ++v) Fv[v] = Fv[v] * gs[3 * ND];
            plain::cons2entropy<T, ND>(Qr, a.gamma, a.errflag, Wr);
            T b = (Wr[0] - Wl[0]) * (Fx[0] - Fv[0]);
#pragma unroll
            for (int v = 1; v < NV; ++v) b = b + (Wr[v] - Wl[v]) * (Fx[v] - Fv[v]);
            const T dl = plain::hybrid_delta(b, a.blend, a.errflag);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                Fx[v] = (T(1) - dl) * Fv[v] + dl * Fx[v];
                sHs[le2][v][ni] = Fx[v];
                Fc[v] = Fx[v];
                Wl[v] = Wr[v];
            }
        }
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const T l = ii == 0 ? -sFn[le][2 * dir][v][f] : sHs[le][v][line_node<ND, N>(dir, f, ii - 1)];
            const T r = ii == N - 1 ? sFn[le][2 * dir + 1][v][f] : sHs[le][v][node];
            acc[v] = acc[v] + (l - r) / w[ii];
        }
    }
}

template <typename T, int ND, int N, int EQ, int VOL, int TP, bool CURV>
__global__ void __launch_bounds__((Cfg<ND, N>::TPB)) k_plain(const RhsArgs<T> a) {
    using C = Cfg<ND, N>;
    constexpr int NPE = C::NPE, NPF = C::NPF, NF = C::NF, EPB = C::EPB, TPB = C::TPB;
    constexpr int NV = EQ == EQ_ADVECTION ? 1 : ND + 2;
    constexpr int NP = EQ == EQ_EULER ? NV + 1 : 1;
    constexpr int NFT = VOL == VOL_STRONG ? ND * NV : 1;
    constexpr int NOPS = ops_size(N);
    constexpr int NG = ND * ND + 1;

    __shared__ T sOps[NOPS];
    __shared__ T sG[EPB][NG];
    __shared__ T sQ[EPB][NV][NPE];
    __shared__ T sP[EPB][NP][NPE];
    __shared__ T sFt[EPB][NFT][NPE];
    __shared__ T sFn[EPB][NF][NV][NPF];
    // HybridDivOperator: subcell-interface fluxes of one direction, [le][v][line][s]
    constexpr int HB = VOL == VOL_HYBRID ? 1 : 0;
    __shared__ T sFb[HB ? EPB : 1][HB ? NV : 1][HB ? NPF : 1][HB ? N : 1];
    // Split and hybrid forms on Gauss nodes: projected end states (Q, primitives, β) of every
    // line and the two-point fluxes of every node with them, one direction at a time
    constexpr int NB = (VOL == VOL_SPLIT_NB || VOL == VOL_HYBRID_NB) ? 1 : 0;
    __shared__ T sNb[NB ? EPB : 1][NB ? NPF : 1][2][NB ? 2 * NV + 1 : 1];
    __shared__ T sNf[NB ? EPB : 1][2][NB ? NV : 1][NB ? NPE : 1];
    // Hybrid form on Gauss nodes: per node the row sum (D♯ F♯)_i ω_i, overwritten along each
    // line by the subcell flux of the interface to the node's right
    constexpr int HN = VOL == VOL_HYBRID_NB ? 1 : 0;
    __shared__ T sHs[HN ? EPB : 1][HN ? NV : 1][HN ? NPE : 1];
    // Split form (Euler, GLL): the two-point fluxes of one direction, each pair of a line
    // evaluated once by the whole block ([le][line][pair][v], pairs (a < b) in row order)
    constexpr int PS = (VOL == VOL_SPLIT && EQ == EQ_EULER) ? 1 : 0;
    constexpr int NPAIR = N * (N - 1) / 2;
    __shared__ T sPair[PS ? EPB : 1][PS ? NPF * NPAIR : 1][PS ? NV : 1];

    const int tid = threadIdx.x;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int e0 = blk * EPB;
    const int64_t Nd = a.ndofs;
    const T avel[3] = {a.a0, a.a1, a.a2};

    // Split form, Euler, 6 or more nodes per direction, FP64, and FP32 except 3-D N=5 (the
    // registers fit without lowering the occupancy there): the face node this thread takes first in the face phase
    // (it = tid) has its face tables loaded before the element's nodes, and on GLL nodes
    // (gather) its neighbour trace, face frame and surface J after them, together with the
    // stage epilogue's second operand (RK register or SSP base) of this thread's node.  All
    // are then in flight during the staging instead of after it.
    constexpr bool PF = EQ == EQ_EULER && VOL == VOL_SPLIT && N >= 6 && (std::is_same<T, double>::value || !(ND == 3 && N == 6));
    constexpr int NFN = EPB * NF * NPF;  // face nodes of the block
    const int pfk = (tid % (NF * NPF)) / NPF, pff = tid % NPF, pfe = e0 + tid / (NF * NPF);
    const bool pf = PF && tid < NFN && pfe < a.ne;
    int pf_info = 0, pf_nbe = -1, pf_fgi = 0;
    if (pf) {
        const int slot = pfe * NF + pfk;
        pf_info = a.finfo[slot];
        pf_nbe = a.nb[slot];
        pf_fgi = a.fgi[slot];
    }
    constexpr int NLD = PF ? (NV * EPB * NPE + TPB - 1) / TPB : 1;  // staging loads per thread
    T ld[NLD];
    if constexpr (PF) {
#pragma unroll
        for (int r = 0; r < NLD; ++r) {
            const int i = tid + r * TPB;
            const int v = i / (EPB * NPE), j = i % (EPB * NPE);
            ld[r] = (i < NV * EPB * NPE && e0 + j / NPE < a.ne) ? a.Q[v * Nd + (int64_t)e0 * NPE + j] : T(0);
        }
    }
    T pf_ext[NV], pf_g[3 * ND + 1], pf_aux[NV];
    if (pf && a.gather && pf_nbe >= 0) {
        const int kp = finfo_nbpos(pf_info);
        const T* Qn = a.Q + (int64_t)pf_nbe * NPE;
        const int nn = line_node<ND, N>(kp >> 1, face_nb_node<ND, N>(pff, pf_info), (kp & 1) ? N - 1 : 0);
#pragma unroll
        for (int v = 0; v < NV; ++v) pf_ext[v] = Qn[v * Nd + nn];
    }
    if (pf) {
        const T* g = a.fgeom +
                     ((int64_t)pf_fgi * (CURV ? NPF : 1) + (CURV ? face_master_node<ND, N>(pff, pf_info) : 0)) *
                         (3 * ND + 1);
#pragma unroll
        for (int c = 0; c < 3 * ND + 1; ++c) pf_g[c] = g[c];
    }
    const T* aux = (a.mode == MODE_2N && !a.first) ? a.reg
                 : ((a.mode == MODE_SSP2 || a.mode == MODE_SSP3) ? a.ubase : nullptr);
    if (PF && aux && e0 + tid / NPE < a.ne) {
#pragma unroll
        for (int v = 0; v < NV; ++v) pf_aux[v] = aux[v * Nd + (int64_t)e0 * NPE + tid];
    }
    for (int i = tid; i < NOPS; i += TPB) sOps[i] = a.ops[i];
    for (int i = tid; i < EPB * NG; i += TPB) {
        const int le = i / NG, c = i % NG, e = e0 + le;
        if (e < a.ne) sG[le][c] = a.egeom[(int64_t)e * a.egeom_stride + c];
    }
    if constexpr (PF) {
#pragma unroll
        for (int r = 0; r < NLD; ++r) {
            const int i = tid + r * TPB;
            const int v = i / (EPB * NPE), j = i % (EPB * NPE);
            if (i < NV * EPB * NPE && e0 + j / NPE < a.ne) sQ[j / NPE][v][j % NPE] = ld[r];
        }
    } else {
        for (int i = tid; i < NV * EPB * NPE; i += TPB) {
            const int v = i / (EPB * NPE), j = i % (EPB * NPE);
            const int le = j / NPE, nd = j % NPE;
            if (e0 + le < a.ne) sQ[le][v][nd] = a.Q[v * Nd + (int64_t)e0 * NPE + j];
        }
    }
    __syncthreads();

    const int le = tid / NPE, node = tid % NPE, e = e0 + le;
    const bool valid = e < a.ne;
    if (valid) {
        T q[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = sQ[le][v][node];
        if constexpr (EQ == EQ_EULER) {
            T P[NV];
            plain::cons2prim<T, ND>(q, a.gamma, P);
#pragma unroll
            for (int v = 0; v < NV; ++v) sP[le][v][node] = P[v];
            sP[le][NV][node] = P[0] / (T(2) * P[ND + 1]);
        }
        if constexpr (VOL == VOL_STRONG) {
            const T* Mp = &sG[le][0];
            T Mn[ND * ND];
            if constexpr (CURV) {
#pragma unroll
                for (int c = 0; c < ND * ND; ++c) Mn[c] = a.ngeom[c * Nd + (int64_t)e * NPE + node];
                Mp = Mn;
            }
#pragma unroll
            for (int dir = 0; dir < ND; ++dir) {
                T Ft[NV];
                plain::contravariant_flux<T, ND, EQ>(q, Mp, dir, a.gamma, avel, Ft);
#pragma unroll
                for (int v = 0; v < NV; ++v) sFt[le][dir * NV + v][node] = Ft[v];
            }
        }
    }

    // Numerical fluxes on the element's faces (interface_fluxes!, applyBCs!)
    for (int it = tid; it < EPB * NF * NPF; it += TPB) {
        const int le2 = it / (NF * NPF), r = it % (NF * NPF);
        const int k = r / NPF, f = r % NPF, e2 = e0 + le2;
        if (e2 >= a.ne) continue;
        const int dir = k >> 1, lr = k & 1;
        T own[NV], ext[NV];
        if (a.gather) {
            const int nn = line_node<ND, N>(dir, f, lr ? N - 1 : 0);
#pragma unroll
            for (int v = 0; v < NV; ++v) own[v] = sQ[le2][v][nn];
        } else {
            const T* l = &sOps[lr ? ops_lp(N) : ops_lm(N)];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                T s = T(0);
                for (int p = 0; p < N; ++p) s = s + l[p] * sQ[le2][v][line_node<ND, N>(dir, f, p)];
                own[v] = s;
            }
        }
        const int slot = e2 * NF + k;
        const bool first = PF && it == tid;  // prefetched before the staging barrier
        const int info = first ? pf_info : a.finfo[slot];
        const int nbe = first ? pf_nbe : a.nb[slot];
        // face geometry per face node (curved) is stored by the master's face-node index
        const T* g = a.fgeom + ((int64_t)(first ? pf_fgi : a.fgi[slot]) * (CURV ? NPF : 1) +
                                (CURV ? face_master_node<ND, N>(f, info) : 0)) *
                                   (3 * ND + 1);
        T gbuf[PF ? 3 * ND + 1 : 1];
        if constexpr (PF) {
#pragma unroll
            for (int c = 0; c < 3 * ND + 1; ++c) gbuf[c] = first ? pf_g[c] : g[c];
            g = gbuf;
        }
        if (nbe >= 0) {
            const int kp = finfo_nbpos(info), ndir = kp >> 1, nlr = kp & 1;
            const int fn = face_nb_node<ND, N>(f, info);  // the neighbour's face node meeting f
            const T* Qn = a.Q + (int64_t)nbe * NPE;
            if (a.gather && first) {
#pragma unroll
                for (int v = 0; v < NV; ++v) ext[v] = pf_ext[v];
            } else if (a.gather) {
                const int nn = line_node<ND, N>(ndir, fn, nlr ? N - 1 : 0);
#pragma unroll
                for (int v = 0; v < NV; ++v) ext[v] = Qn[v * Nd + nn];
            } else {
                const T* l = &sOps[nlr ? ops_lp(N) : ops_lm(N)];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    T s = T(0);
                    for (int p = 0; p < N; ++p) s = s + l[p] * Qn[v * Nd + line_node<ND, N>(ndir, fn, p)];
                    ext[v] = s;
                }
            }
        } else {
            const int bd = finfo_bd(info);
            const int type = a.bctype[bd];
            if (type == BC_DIRICHLET) {
#pragma unroll
                for (int v = 0; v < NV; ++v) ext[v] = a.bcstate[((int64_t)a.fid[slot] * NPF + f) * NV + v];
            } else if (type == BC_INFLOW) {
#pragma unroll
                for (int v = 0; v < NV; ++v) ext[v] = a.bcconst[bd * NV + v];
            } else if (type == BC_OUTFLOW || EQ == EQ_ADVECTION) {
#pragma unroll
                for (int v = 0; v < NV; ++v) ext[v] = own[v];
            } else {
                if constexpr (EQ == EQ_EULER) {
                    T R[NV];
                    plain::rotate2face<T, ND>(own, g, R);
                    R[1] = -R[1];
                    plain::rotate2phys<T, ND>(R, g, ext);
                }
            }
        }
        const bool slave = finfo_side(info);
        T Ql[NV], Qr[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            Ql[v] = slave ? ext[v] : own[v];
            Qr[v] = slave ? own[v] : ext[v];
        }
        T F[NV];
        if constexpr (EQ == EQ_ADVECTION) {
            T an = avel[0] * g[0];
#pragma unroll
            for (int kk = 1; kk < ND; ++kk) an = an + avel[kk] * g[kk];
            F[0] = (an * (Ql[0] + Qr[0])) / T(2);
            if (a.sdiss == DISS_LXF) F[0] = F[0] + ((fabs(an) * (Ql[0] - Qr[0])) / T(2)) * a.intensity;
        } else {
            T Rl[NV], Rr[NV], Fr[NV];
            plain::rotate2face<T, ND>(Ql, g, Rl);
            plain::rotate2face<T, ND>(Qr, g, Rr);
            plain::euler_numericalflux<T, ND>(Rl, Rr, a.savg, a.sdiss, a.typo, a.intensity,
                                              a.gamma, a.errflag, Fr);
            plain::rotate2phys<T, ND>(Fr, g, F);
        }
        const T jac = g[3 * ND];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const T x = F[v] * jac;
            sFn[le2][k][v][f] = slave ? -x : x;
        }
    }
    __syncthreads();

    T acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = T(0);
    if constexpr (VOL == VOL_HYBRID) {
        // Telescopic subcell form (OpDivergence.jl:500-620), direction by direction:
        // F̄[s] on the interface between line positions s-1 and s (F̄[0] = F̄[N] = 0), then
        // dQ[i] += (F̄[i] - F̄[i+1]) / ω[i].
        const T* D = &sOps[ops_D(N)];
        const T* w = &sOps[ops_w(N)];
        for (int dir = 0; dir < ND; ++dir) {
            T gs[3 * ND + 1];
            plain::subgrid_frame<T, ND>(dir, a.sjac[dir], gs);
            for (int it = tid; it < EPB * NPF * (N - 1); it += TPB) {
                const int le2 = it / (NPF * (N - 1)), r = it % (NPF * (N - 1));
                const int f = r / (N - 1), sp = 1 + r % (N - 1);
                if (e0 + le2 >= a.ne) continue;
                const T* M = &sG[le2][0] + ND * dir;
                T Fb[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) Fb[v] = T(0);
                for (int ik = sp; ik < N; ++ik) {
                    const int nk = line_node<ND, N>(dir, f, ik);
                    T qk[NV], Pk[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) { qk[v] = sQ[le2][v][nk]; Pk[v] = sP[le2][v][nk]; }
                    const T bk = sP[le2][NV][nk];
                    for (int il = 0; il < sp; ++il) {
                        const int nl = line_node<ND, N>(dir, f, il);
                        T ql[NV], Pl[NV], Ft[NV];
#pragma unroll
                        for (int v = 0; v < NV; ++v) { ql[v] = sQ[le2][v][nl]; Pl[v] = sP[le2][v][nl]; }
                        const T bl = sP[le2][NV][nl];
                        plain::twopointflux<T, ND, EQ, TP>(ql, qk, Pl, Pk, bl, bk, M, M, a.gamma, avel,
                                                           a.errflag, Ft);
                        const T c = (T(2) * w[il]) * D[il * N + ik];
#pragma unroll
                        for (int v = 0; v < NV; ++v) Fb[v] = Fb[v] + c * Ft[v];
                    }
                }
                // FV flux between positions sp-1 and sp on the subcell-interface frame
                const int nl = line_node<ND, N>(dir, f, sp - 1), nr = line_node<ND, N>(dir, f, sp);
                T Ql[NV], Qr[NV], Rl[NV], Rr[NV], Fr[NV], Fv[NV], Wl[NV], Wr[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) { Ql[v] = sQ[le2][v][nl]; Qr[v] = sQ[le2][v][nr]; }
                plain::rotate2face<T, ND>(Ql, gs, Rl);
                plain::rotate2face<T, ND>(Qr, gs, Rr);
                plain::euler_numericalflux<T, ND>(Rl, Rr, a.savg, a.sdiss, a.typo, a.intensity, a.gamma,
                                                  a.errflag, Fr);
                plain::rotate2phys<T, ND>(Fr, gs, Fv);
#pragma unroll
                for (int v = 0; v < NV; ++v) Fv[v] = Fv[v] * gs[3 * ND];
                // blending on the entropy production of the high-order interface flux
                plain::cons2entropy<T, ND>(Ql, a.gamma, a.errflag, Wl);
                plain::cons2entropy<T, ND>(Qr, a.gamma, a.errflag, Wr);
                T b = (Wr[0] - Wl[0]) * (Fb[0] - Fv[0]);
#pragma unroll
                for (int v = 1; v < NV; ++v) b = b + (Wr[v] - Wl[v]) * (Fb[v] - Fv[v]);
                const T dl = plain::hybrid_delta(b, a.blend, a.errflag);
#pragma unroll
                for (int v = 0; v < NV; ++v) sFb[le2][v][f][sp] = (T(1) - dl) * Fv[v] + dl * Fb[v];
            }
            __syncthreads();
            if (valid) {
                int f, ii;
                node_line<ND, N>(node, dir, f, ii);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const T l = ii == 0 ? T(0) : sFb[le][v][f][ii];
                    const T r = ii == N - 1 ? T(0) : sFb[le][v][f][ii + 1];
                    acc[v] = acc[v] + (l - r) / w[ii];
                }
            }
            __syncthreads();
        }
    }
    // The Gauss split form has block-wide phases after the volume terms: its threads past
    // the last element stay, skipping every per-node step.
    // The split form's pair phases are block-wide too.
    if constexpr (!NB && !PS) {
        if (!valid) return;
    }

    // Volume terms, one thread per node, in the reference's summation order
    const int64_t gnode = (int64_t)e * NPE + node;
    const T* M = &sG[le][0];
    T J = sG[le][ND * ND];
    T Mnode[ND * ND];
    if constexpr (CURV) {
        if (valid) {
#pragma unroll
            for (int c = 0; c < ND * ND; ++c) Mnode[c] = a.ngeom[c * Nd + gnode];
            J = a.ngeom[ND * ND * Nd + gnode];
        }
        M = Mnode;
    }
    if constexpr (PS) {
        // dQ[i] -= D♯[ii, j] F♯[ii, j] for j = 0..N-1 in order (F♯ diagonal = the contravariant
        // flux, pairs with the lower line position first), as the node loop below computes
        // it, but each pair flux is evaluated once per block instead of once per node
        const T* Dsh = &sOps[ops_Dsh(N)];
        for (int dir = 0; dir < ND; ++dir) {
            __syncthreads();  // the previous direction's table has been read
            for (int k = tid; k < EPB * NPF * NPAIR; k += TPB) {
                const int le2 = k / (NPF * NPAIR), r = k - le2 * (NPF * NPAIR);
                const int f2 = r / NPAIR;
                int pa = 0, pb = r - f2 * NPAIR;
                while (pb >= N - 1 - pa) { pb -= N - 1 - pa; ++pa; }
                pb += pa + 1;
                if (e0 + le2 >= a.ne) continue;
                const int na = line_node<ND, N>(dir, f2, pa), nb = line_node<ND, N>(dir, f2, pb);
                T qa[NV], qb[NV], Pa[NV], Pb[NV], F[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    qa[v] = sQ[le2][v][na]; qb[v] = sQ[le2][v][nb];
                    Pa[v] = sP[le2][v][na]; Pb[v] = sP[le2][v][nb];
                }
                const T* Ma = &sG[le2][0] + ND * dir;
                const T* Mb = Ma;
                T Mac[ND], Mbc[ND];
                if constexpr (CURV) {
                    const int64_t ga = (int64_t)(e0 + le2) * NPE + na, gb = (int64_t)(e0 + le2) * NPE + nb;
#pragma unroll
                    for (int c = 0; c < ND; ++c) {
                        Mac[c] = a.ngeom[(c + ND * dir) * Nd + ga];
                        Mbc[c] = a.ngeom[(c + ND * dir) * Nd + gb];
                    }
                    Ma = Mac;
                    Mb = Mbc;
                }
                plain::twopointflux<T, ND, EQ, TP>(qa, qb, Pa, Pb, sP[le2][NV][na], sP[le2][NV][nb], Ma, Mb,
                                                   a.gamma, avel, a.errflag, F);
#pragma unroll
                for (int v = 0; v < NV; ++v) sPair[le2][r][v] = F[v];
            }
            __syncthreads();
            if (valid) {
                int f, ii;
                node_line<ND, N>(node, dir, f, ii);
                T qi[NV], Fd[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) qi[v] = sQ[le][v][node];
                plain::contravariant_flux<T, ND, EQ>(qi, M, dir, a.gamma, avel, Fd);
                for (int j = 0; j < N; ++j) {
                    const T Dij = Dsh[ii * N + j];
                    T F[NV];
                    if (j == ii) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) F[v] = Fd[v];
                    } else {
                        const int lo = j < ii ? j : ii, hi = j < ii ? ii : j;
                        const int r = f * NPAIR + lo * (2 * N - lo - 1) / 2 + (hi - lo - 1);
#pragma unroll
                        for (int v = 0; v < NV; ++v) F[v] = sPair[le][r][v];
                    }
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[v] = acc[v] - Dij * F[v];
                }
            }
        }
        if (!valid) return;
    }
#pragma unroll
    for (int dir = 0; dir < ((VOL == VOL_HYBRID || VOL == VOL_HYBRID_NB || PS) ? 0 : ND) && valid; ++dir) {
        int f, ii;
        node_line<ND, N>(node, dir, f, ii);
        if constexpr (VOL == VOL_STRONG) {
            const T* Ds = &sOps[ops_Ds(N)];
#pragma unroll
            for (int kk = 0; kk < N; ++kk) {
                const int nk = line_node<ND, N>(dir, f, kk);
                const T Dik = Ds[ii * N + kk];
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[v] = acc[v] + Dik * (-sFt[le][dir * NV + v][nk]);
            }
        } else {
            const T* Dsh = &sOps[ops_Dsh(N)];
            T qi[NV], Pi[NV], bi = T(0);
#pragma unroll
            for (int v = 0; v < NV; ++v) qi[v] = sQ[le][v][node];
            if constexpr (EQ == EQ_EULER) {
#pragma unroll
                for (int v = 0; v < NV; ++v) Pi[v] = sP[le][v][node];
                bi = sP[le][NV][node];
            }
            T Fd[NV];
            plain::contravariant_flux<T, ND, EQ>(qi, M, dir, a.gamma, avel, Fd);
            for (int j = 0; j < N; ++j) {
                const int nj = line_node<ND, N>(dir, f, j);
                const T Dij = Dsh[ii * N + j];
                T F[NV];
                if (j == ii) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) F[v] = Fd[v];
                } else {
                    T qj[NV], Pj[NV], bj = T(0);
#pragma unroll
                    for (int v = 0; v < NV; ++v) qj[v] = sQ[le][v][nj];
                    if constexpr (EQ == EQ_EULER) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) Pj[v] = sP[le][v][nj];
                        bj = sP[le][NV][nj];
                    }
                    const T* Mj = M + ND * dir;
                    T Mjc[ND];
                    if constexpr (CURV) {
#pragma unroll
                        for (int k = 0; k < ND; ++k) Mjc[k] = a.ngeom[(k + ND * dir) * Nd + (int64_t)e * NPE + nj];
                        Mj = Mjc;
                    }
                    // F♯ entries are twopointflux(Q_lower position, Q_higher position)
                    if (j > ii)
                        plain::twopointflux<T, ND, EQ, TP>(qi, qj, Pi, Pj, bi, bj, M + ND * dir,
                                                           Mj, a.gamma, avel, a.errflag, F);
                    else
                        plain::twopointflux<T, ND, EQ, TP>(qj, qi, Pj, Pi, bj, bi, Mj,
                                                           M + ND * dir, a.gamma, avel, a.errflag, F);
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[v] = acc[v] - Dij * F[v];
            }
        }
    }

    const T* gm = &sOps[ops_gm(N)];
    const T* gp = &sOps[ops_gp(N)];
    if constexpr (NB) {
        // Entropy-projected surface terms (_splitdiv_nb_surface_contribution!,
        // OpDivergence.jl:300-436), direction by direction: end states u(l·w), u(r·w) of each
        // line; f♯(Q_i, end state) per node; then
        // dQ[i] += ∂gl[i](Fl_i - (l·Fl + Fn_l)) - ∂gr[i](Fr_i - (r·Fr - Fn_r)).
        const T* lm = &sOps[ops_lm(N)];
        const T* lp = &sOps[ops_lp(N)];
        for (int dir = 0; dir < ND; ++dir) {
            for (int it = tid; it < EPB * NPF; it += TPB) {
                const int le2 = it / NPF, f = it % NPF;
                if (e0 + le2 >= a.ne) continue;
                T Wl[NV], Wr[NV];
                for (int k = 0; k < N; ++k) {
                    const int nk = line_node<ND, N>(dir, f, k);
                    T qk[NV], Wk[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) qk[v] = sQ[le2][v][nk];
                    plain::cons2entropy<T, ND>(qk, a.gamma, a.errflag, Wk);
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        Wl[v] = k == 0 ? lm[0] * Wk[v] : Wl[v] + lm[k] * Wk[v];
                        Wr[v] = k == 0 ? lp[0] * Wk[v] : Wr[v] + lp[k] * Wk[v];
                    }
                }
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    T Qe[NV], Pe[NV];
                    plain::entropy2cons<T, ND>(side ? Wr : Wl, a.gamma, a.errflag, Qe);
                    plain::cons2prim<T, ND>(Qe, a.gamma, Pe);
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        sNb[le2][f][side][v] = Qe[v];
                        sNb[le2][f][side][NV + v] = Pe[v];
                    }
                    sNb[le2][f][side][2 * NV] = Pe[0] / (T(2) * Pe[ND + 1]);
                }
            }
            __syncthreads();
            if (valid) {
                int f, ii;
                node_line<ND, N>(node, dir, f, ii);
                T qi[NV], Pi[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) { qi[v] = sQ[le][v][node]; Pi[v] = sP[le][v][node]; }
                const T bi = sP[le][NV][node];
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    const T* E = &sNb[le][f][side][0];
                    T Fs[NV];
                    plain::twopointflux<T, ND, EQ, TP>(qi, E, Pi, E + NV, bi, E[2 * NV], M + ND * dir,
                                                       M + ND * dir, a.gamma, avel, a.errflag, Fs);
#pragma unroll
                    for (int v = 0; v < NV; ++v) sNf[le][side][v][node] = Fs[v];
                }
                if constexpr (HN) {
                    // (D♯[ii,:]·F♯[:,i]) ω[ii]: left fold over the line, F♯ diagonal = the
                    // contravariant flux, pairs with the lower position first
                    const T* Dsh = &sOps[ops_Dsh(N)];
                    const T* w = &sOps[ops_w(N)];
                    T Fd[NV], S[NV];
                    plain::contravariant_flux<T, ND, EQ>(qi, M, dir, a.gamma, avel, Fd);
                    for (int j = 0; j < N; ++j) {
                        const int nj = line_node<ND, N>(dir, f, j);
                        T F[NV];
                        if (j == ii) {
#pragma unroll
                            for (int v = 0; v < NV; ++v) F[v] = Fd[v];
                        } else {
                            T qj[NV], Pj[NV];
#pragma unroll
                            for (int v = 0; v < NV; ++v) { qj[v] = sQ[le][v][nj]; Pj[v] = sP[le][v][nj]; }
                            const T bj = sP[le][NV][nj];
                            if (j > ii)
                                plain::twopointflux<T, ND, EQ, TP>(qi, qj, Pi, Pj, bi, bj, M + ND * dir, M + ND * dir,
                                                                   a.gamma, avel, a.errflag, F);
                            else
                                plain::twopointflux<T, ND, EQ, TP>(qj, qi, Pj, Pi, bj, bi, M + ND * dir, M + ND * dir,
                                                                   a.gamma, avel, a.errflag, F);
                        }
                        const T Dij = Dsh[ii * N + j];
#pragma unroll
                        for (int v = 0; v < NV; ++v) S[v] = j == 0 ? Dij * F[v] : S[v] + Dij * F[v];
                    }
#pragma unroll
                    for (int v = 0; v < NV; ++v) sHs[le][v][node] = S[v] * w[ii];
                }
            }
            __syncthreads();
            if constexpr (HN) {
                hybrid_nb_line_phases<T, ND, N, NV, EPB, TPB>(a, sOps, sQ, sNf, sHs, sFn, dir, e0, le, node, valid,
                                                              tid, acc);
            } else if (valid) {
                int f, ii;
                node_line<ND, N>(node, dir, f, ii);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const int n0 = line_node<ND, N>(dir, f, 0);
                    T lF = lm[0] * sNf[le][0][v][n0], rF = lp[0] * sNf[le][1][v][n0];
                    for (int k = 1; k < N; ++k) {
                        const int nk = line_node<ND, N>(dir, f, k);
                        lF = lF + lm[k] * sNf[le][0][v][nk];
                        rF = rF + lp[k] * sNf[le][1][v][nk];
                    }
                    const T fl = sNf[le][0][v][node] - (lF + sFn[le][2 * dir][v][f]);
                    const T fr = sNf[le][1][v][node] - (rF - sFn[le][2 * dir + 1][v][f]);
                    acc[v] = acc[v] + (gm[ii] * fl - gp[ii] * fr);
                }
            }
            __syncthreads();
        }
        if (!valid) return;
    } else {
        // FR lift: dQ[line[ii]] -= ∂gl[ii]*Fn_left + ∂gr[ii]*Fn_right, direction by direction
#pragma unroll
        for (int dir = 0; dir < ND; ++dir) {
            int f, ii;
            node_line<ND, N>(node, dir, f, ii);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                acc[v] = acc[v] - (gm[ii] * sFn[le][2 * dir][v][f] + gp[ii] * sFn[le][2 * dir + 1][v][f]);
        }
    }

    // Mass matrix (ldiv! by Diagonal(J)) and the stage epilogue
    const int64_t gi = gnode;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const T kv = acc[v] / J;
        const int64_t idx = v * Nd + gi;
        const T uin = sQ[le][v][node];
        switch (a.mode) {
            case MODE_RHS:
                a.out[idx] = kv;
                break;
            case MODE_2N: {
                const T t = a.first ? a.dt * kv : a.A * (PF ? pf_aux[v] : a.reg[idx]) + a.dt * kv;
                a.reg[idx] = t;
                a.out[idx] = uin + a.B * t;
                break;
            }
            case MODE_SSP1:
                a.out[idx] = uin + a.dt * kv;
                break;
            case MODE_SSP2:
                a.out[idx] = ((T(3) * (PF ? pf_aux[v] : a.ubase[idx]) + uin) + a.dt * kv) / T(4);
                break;
            default:
                a.out[idx] = (((PF ? pf_aux[v] : a.ubase[idx]) + T(2) * uin) + (T(2) * a.dt) * kv) / T(3);
                break;
        }
    }
}

template <typename T, int ND, int N, int EQ, int VOL, int TP, bool CURV>
static hipError_t launch_plain(const RhsArgs<T>& a, hipStream_t s) {
    using C = Cfg<ND, N>;
    const int nblocks = (a.ne + C::EPB - 1) / C::EPB;
    hipLaunchKernelGGL((k_plain<T, ND, N, EQ, VOL, TP, CURV>), dim3(nblocks), dim3(C::TPB), 0, s, a);
    return hipGetLastError();
}

template <typename T, int ND, int N, bool CURV>
static hipError_t (*pick_plain(int eq, int vol, int tp))(const RhsArgs<T>&, hipStream_t) {
    if (eq == EQ_ADVECTION) {
        if (vol == VOL_HYBRID || vol == VOL_SPLIT_NB || vol == VOL_HYBRID_NB) return nullptr;
        return vol == VOL_STRONG ? &launch_plain<T, ND, N, EQ_ADVECTION, VOL_STRONG, 0, CURV>
                                 : &launch_plain<T, ND, N, EQ_ADVECTION, VOL_SPLIT, 0, CURV>;
    }
    if (vol == VOL_STRONG) return &launch_plain<T, ND, N, EQ_EULER, VOL_STRONG, 0, CURV>;
    if (vol == VOL_SPLIT_NB) {
        if constexpr (CURV) return nullptr;  // subcell geometry is defined for Cartesian elements
        switch (tp) {
            case TP_STDAVG: return &launch_plain<T, ND, N, EQ_EULER, VOL_SPLIT_NB, TP_STDAVG, false>;
            case TP_CHANDRASEKHAR: return &launch_plain<T, ND, N, EQ_EULER, VOL_SPLIT_NB, TP_CHANDRASEKHAR, false>;
            case TP_KENNEDY_GRUBER: return &launch_plain<T, ND, N, EQ_EULER, VOL_SPLIT_NB, TP_KENNEDY_GRUBER, false>;
            default: return nullptr;
        }
    }
    if (vol == VOL_HYBRID_NB) {
        if constexpr (CURV) return nullptr;  // subcell geometry is defined for Cartesian elements
        switch (tp) {
            case TP_STDAVG: return &launch_plain<T, ND, N, EQ_EULER, VOL_HYBRID_NB, TP_STDAVG, false>;
            case TP_CHANDRASEKHAR: return &launch_plain<T, ND, N, EQ_EULER, VOL_HYBRID_NB, TP_CHANDRASEKHAR, false>;
            case TP_KENNEDY_GRUBER: return &launch_plain<T, ND, N, EQ_EULER, VOL_HYBRID_NB, TP_KENNEDY_GRUBER, false>;
            default: return nullptr;
        }
    }
    if (vol == VOL_HYBRID) {
        if constexpr (CURV) return nullptr;  // subcell geometry is defined for Cartesian elements
        switch (tp) {
            case TP_STDAVG: return &launch_plain<T, ND, N, EQ_EULER, VOL_HYBRID, TP_STDAVG, false>;
            case TP_CHANDRASEKHAR: return &launch_plain<T, ND, N, EQ_EULER, VOL_HYBRID, TP_CHANDRASEKHAR, false>;
            case TP_KENNEDY_GRUBER: return &launch_plain<T, ND, N, EQ_EULER, VOL_HYBRID, TP_KENNEDY_GRUBER, false>;
            default: return nullptr;
        }
    }
    switch (tp) {
        case TP_STDAVG: return &launch_plain<T, ND, N, EQ_EULER, VOL_SPLIT, TP_STDAVG, CURV>;
        case TP_CHANDRASEKHAR: return &launch_plain<T, ND, N, EQ_EULER, VOL_SPLIT, TP_CHANDRASEKHAR, CURV>;
        case TP_KENNEDY_GRUBER: return &launch_plain<T, ND, N, EQ_EULER, VOL_SPLIT, TP_KENNEDY_GRUBER, CURV>;
        default: return nullptr;
    }
}

template <typename T, int ND, bool CURV>
static hipError_t (*pick_plain_n(int n, int eq, int vol, int tp))(const RhsArgs<T>&, hipStream_t) {
    switch (n) {
        case 2: return pick_plain<T, ND, 2, CURV>(eq, vol, tp);
        case 3: return pick_plain<T, ND, 3, CURV>(eq, vol, tp);
        case 4: return pick_plain<T, ND, 4, CURV>(eq, vol, tp);
        case 5: return pick_plain<T, ND, 5, CURV>(eq, vol, tp);
        case 6: return pick_plain<T, ND, 6, CURV>(eq, vol, tp);
        case 7: return pick_plain<T, ND, 7, CURV>(eq, vol, tp);
        case 8: return pick_plain<T, ND, 8, CURV>(eq, vol, tp);
        default: return nullptr;
    }
}

#if defined(FLOUHIP_PREC_F64)
#define FLOUHIP_SCALAR double
#define FLOUHIP_PTAG f64
#elif defined(FLOUHIP_PREC_F32)
#define FLOUHIP_SCALAR float
#define FLOUHIP_PTAG f32
#else
#error "define FLOUHIP_PREC_F64 or FLOUHIP_PREC_F32"
#endif

#define FLOUHIP_CAT3(a, b, c) a##_##b##_##c
#define FLOUHIP_NAME(p, d) FLOUHIP_CAT3(plain, p, d)

#if FLOUHIP_ND == 1
#define FLOUHIP_DTAG 1d
#elif FLOUHIP_ND == 2
#define FLOUHIP_DTAG 2d
#elif FLOUHIP_ND == 3
#define FLOUHIP_DTAG 3d
#else
#error "define FLOUHIP_ND"
#endif

hipError_t (*FLOUHIP_NAME(FLOUHIP_PTAG, FLOUHIP_DTAG)(int n, int eq, int vol, int tp, int curved))(
    const RhsArgs<FLOUHIP_SCALAR>&, hipStream_t) {
#if FLOUHIP_ND > 1
    if (curved) return pick_plain_n<FLOUHIP_SCALAR, FLOUHIP_ND, true>(n, eq, vol, tp);
#else
    if (curved) return nullptr;
#endif
    return pick_plain_n<FLOUHIP_SCALAR, FLOUHIP_ND, false>(n, eq, vol, tp);
}

}  // namespace flouhip
