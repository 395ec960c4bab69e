// Adjoint (vector-Jacobian product) of the rhs, fused with the reverse RK stage epilogue.
//
// Forward (rhs_plain.hip): out_i = (1/J) R_i with
//   R_i = -Σ_d Σ_j D♯_ij F♯_d(Q_i, Q_j)      (split; strong: -Σ_j Ds_ij F̃_d(Q_j))
//         -Σ_d [∂g⁻_ii Fn_{2d}(f) + ∂g⁺_ii Fn_{2d+1}(f)],  Fn = ±jac · F(trace_master, trace_slave)
// (OpDivergence.jl:42-299, Interfaces.jl:25-136, MultielementDiscontinuous.jl:132-137).
// Given a cotangent Y of out, one pass per element block computes
//   V = (∂out/∂Q)^T (c·Y)
// as a gather (no atomics):
//   1. w = c·Y/J for the block's elements (LDS);
//   2. per face node of the element's faces: the cotangent of the face flux,
//      F̄ = -jac (σ_own Σ_i ∂g_i w_own,i + σ_nb Σ_i ∂g_i w_nb,i) (σ = +1 master, -1 slave,
//      the neighbour's w read from HBM), pulled through the rotation and the numerical
//      flux to the element's own trace, and stored in LDS;
//   3. per node: the volume terms (each node gathers ∂/∂Q_i of every pair flux it is part
//      of, with F̄_ij = -D♯_ij w_i - D♯_ji w_j) and the transposed trace projection;
// then the epilogue  o0 = a0·P0 + a1·P1 + V,  o1 = b0·o0 + b1·P1 + b2·P2  (o1 optional),
// which expresses rhs_vjp (Qbar += V) and the reverse stages of SSPRK33 and 2N schemes.
// Compile with -DFLOUHIP_PREC_F64 or -DFLOUHIP_PREC_F32 and -DFLOUHIP_ND=1|2|3.
#include "common.hpp"
#include "physics_plain.hpp"
#include "adjoint.hpp"
#include "dual.hpp"
#include <type_traits>

namespace flouhip {

// LDS slot of b's side of the pair (a, b = a + k mod N) on line f, b at node nb: one slot
// per (k, receiving node), and for even N the k = N/2 pairs (a < N/2) by (f, b - N/2).
template <int N, int NPE>
__device__ inline int pair_slot(int k, int nb, int f, int b) {
    if (N % 2 == 0 && k == N / 2) return (k - 1) * NPE + f * (N / 2) + (b - N / 2);
    return (k - 1) * NPE + nb;
}

// Reverse product of any Euler numerical flux in the face frame (Equations/Euler.jl:
// 99-380: the StdAverage or Chandrasekhar average with LxF, Scalar or Matrix dissipation):
// one forward evaluation of the plain flux in dual numbers per input direction (2 NV),
// Rlb += (∂Φ/∂Rl)^T Fb, Rrb += (∂Φ/∂Rr)^T Fb.  Used where the derivative is not written out
// by hand (adj::numflux_vjp covers StdAverage with LxF or no dissipation).
template <typename T, int ND>
__device__ inline void numflux_vjp_fwd(const T* Rl, const T* Rr, const RhsArgs<T>& a, const T* Fb, T* Rlb,
                                       T* Rrb) {
    using D = adj::Dual1<T>;
    constexpr int NV = ND + 2;
#pragma nounroll
    for (int s = 0; s < 2 * NV; ++s) {
        D L[NV], R[NV], F[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            L[v] = D(Rl[v], s == v ? T(1) : T(0));
            R[v] = D(Rr[v], s == NV + v ? T(1) : T(0));
        }
        plain::euler_numericalflux<D, ND>(L, R, a.savg, a.sdiss, a.typo, D(a.intensity, T(0)), D(a.gamma, T(0)),
                                          a.errflag, F);
        T acc = Fb[0] * F[0].d;
#pragma unroll
        for (int v = 1; v < NV; ++v) acc = acc + Fb[v] * F[v].d;
        if (s < NV) Rlb[s] = Rlb[s] + acc;
        else Rrb[s - NV] = Rrb[s - NV] + acc;
    }
}

// GEN: the face-flux VJP through numflux_vjp_fwd (any Euler surface flux)
template <typename T, int ND, int N, int EQ, int VOL, int TP, bool CURV, bool GEN>
__global__ void __launch_bounds__((Cfg<ND, N>::TPB)) k_vjp(const VjpArgs<T> A) {
    using C = Cfg<ND, N>;
    constexpr int NPE = C::NPE, NPF = C::NPF, NF = C::NF, EPB = C::EPB, TPB = C::TPB;
    constexpr int NV = EQ == EQ_ADVECTION ? 1 : ND + 2;
    constexpr int NOPS = ops_size(N);
    constexpr int NG = ND * ND + 1;
    const RhsArgs<T>& a = A.r;

    __shared__ T sOps[NOPS];
    __shared__ T sG[EPB][NG];
    __shared__ T sQ[EPB][NV][NPE];
    __shared__ T sW[EPB][NV][NPE];
    __shared__ T sTb[EPB][NF][NV][NPF];
    // Split form (Euler): every pair flux of a direction is differentiated once, by the
    // thread of one of its nodes -- node a owns the pairs (a, a+k mod N), k = 1..N/2 (for
    // even N, k = N/2 only when a < N/2) -- which adds a's side of the VJP to its own sum
    // and leaves b's side in sPo (one value per pair and variable, at pair_slot); after a
    // barrier every node adds the sides it received.
    constexpr int NPAIR = N * (N - 1) / 2;
    constexpr bool PS = VOL == VOL_SPLIT && EQ == EQ_EULER && EPB * NV * NPF * NPAIR * (int)sizeof(T) <= 65536;
    // Chandrasekhar: the pair sides are cotangents of the primitives (ρ, u, p), own and
    // received summed over every pair and direction and pulled back to (ρ, m, E) once per
    // node (the pullback is linear), instead of twice per pair
    constexpr bool PP = PS && TP == TP_CHANDRASEKHAR;
    __shared__ T sPo[PS ? EPB : 1][PS ? NV : 1][PS ? NPF * NPAIR : 1];
    // Per-node primitives for the pair phase (adj::node_prims without β, which is rebuilt
    // from ρ and 1/p), computed once per node instead of once per pair and direction, when
    // the extra LDS leaves the number of blocks per CU (160 KB of LDS) unchanged
    constexpr int NR = ND + 4, NRS = ND + 3;
    constexpr int LDS0 = (int)sizeof(T) * (NOPS + EPB * NG + 2 * EPB * NV * NPE + EPB * NF * NV * NPF +
                                           (PS ? EPB * NV * NPF * NPAIR : 1));
    constexpr bool PR = PS && 163840 / (LDS0 + (int)sizeof(T) * EPB * NRS * NPE) == 163840 / LDS0;
    __shared__ T sR[PR ? EPB : 1][PR ? NRS : 1][PR ? NPE : 1];

    const int tid = threadIdx.x;
    const int e0 = xcd_remap(blockIdx.x, gridDim.x) * EPB;
    const int64_t Nd = a.ndofs;
    const T avel[3] = {a.a0, a.a1, a.a2};

    // FP64 split form, Euler, affine, 6 nodes per direction in 3-D (occupancy is set by LDS
    // there and the registers fit without spills or a lower occupancy; in 2-D the registers
    // it takes would lower the occupancy):
    // everything a thread reads from HBM is issued
    // up front -- the element's nodes of Q and Y with their J, the face tables of the face
    // node it takes first in the face phase (it = tid), then (GLL: gather) that face node's
    // neighbour trace, the neighbour line of Y, the neighbour's J and the face frame, and
    // the epilogue operands of its own node -- so that one barrier ends the staging and the
    // face phase waits for nothing.  The arithmetic and its order are unchanged.
    constexpr bool PF = std::is_same<T, double>::value && EQ == EQ_EULER && VOL == VOL_SPLIT && !CURV &&
                        ND == 3 && N == 6;
    constexpr int NFN = EPB * NF * NPF;  // face nodes of the block
    const int pfk = (tid % (NF * NPF)) / NPF, pff = tid % NPF, pfe = e0 + tid / (NF * NPF);
    const bool pf = PF && a.gather && tid < NFN && pfe < a.ne;
    int pf_info = 0, pf_nbe = -1, pf_fgi = 0;
    if (pf) {
        const int slot = pfe * NF + pfk;
        pf_info = a.finfo[slot];
        pf_nbe = a.nb[slot];
        pf_fgi = a.fgi[slot];
    }
    constexpr int NLD = PF ? (NV * EPB * NPE + TPB - 1) / TPB : 1;  // staging loads per thread
    T ldq[NLD], ldy[NLD], ldj[NLD];
    if constexpr (PF) {
#pragma unroll
        for (int r = 0; r < NLD; ++r) {
            const int i = tid + r * TPB;
            const int v = i / (EPB * NPE), j = i % (EPB * NPE);
            const bool in = i < NV * EPB * NPE && e0 + j / NPE < a.ne;
            const int64_t g = v * Nd + (int64_t)e0 * NPE + j;
            ldq[r] = in ? a.Q[g] : T(0);
            ldy[r] = in ? A.Y[g] : T(0);
            ldj[r] = in ? a.egeom[(int64_t)(e0 + j / NPE) * a.egeom_stride + ND * ND] : T(1);
        }
    }
    T pf_ext[NV], pf_yl[N][NV], pf_g[3 * ND + 1], pf_jn = T(1);
    if (pf && pf_nbe >= 0) {
        const int kp = finfo_nbpos(pf_info), ndir = kp >> 1;
        const int fn = face_nb_node<ND, N>(pff, pf_info);
        const int64_t base = (int64_t)pf_nbe * NPE;
        pf_jn = a.egeom[(int64_t)pf_nbe * a.egeom_stride + ND * ND];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            pf_ext[v] = a.Q[v * Nd + base + line_node<ND, N>(ndir, fn, (kp & 1) ? N - 1 : 0)];
#pragma unroll
            for (int p = 0; p < N; ++p) pf_yl[p][v] = A.Y[v * Nd + base + line_node<ND, N>(ndir, fn, p)];
        }
    }
    if (pf) {
        const T* g = a.fgeom + (int64_t)pf_fgi * (3 * ND + 1);
#pragma unroll
        for (int c = 0; c < 3 * ND + 1; ++c) pf_g[c] = g[c];
    }
    T pf_p0[NV], pf_p1[NV], pf_p2[NV];
    if (PF && e0 + tid / NPE < a.ne) {
        const int64_t gi = (int64_t)e0 * NPE + tid;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (A.P0) pf_p0[v] = A.P0[v * Nd + gi];
            if (A.P1) pf_p1[v] = A.P1[v * Nd + gi];
            if (A.P2) pf_p2[v] = A.P2[v * Nd + gi];
        }
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
            if (i < NV * EPB * NPE && e0 + j / NPE < a.ne) {
                sQ[j / NPE][v][j % NPE] = ldq[r];
                sW[j / NPE][v][j % NPE] = A.c * ldy[r] / ldj[r];
            }
        }
    } else {
        __syncthreads();
        for (int i = tid; i < NV * EPB * NPE; i += TPB) {
            const int v = i / (EPB * NPE), j = i % (EPB * NPE);
            const int le = j / NPE, nd = j % NPE;
            if (e0 + le < a.ne) {
                const int64_t g = v * Nd + (int64_t)e0 * NPE + j;
                sQ[le][v][nd] = a.Q[g];
                const T J = CURV ? a.ngeom[ND * ND * Nd + (int64_t)e0 * NPE + j] : sG[le][ND * ND];
                sW[le][v][nd] = A.c * A.Y[g] / J;
            }
        }
    }
    __syncthreads();
    if constexpr (PR) {  // read in the pair phases, after the face phase's barrier
        for (int i = tid; i < EPB * NPE; i += TPB) {
            const int le = i / NPE, nd = i % NPE;
            if (e0 + le >= a.ne) continue;
            T q[NV], r[NR];
#pragma unroll
            for (int v = 0; v < NV; ++v) q[v] = sQ[le][v][nd];
            adj::node_prims<T, ND>(q, a.gamma, r);
#pragma unroll
            for (int c = 0; c < NRS; ++c) sR[le][c][nd] = r[c];
        }
    }

    // ---- face nodes: cotangent of the own trace ----------------------------------------
    for (int it = tid; it < EPB * NF * NPF; it += TPB) {
        const int le = it / (NF * NPF), r = it % (NF * NPF);
        const int k = r / NPF, f = r % NPF, e = e0 + le;
        if (e >= a.ne) continue;
        const int dir = k >> 1, lr = k & 1;
        const T* lk = &sOps[lr ? ops_lp(N) : ops_lm(N)];
        const T* gk = &sOps[lr ? ops_gp(N) : ops_gm(N)];
        T own[NV], ext[NV], Lo[NV];
        const int pend = lr ? N - 1 : 0, nend = line_node<ND, N>(dir, f, pend);
        if (a.gather && a.lift_end) {  // GLL: the trace and the lift live on the end node
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                own[v] = sQ[le][v][nend];
                Lo[v] = gk[pend] * sW[le][v][nend];
            }
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                T s = T(0), t = T(0);
                for (int p = 0; p < N; ++p) {
                    const int nn = line_node<ND, N>(dir, f, p);
                    s = s + lk[p] * sQ[le][v][nn];
                    t = t + gk[p] * sW[le][v][nn];
                }
                own[v] = a.gather ? sQ[le][v][nend] : s;
                Lo[v] = t;
            }
        }
        const int slot = e * NF + k;
        const bool first = pf && it == tid;  // issued before the staging barrier
        const int info = first ? pf_info : a.finfo[slot];
        const int nbe = first ? pf_nbe : a.nb[slot];
        const bool slave = finfo_side(info);
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
        const T jac = g[3 * ND];
        T Fb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) Fb[v] = slave ? Lo[v] : -Lo[v];  // -σ_own L_own
        int type = -1;
        if (nbe >= 0) {
            const int kp = finfo_nbpos(info), ndir = kp >> 1, nlr = kp & 1;
            const T* lq = &sOps[nlr ? ops_lp(N) : ops_lm(N)];
            const T* gq = &sOps[nlr ? ops_gp(N) : ops_gm(N)];
            const T Jn = first ? pf_jn : a.egeom[(int64_t)nbe * a.egeom_stride + ND * ND];
            const int64_t base = (int64_t)nbe * NPE;
            const int fn = face_nb_node<ND, N>(f, info);  // the neighbour's face node meeting f
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                T s = T(0), t = T(0);
                if (first) {
#pragma unroll
                    for (int p = 0; p < N; ++p) t = t + gq[p] * pf_yl[p][v];
                    ext[v] = pf_ext[v];
                } else if (a.gather && a.lift_end) {  // one node of the neighbour's line
                    const int pn = nlr ? N - 1 : 0;
                    const int64_t nn = base + line_node<ND, N>(ndir, fn, pn);
                    const int64_t gi = v * Nd + nn;
                    t = gq[pn] * (CURV ? A.Y[gi] / a.ngeom[ND * ND * Nd + nn] : A.Y[gi]);
                    ext[v] = a.Q[gi];
                } else {
                    for (int p = 0; p < N; ++p) {
                        const int64_t nn = base + line_node<ND, N>(ndir, fn, p);
                        const int64_t gi = v * Nd + nn;
                        s = s + lq[p] * a.Q[gi];
                        t = t + gq[p] * (CURV ? A.Y[gi] / a.ngeom[ND * ND * Nd + nn] : A.Y[gi]);
                    }
                    ext[v] = a.gather ? a.Q[v * Nd + base + line_node<ND, N>(ndir, fn, nlr ? N - 1 : 0)] : s;
                }
                const T Ln = CURV ? A.c * t : A.c * t / Jn;
                Fb[v] = Fb[v] + (slave ? -Ln : Ln);  // -σ_nb L_nb, σ_nb = -σ_own
            }
        } else {
            const int bd = finfo_bd(info);
            type = a.bctype[bd];
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
#pragma unroll
        for (int v = 0; v < NV; ++v) Fb[v] = Fb[v] * jac;
        const T* Ql = slave ? ext : own;
        const T* Qr = slave ? own : ext;
        T Qlb[NV], Qrb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) { Qlb[v] = T(0); Qrb[v] = T(0); }
        if constexpr (EQ == EQ_ADVECTION) {
            T an = avel[0] * g[0];
#pragma unroll
            for (int kk = 1; kk < ND; ++kk) an = an + avel[kk] * g[kk];
            const T diss = a.sdiss == DISS_LXF ? fabs(an) * a.intensity : T(0);
            Qlb[0] = Fb[0] * (an + diss) / T(2);
            Qrb[0] = Fb[0] * (an - diss) / T(2);
        } else {
            T Rl[NV], Rr[NV], Rb[NV], Rlb[NV], Rrb[NV];
            plain::rotate2face<T, ND>(Ql, g, Rl);
            plain::rotate2face<T, ND>(Qr, g, Rr);
            plain::rotate2face<T, ND>(Fb, g, Rb);
#pragma unroll
            for (int v = 0; v < NV; ++v) { Rlb[v] = T(0); Rrb[v] = T(0); }
            if constexpr (GEN) numflux_vjp_fwd<T, ND>(Rl, Rr, a, Rb, Rlb, Rrb);
            else adj::numflux_vjp<T, ND>(Rl, Rr, a.sdiss == DISS_LXF, a.intensity, a.gamma, Rb, Rlb, Rrb);
            plain::rotate2phys<T, ND>(Rlb, g, Qlb);
            plain::rotate2phys<T, ND>(Rrb, g, Qrb);
        }
        T tb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) tb[v] = slave ? Qrb[v] : Qlb[v];
        if (nbe < 0) {  // boundary: the exterior state is a function of the own trace
            if (type == BC_OUTFLOW || (EQ == EQ_ADVECTION && type != BC_DIRICHLET && type != BC_INFLOW)) {
#pragma unroll
                for (int v = 0; v < NV; ++v) tb[v] = tb[v] + Qrb[v];
            } else if (type == BC_SLIP) {
                if constexpr (EQ == EQ_EULER) {
                    T R[NV], B[NV];
                    plain::rotate2face<T, ND>(Qrb, g, R);
                    R[1] = -R[1];
                    plain::rotate2phys<T, ND>(R, g, B);
#pragma unroll
                    for (int v = 0; v < NV; ++v) tb[v] = tb[v] + B[v];
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) sTb[le][k][v][f] = tb[v];
    }
    __syncthreads();

    const int le = tid / NPE, node = tid % NPE, e = e0 + le;
    const bool valid = e < a.ne;
    // the pair phases are block-wide: threads past the last element stay until they end
    if constexpr (!PS) {
        if (!valid) return;
    }

    T qi[NV], wi[NV], V[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        qi[v] = sQ[le][v][node];
        wi[v] = sW[le][v][node];
        V[v] = T(0);
    }
    const T* M = &sG[le][0];
    T Mnode[ND * ND];
    if constexpr (CURV) {
        if (valid) {
#pragma unroll
            for (int c = 0; c < ND * ND; ++c) Mnode[c] = a.ngeom[c * Nd + (int64_t)e * NPE + node];
        }
        M = Mnode;
    }
    T ri[PS ? NR : 1], Vp[PP ? NV : 1];
    if constexpr (PS) {
        if (valid) adj::node_prims<T, ND>(qi, a.gamma, ri);
    }
#pragma unroll
    for (int v = 0; v < (PP ? NV : 1); ++v) Vp[v] = T(0);
#pragma unroll
    for (int dir = 0; dir < ND; ++dir) {
        int f, ii;
        node_line<ND, N>(node, dir, f, ii);
        T n[ND];
#pragma unroll
        for (int k = 0; k < ND; ++k) n[k] = valid ? M[k + ND * dir] : T(0);
        if constexpr (PS) {
            // pair phase: Gb = -D♯_ab w_a - D♯_ba w_b of the owned pairs, pulled back to both
            const T* Dsh = &sOps[ops_Dsh(N)];
            const T ig1 = T(1) / (a.gamma - T(1));
            __syncthreads();  // the previous direction's received sides have been read
            if (valid) {
                for (int k = 1; k <= N / 2; ++k) {
                    if (N % 2 == 0 && k == N / 2 && ii >= N / 2) break;
                    const int jj = ii + k < N ? ii + k : ii + k - N;
                    const int nj = line_node<ND, N>(dir, f, jj);
                    const T dab = Dsh[ii * N + jj], dba = Dsh[jj * N + ii];
                    T qj[NV], rj[NR], Gb[NV], Qbb[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        qj[v] = sQ[le][v][nj];
                        Gb[v] = -dab * wi[v] - dba * sW[le][v][nj];
                        Qbb[v] = T(0);
                    }
                    if constexpr (PR) {
#pragma unroll
                        for (int c = 0; c < NRS; ++c) rj[c] = sR[le][c][nj];
                        rj[ND + 3] = T(0.5) * qj[0] * rj[ND + 2];
                    } else {
                        adj::node_prims<T, ND>(qj, a.gamma, rj);
                    }
                    T nbar[ND];  // averaged metric (Ja_a + Ja_b)/2 of the pair
#pragma unroll
                    for (int c = 0; c < ND; ++c)
                        nbar[c] = CURV ? (n[c] + a.ngeom[(c + ND * dir) * Nd + (int64_t)e * NPE + nj]) / T(2) : n[c];
                    adj::twopoint_vjp2p<T, ND, TP, PP>(qi, ri, qj, rj, nbar, a.gamma, ig1, Gb, PP ? Vp : V, Qbb);
                    const int slot = pair_slot<N, NPE>(k, nj, f, jj);
#pragma unroll
                    for (int v = 0; v < NV; ++v) sPo[le][v][slot] = Qbb[v];
                }
            }
            __syncthreads();
            if (!valid) continue;
            for (int k = 1; k <= N / 2; ++k) {  // the sides received from the owners ii - k
                if (N % 2 == 0 && k == N / 2 && ii < N / 2) break;
                const int slot = pair_slot<N, NPE>(k, node, f, ii);
                T* acc = PP ? Vp : V;
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[v] = acc[v] + sPo[le][v][slot];
            }
            const T d = Dsh[ii * N + ii];
            if (d != T(0)) {  // zero on GLL nodes
                T Gb[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) Gb[v] = -d * wi[v];
                adj::flux_vjp<T, ND>(qi, n, a.gamma, Gb, V);
            }
        } else if constexpr (VOL == VOL_STRONG) {
            const T* Ds = &sOps[ops_Ds(N)];
            T cb[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) cb[v] = T(0);
            for (int kk = 0; kk < N; ++kk) {
                const int nk = line_node<ND, N>(dir, f, kk);
                const T Dki = Ds[kk * N + ii];
#pragma unroll
                for (int v = 0; v < NV; ++v) cb[v] = cb[v] - Dki * sW[le][v][nk];
            }
            if constexpr (EQ == EQ_ADVECTION) {
                T an = T(0);
#pragma unroll
                for (int k = 0; k < ND; ++k) an = an + avel[k] * n[k];
                V[0] = V[0] + an * cb[0];
            } else {
                adj::flux_vjp<T, ND>(qi, n, a.gamma, cb, V);
            }
        } else {
            const T* Dsh = &sOps[ops_Dsh(N)];
            for (int j = 0; j < N; ++j) {
                const int nj = line_node<ND, N>(dir, f, j);
                T Gb[NV];
                if (j == ii) {
                    const T d = Dsh[ii * N + ii];
#pragma unroll
                    for (int v = 0; v < NV; ++v) Gb[v] = -d * wi[v];
                    if constexpr (EQ == EQ_ADVECTION) {
                        T an = T(0);
#pragma unroll
                        for (int k = 0; k < ND; ++k) an = an + avel[k] * n[k];
                        V[0] = V[0] + an * Gb[0];
                    } else {
                        adj::flux_vjp<T, ND>(qi, n, a.gamma, Gb, V);
                    }
                } else {
                    const T dij = Dsh[ii * N + j], dji = Dsh[j * N + ii];
                    T qj[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        qj[v] = sQ[le][v][nj];
                        Gb[v] = -dij * wi[v] - dji * sW[le][v][nj];
                    }
                    T nb[ND];  // averaged metric (Ja_i + Ja_j)/2 of the pair
#pragma unroll
                    for (int k = 0; k < ND; ++k)
                        nb[k] = CURV ? (n[k] + a.ngeom[(k + ND * dir) * Nd + (int64_t)e * NPE + nj]) / T(2) : n[k];
                    if constexpr (EQ == EQ_ADVECTION) {
                        T an = T(0);
#pragma unroll
                        for (int k = 0; k < ND; ++k) an = an + avel[k] * nb[k];
                        V[0] = V[0] + an * Gb[0] / T(2);
                    } else {
                        adj::twopoint_vjp1<T, ND, TP>(qi, qj, nb, a.gamma, Gb, V);
                    }
                }
            }
        }
        // transposed trace projection of the two faces of this line
        const T lmi = sOps[ops_lm(N) + ii], lpi = sOps[ops_lp(N) + ii];
#pragma unroll
        for (int v = 0; v < NV; ++v)
            V[v] = V[v] + lmi * sTb[le][2 * dir][v][f] + lpi * sTb[le][2 * dir + 1][v][f];
    }

    if (!valid) return;
    if constexpr (PP) adj::prim_pullback_u<T, ND>(ri[ND + 1], ri, a.gamma, Vp[0], Vp + 1, Vp[ND + 1], V);
    // epilogue
    const int64_t gi = (int64_t)e * NPE + node;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const int64_t idx = v * Nd + gi;
        T o = V[v];
        if (A.P0) o = o + A.a0 * (PF ? pf_p0[v] : A.P0[idx]);
        if (A.P1) o = o + A.a1 * (PF ? pf_p1[v] : A.P1[idx]);
        if (A.o1) {
            T x = A.b0 * o;
            if (A.P1) x = x + A.b1 * (PF ? pf_p1[v] : A.P1[idx]);
            if (A.P2) x = x + A.b2 * (PF ? pf_p2[v] : A.P2[idx]);
            A.o1[idx] = x;
        }
        A.o0[idx] = o;
    }
}

template <typename T, int ND, int N, int EQ, int VOL, int TP, bool CURV, bool GEN = false>
static hipError_t launch_vjp(const VjpArgs<T>& a, hipStream_t s) {
    using C = Cfg<ND, N>;
    const int nblocks = (a.r.ne + C::EPB - 1) / C::EPB;
    hipLaunchKernelGGL((k_vjp<T, ND, N, EQ, VOL, TP, CURV, GEN>), dim3(nblocks), dim3(C::TPB), 0, s, a);
    return hipGetLastError();
}

template <typename T, int ND, int N, bool CURV>
static hipError_t (*pick_vjp(int eq, int vol, int tp, int gen))(const VjpArgs<T>&, hipStream_t) {
    if (eq == EQ_ADVECTION)
        return vol == VOL_STRONG ? &launch_vjp<T, ND, N, EQ_ADVECTION, VOL_STRONG, 0, CURV>
                                 : &launch_vjp<T, ND, N, EQ_ADVECTION, VOL_SPLIT, 0, CURV>;
    if (gen) {
        if (vol == VOL_STRONG) return &launch_vjp<T, ND, N, EQ_EULER, VOL_STRONG, 0, CURV, true>;
        switch (tp) {
            case TP_STDAVG: return &launch_vjp<T, ND, N, EQ_EULER, VOL_SPLIT, TP_STDAVG, CURV, true>;
            case TP_CHANDRASEKHAR: return &launch_vjp<T, ND, N, EQ_EULER, VOL_SPLIT, TP_CHANDRASEKHAR, CURV, true>;
            case TP_KENNEDY_GRUBER: return &launch_vjp<T, ND, N, EQ_EULER, VOL_SPLIT, TP_KENNEDY_GRUBER, CURV, true>;
            default: return nullptr;
        }
    }
    if (vol == VOL_STRONG) return &launch_vjp<T, ND, N, EQ_EULER, VOL_STRONG, 0, CURV>;
    switch (tp) {
        case TP_STDAVG: return &launch_vjp<T, ND, N, EQ_EULER, VOL_SPLIT, TP_STDAVG, CURV>;
        case TP_CHANDRASEKHAR: return &launch_vjp<T, ND, N, EQ_EULER, VOL_SPLIT, TP_CHANDRASEKHAR, CURV>;
        case TP_KENNEDY_GRUBER: return &launch_vjp<T, ND, N, EQ_EULER, VOL_SPLIT, TP_KENNEDY_GRUBER, CURV>;
        default: return nullptr;
    }
}

template <typename T, int ND, bool CURV>
static hipError_t (*pick_vjp_n(int n, int eq, int vol, int tp, int gen))(const VjpArgs<T>&, hipStream_t) {
    switch (n) {
        case 2: return pick_vjp<T, ND, 2, CURV>(eq, vol, tp, gen);
        case 3: return pick_vjp<T, ND, 3, CURV>(eq, vol, tp, gen);
        case 4: return pick_vjp<T, ND, 4, CURV>(eq, vol, tp, gen);
        case 5: return pick_vjp<T, ND, 5, CURV>(eq, vol, tp, gen);
        case 6: return pick_vjp<T, ND, 6, CURV>(eq, vol, tp, gen);
        case 7: return pick_vjp<T, ND, 7, CURV>(eq, vol, tp, gen);
        case 8: return pick_vjp<T, ND, 8, CURV>(eq, vol, tp, gen);
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
#define FLOUHIP_NAME(p, d) FLOUHIP_CAT3(vjp, p, d)

#if FLOUHIP_ND == 1
#define FLOUHIP_DTAG 1d
#elif FLOUHIP_ND == 2
#define FLOUHIP_DTAG 2d
#elif FLOUHIP_ND == 3
#define FLOUHIP_DTAG 3d
#else
#error "define FLOUHIP_ND"
#endif

hipError_t (*FLOUHIP_NAME(FLOUHIP_PTAG, FLOUHIP_DTAG)(int n, int eq, int vol, int tp, int curved, int gen))(
    const VjpArgs<FLOUHIP_SCALAR>&, hipStream_t) {
#if FLOUHIP_ND > 1
    if (curved) return pick_vjp_n<FLOUHIP_SCALAR, FLOUHIP_ND, true>(n, eq, vol, tp, gen);
#else
    if (curved) return nullptr;
#endif
    return pick_vjp_n<FLOUHIP_SCALAR, FLOUHIP_ND, false>(n, eq, vol, tp, gen);
}

}  // namespace flouhip
