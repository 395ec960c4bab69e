// FP32 ("chosen form") fused element kernel for the 3-D split-form Euler rhs on
// axis-aligned affine hexahedra (Flou's CartesianMesh), GLL nodes, LxF(StdAverage) or
// StdAverage surface flux, with the RK stage epilogue.
//
// Same mathematics as SplitDivOperator + interface_fluxes! + the FR lift
// (OpDivergence.jl:200-299, Interfaces.jl:111-136), organised for gfx950:
//   * EPB elements per block.  Staging: every thread loads all five variables of its
//     nodes (coalesced, one uniform base per variable) and forms the two-point flux
//     inputs in registers; the exterior traces of the block's 6*EPB faces (neighbour
//     nodes or boundary states) are staged in LDS at the same time, so their latency
//     overlaps the element loads;
//   * one thread per (element, direction, line) -- every thread does the same work:
//       - all N(N-1)/2 pairs of its line, each evaluated once (F♯ is symmetric), both rows
//         accumulated in registers (F♯ never touches memory);
//       - the numerical flux on the two face nodes where the line ends and the FR lift of
//         those fluxes into the two end rows (GLL: ∂g lives on the ends);
//     loads and stores of a line are specialised per direction, so node strides and the
//     velocity permutation into the line frame are compile-time LDS offsets;
//   * per node: Σ_d (Ja_dd/J)·(row sum of direction d) and the RK stage epilogue.  The row
//     sums of the three directions go to three LDS regions, two of which reuse the flux
//     inputs and the exterior traces once every line is done.
// FP32 reformulations (the FP64 plain kernel keeps the reference's operation order):
//   - node inputs are pre-halved (ρ/2, u/2) so a mean is one add;
//   - log-mean ρ̂ = (a+b)/(2F), F = atanh(f)/f, f = (a-b)/(a+b), u = f²
//     (Utilities.jl:34-44).  Each line picks a branch-free path from the spread of ρ and
//     β along it: u < 1e-4: F = 1 + u/3, 1/F = 1 - u/3 (truncation < 2e-9), with
//     1/(a+b) ≈ (1/a + 1/b)/4 inside the O(u) correction (relative error u);
//     u < 1e-2: the reference's three-term series; otherwise per pair the series or
//     F = (ln a - ln b)/(2f);
//   - 1/(2β̂) = F_β/(β1+β2) and p̂/ρ̂ = F_ρ/(β1+β2) (algebraically identical);
//   - energy: |ū|² - ½·avg(|u|²) = ½ u1·u2, and the energy row is accumulated halved;
//   - the diagonal metric and 1/J are applied once per node, the lift coefficients
//     ½·∂g/Ja_dd·(surface J) are folded per face;
//   - Rusanov in the element's own frame: F·n_own = ½[s(f_own+f_ext) + λ(Q_own-Q_ext)]
//     with s the own outward normal along the line; the two elements of a face obtain
//     exactly opposite values (the same sums, λ symmetric).
#include "common.hpp"
#include "fast.hpp"

namespace flouhip {
namespace fast3d {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float lg(float x) { return __logf(x); }
__device__ __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); }

constexpr int r4(int x) { return (x + 3) & ~3; }
#ifndef FAST_EPB6
#define FAST_EPB6 2
#endif

template <int N>
struct FCfg {
    static constexpr int NPE = N * N * N;
    static constexpr int NPF = N * N;
    static constexpr int NL = 3 * N * N;                             // line tasks per element
    static constexpr int EPB = N == 4 ? 4 : (N == 5 ? 3 : (N == 6 ? FAST_EPB6 : 1));
    static constexpr int TPB = ((EPB * NL + 63) / 64) * 64;          // threads per block
    static constexpr int NN = EPB * NPE;                             // nodes per block
    static constexpr int NFN = EPB * 6 * NPF;                        // face nodes per block
    static constexpr int NPT = (NN + TPB - 1) / TPB;                 // nodes per thread
    static constexpr int FPT = (NFN + TPB - 1) / TPB;                // face nodes per thread
    // LDS layout in floats.  Q: [5][NN]; P: [6][NN] two-point flux inputs;
    // X: [EPB][6][5][NPF] exterior traces; C: [5][NN]; G: [EPB][4]; F: [EPB][6] float2.
    // Row sums of direction 0 reuse X, direction 1 reuses P, direction 2 uses C.
    // Arrays are indexed by block node, so lanes on consecutive lines hit distinct banks.
    static constexpr int OFF_Q = 0;
    static constexpr int OFF_P = r4(5 * NN);
    static constexpr int OFF_X = r4(OFF_P + 6 * NN);
    static constexpr int SZ_X = 30 * EPB * NPF > 5 * NN ? 30 * EPB * NPF : 5 * NN;
    static constexpr int OFF_C = r4(OFF_X + SZ_X);
    static constexpr int OFF_G = r4(OFF_C + 5 * NN);
    static constexpr int OFF_F = OFF_G + 4 * EPB;
    static constexpr int LDS = OFF_F + 12 * EPB;
};

// Node inputs (6 slots):
//   Chandrasekhar:  ρ/2, u/2, v/2, w/2 | β = ρ/(2p), 2/ρ
//   Kennedy-Gruber: ρ/2, u/2, v/2, w/2 | p/2, e/2
//   StdAverage:     ρ,   u,   v,   w   | p,   ρe
// In a line task slot 1 holds the velocity along the line, slots 2, 3 the other two
// (cyclic order).  Chandrasekhar returns the flux factored as (m, F_1, u_t1, u_t2, h/2):
// F = (m, F_1, m·u_t1, m·u_t2, m·h/2) with the energy component halved.
template <int TP, int PATH>
__device__ __forceinline__ void twopoint(const float* Pi, const float* Pj, float inv_gm1, float* F) {
    if constexpr (TP == TP_CHANDRASEKHAR) {
        const float hs = Pi[0] + Pj[0], dA = Pi[0] - Pj[0];
        const float sb = Pi[4] + Pj[4], db = Pi[4] - Pj[4];
        const float rb = rcp(sb);
        const float fb = db * rb;
        const float ub = fb * fb;
        float rho, X;  // X = (F_β/(γ-1) + F_ρ)/2
        if constexpr (PATH == 1) {
            const float fr4 = dA * (Pi[5] + Pj[5]);  // ≈ 4f
            const float ur16 = fr4 * fr4;
            X = fmaf(ub, inv_gm1 * (1.0f / 6.0f), fmaf(ur16, 1.0f / 96.0f, 0.5f * (inv_gm1 + 1.0f)));
            rho = fmaf(dA * fr4, -1.0f / 12.0f, hs);
        } else {
            const float fr = dA * rcp(hs);
            const float ur = fr * fr;
            float Fr, Fb;
            if constexpr (PATH == 3) {
                Fr = fmaf(ur, fmaf(ur, fmaf(ur, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
                Fb = fmaf(ub, fmaf(ub, fmaf(ub, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
            } else {
                if (ur < 0.01f) Fr = fmaf(ur, fmaf(ur, fmaf(ur, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
                else Fr = 0.5f * (lg(Pi[0]) - lg(Pj[0])) * rcp(fr);
                if (ub < 0.01f) Fb = fmaf(ub, fmaf(ub, fmaf(ub, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
                else Fb = 0.5f * (lg(Pi[4]) - lg(Pj[4])) * rcp(fb);
            }
            X = fmaf(Fb, 0.5f * inv_gm1, 0.5f * Fr);
            rho = hs * rcp(Fr);
        }
        const float p = hs * rb;
        const float un = Pi[1] + Pj[1];
        const float ut1 = Pi[2] + Pj[2];
        const float ut2 = Pi[3] + Pj[3];
        const float dot = fmaf(Pi[1], Pj[1], fmaf(Pi[2], Pj[2], Pi[3] * Pj[3]));
        const float m = rho * un;
        // factored: F = (m, F[1], m·F[2], m·F[3], m·F[4])
        F[0] = m;
        F[1] = fmaf(m, un, p);
        F[2] = ut1;
        F[3] = ut2;
        F[4] = fmaf(rb, X, dot);
    } else if constexpr (TP == TP_KENNEDY_GRUBER) {
        const float rho = Pi[0] + Pj[0];
        const float un = Pi[1] + Pj[1];
        const float ut1 = Pi[2] + Pj[2];
        const float ut2 = Pi[3] + Pj[3];
        const float p = Pi[4] + Pj[4];
        const float e = Pi[5] + Pj[5];
        const float m = rho * un;
        F[0] = m;
        F[1] = fmaf(m, un, p);
        F[2] = m * ut1;
        F[3] = m * ut2;
        F[4] = fmaf(m, e, p * un);
    } else {
        const float mi = Pi[0] * Pi[1], mj = Pj[0] * Pj[1];
        F[0] = 0.5f * (mi + mj);
        F[1] = 0.5f * (fmaf(mi, Pi[1], Pi[4]) + fmaf(mj, Pj[1], Pj[4]));
        F[2] = 0.5f * (mi * Pi[2] + mj * Pj[2]);
        F[3] = 0.5f * (mi * Pi[3] + mj * Pj[3]);
        F[4] = 0.5f * ((Pi[5] + Pi[4]) * Pi[1] + (Pj[5] + Pj[4]) * Pj[1]);
    }
}

// Rows i and j of the line accumulate -D♯_ij F and -D♯_ji F.  The pairs (0, j) carry the
// first contribution to every row, so they set the accumulators.
template <int N, int TP, int PATH>
__device__ __forceinline__ void line_pairs(const float (&P)[N][6], float (&acc)[N][5],
                                           const float* __restrict__ Dsh, float inv_gm1) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = i + 1; j < N; ++j) {
            float F[5];
            twopoint<TP, PATH>(P[i], P[j], inv_gm1, F);
            const float cij = Dsh[i * N + j], cji = Dsh[j * N + i];
            if constexpr (TP == TP_CHANDRASEKHAR) {
                const float mi = cij * F[0], mj = cji * F[0];
                if (i == 0) {
                    acc[i][0] = (j == 1) ? -mi : acc[i][0] - mi;
                    acc[j][0] = -mj;
                    acc[i][1] = (j == 1) ? -cij * F[1] : fmaf(-cij, F[1], acc[i][1]);
                    acc[j][1] = -cji * F[1];
#pragma unroll
                    for (int v = 2; v < 5; ++v) {
                        acc[i][v] = (j == 1) ? -mi * F[v] : fmaf(-mi, F[v], acc[i][v]);
                        acc[j][v] = -mj * F[v];
                    }
                } else {
                    acc[i][0] -= mi;
                    acc[j][0] -= mj;
                    acc[i][1] = fmaf(-cij, F[1], acc[i][1]);
                    acc[j][1] = fmaf(-cji, F[1], acc[j][1]);
#pragma unroll
                    for (int v = 2; v < 5; ++v) {
                        acc[i][v] = fmaf(-mi, F[v], acc[i][v]);
                        acc[j][v] = fmaf(-mj, F[v], acc[j][v]);
                    }
                }
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    if (i == 0) {
                        acc[i][v] = (j == 1) ? -cij * F[v] : fmaf(-cij, F[v], acc[i][v]);
                        acc[j][v] = -cji * F[v];
                    } else {
                        acc[i][v] = fmaf(-cij, F[v], acc[i][v]);
                        acc[j][v] = fmaf(-cji, F[v], acc[j][v]);
                    }
                }
            }
        }
    }
}

// Line addressing.  Node p of a line is block node base + p*stride; the momentum along the
// line and the two others (line-frame slots 1, 2, 3) are physical variables m[0..2] --
// per-lane offsets, so one instruction stream serves lines of every direction.
struct LineAddr {
    int base, stride;
    int m[3];
};

// flux inputs of the N nodes of a line, line-frame order (slot arrays of NN floats)
template <int N, int NN>
__device__ __forceinline__ void load_line(const float* __restrict__ sP, const LineAddr& L, float (&P)[N][6]) {
#pragma unroll
    for (int p = 0; p < N; ++p) {
        const float* q = sP + L.base + p * L.stride;
        P[p][0] = q[0];
        P[p][1] = q[L.m[0] * NN];
        P[p][2] = q[L.m[1] * NN];
        P[p][3] = q[L.m[2] * NN];
        P[p][4] = q[4 * NN];
        P[p][5] = q[5 * NN];
    }
}

// five variables in line-frame order from an array with variable stride VS
template <int VS>
__device__ __forceinline__ void load_vars(const float* __restrict__ s, const LineAddr& L, float (&o)[5]) {
    o[0] = s[0];
    o[1] = s[L.m[0] * VS];
    o[2] = s[L.m[1] * VS];
    o[3] = s[L.m[2] * VS];
    o[4] = s[4 * VS];
}

template <int N, int NN>
__device__ __forceinline__ void store_line(float* __restrict__ R, const LineAddr& L, const float (&acc)[N][5]) {
#pragma unroll
    for (int p = 0; p < N; ++p) {
        float* r = R + L.base + p * L.stride;
        r[0] = acc[p][0];
        r[L.m[0] * NN] = acc[p][1];
        r[L.m[1] * NN] = acc[p][2];
        r[L.m[2] * NN] = acc[p][3];
        r[4 * NN] = acc[p][4];
    }
}

// Numerical flux on the face where a line ends (row R of the line) and its lift:
// acc[R] -= coef·[s(f_own + f_ext) + λ(Q_own - Q_ext)], all in the line frame
// (slot 1 = momentum along the line).
template <int N, int R>
__device__ __forceinline__ void face_lift(const float (&o)[5], const float (&x)[5], float s, float coef,
                                          float gm1, float gamma, float diss, float (&acc)[N][5],
                                          bool& bad) {
    float fo[5], fx[5], ao, ax;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const float* Q = side ? x : o;
        float* f = side ? fx : fo;
        const float ir = rcp(Q[0]);
        const float m2 = fmaf(Q[1], Q[1], fmaf(Q[2], Q[2], Q[3] * Q[3]));
        const float p = gm1 * fmaf(-0.5f * m2, ir, Q[4]);
        const float un = Q[1] * ir;
        const float c2 = gamma * p * ir;
        bad |= !(c2 >= 0.0f);
        const float a = fabsf(un) + sq(c2);
        f[0] = Q[1];
        f[1] = fmaf(Q[1], un, p);
        f[2] = Q[2] * un;
        f[3] = Q[3] * un;
        f[4] = (Q[4] + p) * un;
        if (side) ax = a; else ao = a;
    }
    const float lam = fmaxf(ao, ax) * diss;
#pragma unroll
    for (int v = 0; v < 5; ++v)
        acc[R][v] = fmaf(-coef, fmaf(s, fo[v] + fx[v], lam * (o[v] - x[v])), acc[R][v]);
}

// Buffer resources: a uniform base in SGPRs, 32-bit lane byte offsets, and the
// hardware range check (loads past num_records return 0, stores are dropped).  The fast
// path is only enabled for states of fewer than 2^30 nodes, so every per-variable byte
// offset fits in 32 bits.  Word 3 = 0x00020000: 32-bit data format (gfx9 family).
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const float* base, uint32_t nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, (int)nbytes, 0x00020000);
}
__device__ __forceinline__ float bload(rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bstore(rsrc_t r, uint32_t off, float x) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, (int)off, 0, 0);
}

template <int N, int TP>
__global__ void __launch_bounds__((FCfg<N>::TPB)) k_fast3d(const FastArgs a) {
    using C = FCfg<N>;
    constexpr int NPE = C::NPE, NPF = C::NPF, NL = C::NL, EPB = C::EPB, TPB = C::TPB;
    constexpr int NN = C::NN, NFN = C::NFN, NPT = C::NPT, FPT = C::FPT;
    __shared__ float4 smem[(C::LDS + 3) / 4];
    float* const sm = reinterpret_cast<float*>(smem);
    float* const sQ = sm + C::OFF_Q;
    float* const sP = sm + C::OFF_P;
    float* const sX = sm + C::OFF_X;
    float* const sC = sm + C::OFF_C;
    float* const sG = sm + C::OFF_G;
    float2* const sF = reinterpret_cast<float2*>(sm + C::OFF_F);

    const int tid = threadIdx.x;
    const int e0 = xcd_remap(blockIdx.x, gridDim.x) * EPB;
    const int64_t Nd = a.ndofs;
    const float gamma = a.gamma, gm1 = a.gamma - 1.0f, inv_gm1 = a.inv_gm1;
    const int nel = a.ne - e0 < EPB ? a.ne - e0 : EPB;
    const uint32_t n0 = (uint32_t)e0 * NPE;  // first node of the block
    int err = 0;                              // domain-error bits, reported once at the end

    // ---- staging ---------------------------------------------------------------------
    // face tables of the face nodes this thread stages (first: they feed addresses)
    int fnb[FPT], finf[FPT];
#pragma unroll
    for (int r = 0; r < FPT; ++r) {
        const int fn = tid + r * TPB;
        const int le = fn / (6 * NPF), k = (fn - le * 6 * NPF) / NPF;
        fnb[r] = -2;
        finf[r] = 0;
        if (fn < NFN && le < nel) {
            fnb[r] = a.nb[(e0 + le) * 6 + k];
            finf[r] = a.finfo[(e0 + le) * 6 + k];
        }
    }
    // element nodes, all five variables per node (lanes past the block's last element read 0)
    const uint32_t blk_bytes = (uint32_t)(nel * NPE) * 4u;
    float q[NPT][5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const rsrc_t rq = make_rsrc(a.Q + v * Nd + n0, blk_bytes);
#pragma unroll
        for (int r = 0; r < NPT; ++r) q[r][v] = bload(rq, (uint32_t)(tid + r * TPB) * 4u);
    }
    // exterior traces: neighbour node (interior faces, branch-free; other lanes read 0),
    // then boundary states or the own node (outflow, slip) where the face is a boundary
    float x[FPT][5];
    bool slip[FPT];
    uint32_t xoff[FPT];
#pragma unroll
    for (int r = 0; r < FPT; ++r) {
        const int f = ((tid + r * TPB) % (6 * NPF)) % NPF;
        const int kp = finfo_nbpos(finf[r]);
        xoff[r] = fnb[r] >= 0 ? ((uint32_t)fnb[r] * NPE + line_node<3, N>(kp >> 1, f, (kp & 1) ? N - 1 : 0)) * 4u
                              : 0xffffffffu;
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const rsrc_t rv = make_rsrc(a.Q + v * Nd, (uint32_t)Nd * 4u);
#pragma unroll
        for (int r = 0; r < FPT; ++r) x[r][v] = bload(rv, xoff[r]);
    }
#pragma unroll
    for (int r = 0; r < FPT; ++r) {
        slip[r] = false;
        if (fnb[r] == -1) {
            const int fn = tid + r * TPB;
            const int le = fn / (6 * NPF), rem = fn - le * 6 * NPF, k = rem / NPF, f = rem - k * NPF;
            const int bd = finfo_bd(finf[r]);
            const int type = a.bctype[bd];
            const float* src;
            int64_t vs = 1;
            if (type == BC_DIRICHLET) {
                src = a.bcstate + ((int64_t)a.fid[(e0 + le) * 6 + k] * NPF + f) * 5;
            } else if (type == BC_INFLOW) {
                src = a.bcconst + bd * 5;
            } else {
                src = a.Q + n0 + le * NPE + line_node<3, N>(k >> 1, f, (k & 1) ? N - 1 : 0);
                vs = Nd;
                slip[r] = type == BC_SLIP;
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) x[r][v] = src[v * vs];
        }
    }
    // per face: own outward normal sign along the axis and the lift coefficient
    if (tid < 6 * EPB) {
        const int le = tid / 6, k = tid - (tid / 6) * 6;
        if (le < nel) {
            const int e = e0 + le;
            const float* eg = a.egeom + (int64_t)e * a.egeom_stride;
            const int gi = a.fgi[e * 6 + k];
            const int side = finfo_side(a.finfo[e * 6 + k]);
            const float sn = a.fgeom[(int64_t)gi * 10 + (k >> 1)];
            const float jac = a.fgeom[(int64_t)gi * 10 + 9];
            const float lift = (k & 1) ? a.lift_p : a.lift_m;
            sF[tid] = make_float2(side ? -sn : sn, 0.5f * eg[4 + (k >> 1)] * lift * jac);
            if (k < 3) sG[le * 4 + k] = eg[k];
        }
    }
    // flux inputs and LDS writes
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int j = tid + r * TPB;
        if (j >= NN) continue;
        const float rho = q[r][0];
        const float ir = rcp(rho);
        const float u = q[r][1] * ir, v = q[r][2] * ir, w = q[r][3] * ir;
        const float qq = fmaf(u, u, fmaf(v, v, w * w));
        const float p = gm1 * fmaf(-0.5f * rho, qq, q[r][4]);
        if (j < nel * NPE && (!(rho > 0.0f) || !(p > 0.0f))) err |= ERR_LOG_NEG;
#pragma unroll
        for (int c = 0; c < 5; ++c) sQ[c * NN + j] = q[r][c];
        float in[6];
        if constexpr (TP == TP_CHANDRASEKHAR) {
            in[0] = 0.5f * rho; in[1] = 0.5f * u; in[2] = 0.5f * v; in[3] = 0.5f * w;
            in[4] = 0.5f * rho * rcp(p); in[5] = 2.0f * ir;
        } else if constexpr (TP == TP_KENNEDY_GRUBER) {
            in[0] = 0.5f * rho; in[1] = 0.5f * u; in[2] = 0.5f * v; in[3] = 0.5f * w;
            in[4] = 0.5f * p; in[5] = 0.5f * q[r][4] * ir;
        } else {
            in[0] = rho; in[1] = u; in[2] = v; in[3] = w; in[4] = p; in[5] = q[r][4];
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) sP[c * NN + j] = in[c];
    }
#pragma unroll
    for (int r = 0; r < FPT; ++r) {
        const int fn = tid + r * TPB;
        if (fn >= NFN) continue;
        const int le = fn / (6 * NPF), rem = fn - le * 6 * NPF, k = rem / NPF, f = rem - k * NPF;
        const int d = k >> 1;
        float* dst = sX + (le * 6 + k) * 5 * NPF + f;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const bool flip = slip[r] && c == 1 + d;  // slip wall: reflect the normal momentum
            dst[c * NPF] = flip ? -x[r][c] : x[r][c];
        }
    }
    __syncthreads();

    // ---- line tasks ------------------------------------------------------------------
    // Lanes past the block's line tasks repeat a valid line (their results are dropped),
    // so the whole wave runs one instruction stream.
    const int tl = tid < EPB * NL ? tid : tid - EPB * NL;
    const int lle = tl / NL, lt = tl - (tl / NL) * NL;
    const int ldir = lt / NPF, lf = lt - (lt / NPF) * NPF;
    const bool line_ok = tid < EPB * NL && lle < nel;
    LineAddr L;
    L.base = lle * NPE + (ldir == 0 ? N * lf : (ldir == 1 ? lf % N + NPF * (lf / N) : lf));
    L.stride = ldir == 0 ? 1 : (ldir == 1 ? N : NPF);
    L.m[0] = 1 + ldir;
    L.m[1] = ldir == 2 ? 1 : 2 + ldir;
    L.m[2] = ldir == 0 ? 3 : ldir;
    float acc[N][5];  // set by the first pair of each row
    {
        float P[N][6];
        load_line<N, NN>(sP, L, P);
        if constexpr (TP == TP_CHANDRASEKHAR) {
            float amax = P[0][0], amin = P[0][0], bmax = P[0][4], bmin = P[0][4];
#pragma unroll
            for (int p = 1; p < N; ++p) {
                amax = fmaxf(amax, P[p][0]); amin = fminf(amin, P[p][0]);
                bmax = fmaxf(bmax, P[p][4]); bmin = fminf(bmin, P[p][4]);
            }
            const float fa = (amax - amin) * rcp(amax + amin);
            const float fb = (bmax - bmin) * rcp(bmax + bmin);
            const float um = fmaxf(fa * fa, fb * fb) * 1.0001f;  // margin for rcp rounding
            if (um < 1e-4f) line_pairs<N, TP, 1>(P, acc, a.dsharp, inv_gm1);
            else if (um < 1e-2f) line_pairs<N, TP, 3>(P, acc, a.dsharp, inv_gm1);
            else line_pairs<N, TP, 0>(P, acc, a.dsharp, inv_gm1);
#pragma unroll
            for (int p = 0; p < N; ++p) acc[p][4] += acc[p][4];  // energy row was halved
        } else {
            line_pairs<N, TP, 0>(P, acc, a.dsharp, inv_gm1);
        }
        // face fluxes at the two line ends and their lift
        float o0[5], x0[5], o1[5], x1[5];
        load_vars<NN>(sQ + L.base, L, o0);
        load_vars<NN>(sQ + L.base + (N - 1) * L.stride, L, o1);
        const float* xe = sX + (lle * 6 + 2 * ldir) * 5 * NPF + lf;
        load_vars<NPF>(xe, L, x0);
        load_vars<NPF>(xe + 5 * NPF, L, x1);
        const float2 g0 = sF[lle * 6 + 2 * ldir], g1 = sF[lle * 6 + 2 * ldir + 1];
        bool bad = false;
        face_lift<N, 0>(o0, x0, g0.x, g0.y, gm1, gamma, a.diss, acc, bad);
        face_lift<N, N - 1>(o1, x1, g1.x, g1.y, gm1, gamma, a.diss, acc, bad);
        if (bad && line_ok) err |= ERR_SQRT_NEG;
    }

    // the stage epilogue's second operand (RK register or SSP base), loaded before the
    // barrier so its latency overlaps the other waves' line tasks
    const float* aux = (a.mode == MODE_2N && !a.first) ? a.reg
                     : ((a.mode == MODE_SSP2 || a.mode == MODE_SSP3) ? a.ubase : nullptr);
    float xv[NPT][5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const rsrc_t rx = make_rsrc(aux ? aux + v * Nd + n0 : a.Q, aux ? blk_bytes : 0u);
#pragma unroll
        for (int r = 0; r < NPT; ++r) xv[r][v] = bload(rx, (uint32_t)(tid + r * TPB) * 4u);
    }
    __syncthreads();  // flux inputs and exterior traces are dead from here
    if (line_ok) store_line<N, NN>(ldir == 0 ? sX : (ldir == 1 ? sP : sC), L, acc);
    __syncthreads();

    // ---- per node: Σ_d (Ja_dd/J) row sums and the stage epilogue -----------------------
    // (stores of lanes past the block's last element are dropped by the range check)
    float kv[NPT][5];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int j = tid + r * TPB;
        const int jj = j < NN ? j : 0;
        const int le = jj / NPE;
        const float s0 = sG[le * 4 + 0], s1 = sG[le * 4 + 1], s2 = sG[le * 4 + 2];
#pragma unroll
        for (int v = 0; v < 5; ++v)
            kv[r][v] = fmaf(s0, sX[v * NN + jj], fmaf(s1, sP[v * NN + jj], s2 * sC[v * NN + jj]));
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const rsrc_t ro = make_rsrc(a.out + v * Nd + n0, blk_bytes);
        const rsrc_t rr = make_rsrc(a.reg + v * Nd + n0, a.mode == MODE_2N ? blk_bytes : 0u);
#pragma unroll
        for (int r = 0; r < NPT; ++r) {
            const int j = tid + r * TPB;
            const uint32_t off = (uint32_t)j * 4u;
            const float uin = sQ[v * NN + (j < NN ? j : 0)];
            const float k = kv[r][v], xx = xv[r][v];
            float out;
            if (a.mode == MODE_RHS) {
                out = k;
            } else if (a.mode == MODE_2N) {
                const float t = a.first ? a.dt * k : fmaf(a.A, xx, a.dt * k);
                bstore(rr, off, t);
                out = fmaf(a.B, t, uin);
            } else if (a.mode == MODE_SSP1) {
                out = fmaf(a.dt, k, uin);
            } else if (a.mode == MODE_SSP2) {
                out = 0.25f * (3.0f * xx + uin + a.dt * k);
            } else {
                out = (1.0f / 3.0f) * (xx + 2.0f * uin + 2.0f * a.dt * k);
            }
            bstore(ro, off, out);
        }
    }
    if (err) atomicOr(a.errflag, err);
}

template <int N, int TP>
static hipError_t launch(const FastArgs& a, hipStream_t s) {
    const int nblocks = (a.ne + FCfg<N>::EPB - 1) / FCfg<N>::EPB;
    hipLaunchKernelGGL((k_fast3d<N, TP>), dim3(nblocks), dim3(FCfg<N>::TPB), 0, s, a);
    return hipGetLastError();
}

template <int N>
static FastLauncher pick_tp(int tp) {
    switch (tp) {
        case TP_CHANDRASEKHAR: return &launch<N, TP_CHANDRASEKHAR>;
        case TP_KENNEDY_GRUBER: return &launch<N, TP_KENNEDY_GRUBER>;
        case TP_STDAVG: return &launch<N, TP_STDAVG>;
        default: return nullptr;
    }
}

}  // namespace fast3d

FastLauncher fast3d_launcher(int n, int tp) {
    switch (n) {
        case 4: return fast3d::pick_tp<4>(tp);
        case 5: return fast3d::pick_tp<5>(tp);
        case 6: return fast3d::pick_tp<6>(tp);
        case 7: return fast3d::pick_tp<7>(tp);
        case 8: return fast3d::pick_tp<8>(tp);
        default: return nullptr;
    }
}

}  // namespace flouhip
