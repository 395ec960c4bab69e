// FP32 ("chosen form") fused element kernel for the 3-D split-form Euler rhs on
// axis-aligned affine hexahedra (Flou's CartesianMesh), GLL nodes, LxF(StdAverage) or
// StdAverage surface flux, with the RK stage epilogue.
//
// Same mathematics as SplitDivOperator + interface_fluxes! + the FR lift
// (OpDivergence.jl:200-299, Interfaces.jl:111-136), organised for gfx950:
//   * EPB elements per block, one thread per (element, direction, line) -- every thread
//     does the same work.  Each thread first loads the face tables and then the two
//     exterior traces of its line (neighbour nodes or boundary states, into registers),
//     issued after the block's own nodes so that the staging waits for those only; a
//     neighbour inside the block is read from LDS after staging instead;
//   * staging: every thread loads all five variables of its nodes (coalesced, one
//     uniform base per variable) and writes the node states and the two-point flux
//     inputs to LDS;
//   * line phase: the numerical flux on the two face nodes where the line ends, lifted
//     into the two end rows (GLL: ∂g lives on the ends), then all N(N-1)/2 pairs of the
//     line, each evaluated once (F♯ is symmetric), both rows accumulated in registers
//     (F♯ never touches memory).  Line-frame velocity slots are consecutive in LDS
//     (u, v, w, u, v), so one per-lane offset serves every direction;
//   * per node: Σ_d (Ja_dd/J)·(row sum of direction d) and the RK stage epilogue.  The row
//     sums go to LDS in line-frame order over the dead flux inputs and node states.
// FP32 reformulations (the FP64 plain kernel keeps the reference's operation order):
//   - node inputs are pre-halved (ρ/2, u/2) so a mean is one add;
//   - log-mean ρ̂ = (a+b)/(2F), F = atanh(f)/f, f = (a-b)/(a+b), u = f²
//     (Utilities.jl:34-44).  Each line picks a branch-free path from the spread of ρ and
//     β along it: u < 1e-4: F = 1 + u/3, 1/F = 1 - u/3 (truncation < 2e-9) with no
//     reciprocal per pair (precomputed per-node inputs, see twopoint); u < 1e-2: the
//     reference's three-term series; otherwise per pair the series or
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
#ifndef FAST_WAVES
#define FAST_WAVES
#endif

template <int N>
struct FCfg {
    static constexpr int NPE = N * N * N;
    static constexpr int NPF = N * N;
    static constexpr int NL = 3 * N * N;                             // line tasks per element
    static constexpr int EPB = N == 4 ? 4 : (N == 5 ? 3 : (N == 6 ? FAST_EPB6 : 1));
    static constexpr int TPB = ((EPB * NL + 63) / 64) * 64;          // threads per block
    static constexpr int NN = EPB * NPE;                             // nodes per block
    static constexpr int NPT = (NN + TPB - 1) / TPB;                 // nodes per thread
    // LDS layout in floats.  Q: [5][NN] node states; P: [10][NN] two-point flux inputs
    // (slots 1..5 hold u, v, w, u, v, so the three velocity slots of a line in any
    // direction are consecutive; slot 8 is an extra Chandrasekhar input; slot 9 is spare
    // room for the row sums); G: [EPB][4]; F: [EPB][6] float2.  Once every line is done
    // the row sums (line-frame variable order) of directions 0 and 1 go to P, those of
    // direction 2 to Q.  Arrays are indexed by block node.
    static constexpr int OFF_Q = 0;
    static constexpr int OFF_P = r4(5 * NN);
    static constexpr int OFF_G = r4(OFF_P + 10 * NN);
    static constexpr int OFF_F = OFF_G + 4 * EPB;
    static constexpr int LDS = OFF_F + 12 * EPB;
};

// Node inputs (6 values, stored in 8 LDS slots with u, v repeated after w; a 7th,
// H = 1/(4β) = p/(2ρ), for the Chandrasekhar flux):
//   Chandrasekhar:  ρ/2, u/2, v/2, w/2 | β = ρ/(2p), -(1/6)/ρ | H
//   Kennedy-Gruber: ρ/2, u/2, v/2, w/2 | p/2, e/2
//   StdAverage:     ρ,   u,   v,   w   | p,   ρe
// In a line task value 1 is the velocity along the line, values 2, 3 the other two
// (cyclic order: LDS slots 1+d, 2+d, 3+d for direction d).  Chandrasekhar returns the
// flux factored as (m, F_1, u_t1, u_t2, h/2): F = (m, F_1, m·u_t1, m·u_t2, m·h/2) with the
// energy component halved (line_pairs uses the node velocities in place of u_t1, u_t2).
template <int TP, int PATH>
__device__ __forceinline__ void twopoint(const float* Pi, const float* Pj, float inv_gm1, float* F) {
    if constexpr (TP == TP_CHANDRASEKHAR) {
        const float hs = Pi[0] + Pj[0], dA = Pi[0] - Pj[0];
        const float db = Pi[4] - Pj[4];
        const float rb = PATH == 2 ? 0.0f : rcp(Pi[4] + Pj[4]);
        const float fb = db * rb;
        const float ub = fb * fb;
        float rho, X;  // X = (F_β/(γ-1) + F_ρ)/2
        if constexpr (PATH == 2) {
            // Both log-means to first order in u (u < 1e-4: F = 1 + u/3, truncation < 2e-9),
            // without a reciprocal per pair:
            //   1/(β1+β2) = (H1+H2)(1 - f_β²) exactly, and f_β = (β1-β2)/(β1+β2) ≈
            //   (β1-β2)(H1+H2) inside the O(u) terms (relative error 2f⁴ < 2e-8);
            //   slot 5 holds -(1/6)/ρ, so fr = dA·(slot5_1 + slot5_2) = -f_ρ/3·(1 + O(f²)),
            //   ρ̂ = ρ̄(1 - f_ρ²/3) = hs + dA·fr and the F_ρ term f_ρ²/6 of X is 3/2·fr².
            const float sh = Pi[6] + Pj[6];
            const float fh = db * sh;
            const float uh = fh * fh;
            const float rbh = fmaf(-sh, uh, sh);
            const float fr = dA * (Pi[5] + Pj[5]);
            X = fmaf(uh, inv_gm1 * (1.0f / 6.0f), fmaf(fr * fr, 1.5f, 0.5f * (inv_gm1 + 1.0f)));
            rho = fmaf(dA, fr, hs);
            const float p = hs * rbh;
            const float un = Pi[1] + Pj[1];
            const float ut1 = Pi[2] + Pj[2];
            const float ut2 = Pi[3] + Pj[3];
            const float dot = fmaf(Pi[1], Pj[1], fmaf(Pi[2], Pj[2], Pi[3] * Pj[3]));
            const float m = rho * un;
            F[0] = m;
            F[1] = fmaf(m, un, p);
            F[2] = ut1;
            F[3] = ut2;
            F[4] = fmaf(rbh, X, dot);
            return;
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

// Rows i and j of the line accumulate -D♯_ij F and -D♯_ji F.  Rows 0 and N-1 already
// hold the lifted face fluxes; the pairs (0, j) carry the first contribution to every
// other row, so they set those accumulators.  Chandrasekhar: the transverse momentum
// components m (u_i + u_j)/2 are split, -Σ_j D♯_ij m_ij u_i = u_i × (the row's mass sum),
// which saves the two velocity sums of every pair.
template <int N, int TP, int PATH>
__device__ __forceinline__ void line_pairs(const float (&P)[N][7], float (&acc)[N][5],
                                           const float* __restrict__ Dsh, float inv_gm1) {
    const float f0 = acc[0][0], f1 = acc[N - 1][0];  // face terms preset in rows 0, N-1
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = i + 1; j < N; ++j) {
            float F[5];
            twopoint<TP, PATH>(P[i], P[j], inv_gm1, F);
            const float cij = Dsh[i * N + j], cji = Dsh[j * N + i];
            const bool setj = i == 0 && j < N - 1;  // first contribution to row j
            if constexpr (TP == TP_CHANDRASEKHAR) {
                const float mi = cij * F[0], mj = cji * F[0];
                acc[i][0] -= mi;
                acc[i][1] = fmaf(-cij, F[1], acc[i][1]);
                acc[j][0] = setj ? -mj : acc[j][0] - mj;
                acc[j][1] = setj ? -cji * F[1] : fmaf(-cji, F[1], acc[j][1]);
                // transverse momenta: -m (u_i + u_j)/2 per row; only the partner's half is
                // accumulated here, the own half is u_i times the row's mass sum (below)
#pragma unroll
                for (int v = 2; v < 4; ++v) {
                    acc[i][v] = fmaf(-mi, P[j][v], acc[i][v]);
                    acc[j][v] = setj ? -mj * P[i][v] : fmaf(-mj, P[i][v], acc[j][v]);
                }
                acc[i][4] = fmaf(-mi, F[4], acc[i][4]);
                acc[j][4] = setj ? -mj * F[4] : fmaf(-mj, F[4], acc[j][4]);
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    acc[i][v] = fmaf(-cij, F[v], acc[i][v]);
                    acc[j][v] = setj ? -cji * F[v] : fmaf(-cji, F[v], acc[j][v]);
                }
            }
        }
    }
    if constexpr (TP == TP_CHANDRASEKHAR) {
#pragma unroll
        for (int p = 0; p < N; ++p) {
            const float m = p == 0 ? acc[p][0] - f0 : (p == N - 1 ? acc[p][0] - f1 : acc[p][0]);
            acc[p][2] = fmaf(P[p][2], m, acc[p][2]);
            acc[p][3] = fmaf(P[p][3], m, acc[p][3]);
        }
    }
}

// Line addressing.  Node p of a line is block node base + p*stride; the momentum along the
// line and the two others (line-frame slots 1, 2, 3) are physical variables m[0..2] --
// per-lane offsets, so one instruction stream serves lines of every direction.  vel is
// the per-lane offset (1 + direction)*NN of the line's first velocity slot in P.
struct LineAddr {
    int base, stride, vel;
    int m[3];
};

// flux inputs of the N nodes of a line, line-frame order (slot arrays of NN floats)
template <int N, int NN, bool H>
__device__ __forceinline__ void load_line(const float* __restrict__ sP, const LineAddr& L, float (&P)[N][7]) {
#pragma unroll
    for (int p = 0; p < N; ++p) {
        const float* q = sP + L.base + p * L.stride;
        const float* qv = q + L.vel;
        P[p][0] = q[0];
        P[p][1] = qv[0];
        P[p][2] = qv[NN];
        P[p][3] = qv[2 * NN];
        P[p][4] = q[6 * NN];
        P[p][5] = q[7 * NN];
        P[p][6] = H ? q[8 * NN] : 0.0f;
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

// row sums of a line in line-frame variable order (the node phase undoes the permutation)
template <int N, int NN>
__device__ __forceinline__ void store_line(float* __restrict__ R, const LineAddr& L, const float (&acc)[N][5]) {
#pragma unroll
    for (int p = 0; p < N; ++p) {
        float* r = R + L.base + p * L.stride;
#pragma unroll
        for (int c = 0; c < 5; ++c) r[c * NN] = acc[p][c];
    }
}

// line-frame slot of physical variable v in a line of direction d (momenta cycle)
__host__ __device__ constexpr int lslot(int d, int v) { return (v == 0 || v == 4) ? v : ((v - 1 - d + 3) % 3) + 1; }

// Numerical flux on the face where a line ends (row R of the line) and its lift, which
// starts row R: acc[R] = -coef·[s(f_own + f_ext) + λ(Q_own - Q_ext)], in the line frame
// (slot 1 = momentum along the line); the energy entry is scaled by ehalf.  With the
// Chandrasekhar inputs (PO: ρ/2, u_n/2, H = p/(2ρ) of the own node in Po) the own side
// takes u_n, p and c² from them instead of dividing by ρ; the own node's positivity is
// already checked at staging.
template <int N, int R, bool PO>
__device__ __forceinline__ void face_init(const float (&o)[5], const float (&x)[5], const float* Po, float s,
                                          float coef, float gm1, float gamma, float diss, float ehalf,
                                          float (&acc)[N][5], bool& bad) {
    float fo[5], fx[5], ao, ax;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const float* Q = side ? x : o;
        float* f = side ? fx : fo;
        float p, un, c2;
        if (PO && !side) {
            un = Po[1] + Po[1];
            p = 4.0f * Po[0] * Po[6];
            c2 = (2.0f * gamma) * Po[6];
        } else {
            const float ir = rcp(Q[0]);
            const float m2 = fmaf(Q[1], Q[1], fmaf(Q[2], Q[2], Q[3] * Q[3]));
            p = gm1 * fmaf(-0.5f * m2, ir, Q[4]);
            un = Q[1] * ir;
            c2 = gamma * p * ir;
            bad |= !(c2 >= 0.0f);
        }
        const float a = fabsf(un) + sq(c2);
        f[0] = Q[1];
        f[1] = fmaf(Q[1], un, p);
        f[2] = Q[2] * un;
        f[3] = Q[3] * un;
        f[4] = (Q[4] + p) * un;
        if (side) ax = a; else ao = a;
    }
    const float cs = coef * s, cl = coef * fmaxf(ao, ax) * diss;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const float t = fmaf(-cs, fo[v] + fx[v], -cl * (o[v] - x[v]));
        acc[R][v] = v == 4 ? ehalf * t : t;
    }
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

// FIRST: stage 0 of a 2N scheme (the register is written, not read)
template <int N, int TP, int MODE, bool FIRST>
__global__ void __launch_bounds__((FCfg<N>::TPB)) FAST_WAVES k_fast3d(const FastArgs a) {
    using C = FCfg<N>;
    constexpr int NPE = C::NPE, NPF = C::NPF, NL = C::NL, EPB = C::EPB, TPB = C::TPB;
    constexpr int NN = C::NN, NPT = C::NPT;
    __shared__ float4 smem[(C::LDS + 3) / 4];
    float* const sm = reinterpret_cast<float*>(smem);
    float* const sQ = sm + C::OFF_Q;
    float* const sP = sm + C::OFF_P;
    float* const sG = sm + C::OFF_G;
    float2* const sF = reinterpret_cast<float2*>(sm + C::OFF_F);

    const int tid = threadIdx.x;
    const int e0 = xcd_remap(blockIdx.x, gridDim.x) * EPB;
    const int64_t Nd = a.ndofs;
    const float gamma = a.gamma, gm1 = a.gamma - 1.0f, inv_gm1 = a.inv_gm1;
    const int nel = a.ne - e0 < EPB ? a.ne - e0 : EPB;
    const uint32_t n0 = (uint32_t)e0 * NPE;  // first node of the block
    int err = 0;                              // domain-error bits, reported once at the end

    // ---- this thread's line task ----------------------------------------------------
    // One thread per (element, direction, line); lanes past the block's line tasks repeat
    // a valid line (their results are dropped), so the whole wave runs one stream.
    const int tl = tid < EPB * NL ? tid : tid - EPB * NL;
    const int lle = tl / NL, lt = tl - (tl / NL) * NL;
    const int ldir = lt / NPF, lf = lt - (lt / NPF) * NPF;
    const bool line_in = lle < nel;
    const bool line_ok = tid < EPB * NL && line_in;

    // face tables of both line ends (faces 2*ldir and 2*ldir+1), first: they feed the
    // exterior addresses
    int xnb[2], xinf[2];
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const int slot = (e0 + lle) * 6 + 2 * ldir + sd;
        xnb[sd] = line_ok ? a.nb[slot] : -2;
        xinf[sd] = line_ok ? a.finfo[slot] : 0;
    }

    // per face (lanes tid < 6*EPB): the face-geometry index and the side, loaded with the
    // face tables so that the dependent geometry loads below wait for these only
    const int fle = tid / 6, fk = tid - (tid / 6) * 6;
    const bool face_lane = tid < 6 * EPB && fle < nel;
    int fgi = 0, fside = 0;
    if (face_lane) {
        fgi = a.fgi[(e0 + fle) * 6 + fk];
        fside = finfo_side(a.finfo[(e0 + fle) * 6 + fk]);
    }

    // ---- staging ---------------------------------------------------------------------
    // element nodes, all five variables per node (lanes past the block's last element read 0)
    const uint32_t blk_bytes = (uint32_t)(nel * NPE) * 4u;
    float q[NPT][5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const rsrc_t rq = make_rsrc(a.Q + v * Nd + n0, blk_bytes);
#pragma unroll
        for (int r = 0; r < NPT; ++r) q[r][v] = bload(rq, (uint32_t)(tid + r * TPB) * 4u);
    }
    // own outward normal sign along the axis and the lift coefficient of each face,
    // issued before the exterior traces so that waiting for them does not wait for those
    float fsn = 0.0f, fjac = 0.0f, feg = 0.0f, fed = 0.0f;
    if (face_lane) {
        const float* eg = a.egeom + (int64_t)(e0 + fle) * a.egeom_stride;
        fsn = a.fgeom[(int64_t)fgi * 10 + (fk >> 1)];
        fjac = a.fgeom[(int64_t)fgi * 10 + 9];
        feg = eg[4 + (fk >> 1)];
        if (fk < 3) fed = eg[fk];
    }
    // exterior traces at both line ends: the neighbour's node on interior faces
    // (branch-free; other lanes, including those past the block's line tasks, read 0
    // without a memory access), issued after the element nodes so that the staging below
    // waits only for those
    uint32_t xo[2];
    bool inblk[2];  // neighbour is one of this block's elements: read from LDS after staging
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const int kp = finfo_nbpos(xinf[sd]);
        inblk[sd] = xnb[sd] >= e0 && xnb[sd] < e0 + nel;
        xo[sd] = xnb[sd] >= 0 && !inblk[sd]
                     ? ((uint32_t)xnb[sd] * NPE + line_node<3, N>(kp >> 1, lf, (kp & 1) ? N - 1 : 0)) * 4u
                     : 0xffffffffu;
    }
    float xr[2][5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const rsrc_t rv = make_rsrc(a.Q + v * Nd, (uint32_t)Nd * 4u);
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) xr[sd][v] = bload(rv, xo[sd]);
    }
    // flux inputs and LDS writes
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int j = tid + r * TPB;
        if (j >= NN) continue;
        const float rho = q[r][0];
        const float ir = rcp(rho);
        float in[6], u, v, w, p, hir = 0.0f;
        if constexpr (TP == TP_CHANDRASEKHAR) {
            // halved velocities straight from 1/(2ρ): 0.5ρ|u|² = 2ρ|u/2|²
            hir = 0.5f * ir;
            u = q[r][1] * hir; v = q[r][2] * hir; w = q[r][3] * hir;
            p = gm1 * fmaf(-2.0f * rho, fmaf(u, u, fmaf(v, v, w * w)), q[r][4]);
        } else {
            u = q[r][1] * ir; v = q[r][2] * ir; w = q[r][3] * ir;
            p = gm1 * fmaf(-0.5f * rho, fmaf(u, u, fmaf(v, v, w * w)), q[r][4]);
        }
        if (j < nel * NPE && (!(rho > 0.0f) || !(p > 0.0f))) err |= ERR_LOG_NEG;
#pragma unroll
        for (int c = 0; c < 5; ++c) sQ[c * NN + j] = q[r][c];
        if constexpr (TP == TP_CHANDRASEKHAR) {
            in[0] = 0.5f * rho; in[1] = u; in[2] = v; in[3] = w;
            in[4] = in[0] * rcp(p); in[5] = (-1.0f / 6.0f) * ir;
        } else if constexpr (TP == TP_KENNEDY_GRUBER) {
            in[0] = 0.5f * rho; in[1] = 0.5f * u; in[2] = 0.5f * v; in[3] = 0.5f * w;
            in[4] = 0.5f * p; in[5] = 0.5f * q[r][4] * ir;
        } else {
            in[0] = rho; in[1] = u; in[2] = v; in[3] = w; in[4] = p; in[5] = q[r][4];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) sP[c * NN + j] = in[c];
        sP[4 * NN + j] = in[1];
        sP[5 * NN + j] = in[2];
        sP[6 * NN + j] = in[4];
        sP[7 * NN + j] = in[5];
        if constexpr (TP == TP_CHANDRASEKHAR) sP[8 * NN + j] = p * hir;
    }
    if (face_lane) {
        const float lift = (fk & 1) ? a.lift_p : a.lift_m;
        sF[tid] = make_float2(fside ? -fsn : fsn, 0.5f * feg * lift * fjac);
        if (fk < 3) sG[fle * 4 + fk] = fed;
    }
    // boundary faces: Dirichlet or inflow states, or the own node (outflow, slip)
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        if (xnb[sd] == -1) {
            const int k = 2 * ldir + sd;
            const int bd = finfo_bd(xinf[sd]);
            const int type = a.bctype[bd];
            const float* src;
            int64_t vs = 1;
            if (type == BC_DIRICHLET) {
                src = a.bcstate + ((int64_t)a.fid[(e0 + lle) * 6 + k] * NPF + lf) * 5;
            } else if (type == BC_INFLOW) {
                src = a.bcconst + bd * 5;
            } else {
                src = a.Q + n0 + lle * NPE + line_node<3, N>(ldir, lf, sd ? N - 1 : 0);
                vs = Nd;
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) xr[sd][v] = src[v * vs];
        }
    }
    __syncthreads();
    // neighbours in this block (the x neighbour of every other element): their node
    // states are already in LDS
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        if (inblk[sd]) {
            const int kp = finfo_nbpos(xinf[sd]);
            const int node = (xnb[sd] - e0) * NPE + line_node<3, N>(kp >> 1, lf, (kp & 1) ? N - 1 : 0);
#pragma unroll
            for (int v = 0; v < 5; ++v) xr[sd][v] = sQ[v * NN + node];
        }
    }
    // exterior traces in the line frame; a slip wall reflects the normal momentum
    float xl[2][5];
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const bool slip = xnb[sd] == -1 && a.bctype[finfo_bd(xinf[sd])] == BC_SLIP;
        const float un = ldir == 0 ? xr[sd][1] : (ldir == 1 ? xr[sd][2] : xr[sd][3]);
        xl[sd][0] = xr[sd][0];
        xl[sd][1] = slip ? -un : un;
        xl[sd][2] = ldir == 0 ? xr[sd][2] : (ldir == 1 ? xr[sd][3] : xr[sd][1]);
        xl[sd][3] = ldir == 0 ? xr[sd][3] : (ldir == 1 ? xr[sd][1] : xr[sd][2]);
        xl[sd][4] = xr[sd][4];
    }

    // ---- line tasks ------------------------------------------------------------------
    LineAddr L;
    L.base = lle * NPE + (ldir == 0 ? N * lf : (ldir == 1 ? lf % N + NPF * (lf / N) : lf));
    L.stride = ldir == 0 ? 1 : (ldir == 1 ? N : NPF);
    L.vel = (1 + ldir) * NN;
    L.m[0] = 1 + ldir;
    L.m[1] = ldir == 2 ? 1 : 2 + ldir;
    L.m[2] = ldir == 0 ? 3 : ldir;
    float acc[N][5];
    {
        // face fluxes at the two line ends start rows 0 and N-1
        float o0[5], o1[5];
        load_vars<NN>(sQ + L.base, L, o0);
        load_vars<NN>(sQ + L.base + (N - 1) * L.stride, L, o1);
        const float2 g0 = sF[lle * 6 + 2 * ldir], g1 = sF[lle * 6 + 2 * ldir + 1];
        const float ehalf = TP == TP_CHANDRASEKHAR ? 0.5f : 1.0f;  // energy row accumulated halved
        constexpr bool PO = TP == TP_CHANDRASEKHAR;
        float P[N][7];
        load_line<N, NN, PO>(sP, L, P);
        bool bad = false;
        face_init<N, 0, PO>(o0, xl[0], P[0], g0.x, g0.y, gm1, gamma, a.diss, ehalf, acc, bad);
        face_init<N, N - 1, PO>(o1, xl[1], P[N - 1], g1.x, g1.y, gm1, gamma, a.diss, ehalf, acc, bad);
        if (bad && line_ok) err |= ERR_SQRT_NEG;

        if constexpr (TP == TP_CHANDRASEKHAR) {
            float amax = P[0][0], amin = P[0][0], bmax = P[0][4], bmin = P[0][4];
#pragma unroll
            for (int p = 1; p < N; ++p) {
                amax = fmaxf(amax, P[p][0]); amin = fminf(amin, P[p][0]);
                bmax = fmaxf(bmax, P[p][4]); bmin = fminf(bmin, P[p][4]);
            }
            // u = ((max - min)/(max + min))² of ρ and β along the line, compared without a
            // division: (max - min)² < u0 (max + min)² (a margin of 1e-4 relative for the
            // rounding of the squares)
            const float da = amax - amin, sa = amax + amin, db = bmax - bmin, sb = bmax + bmin;
            const float qa = da * da * 1.0001f, pa = sa * sa, qb = db * db * 1.0001f, pb = sb * sb;
            if (qa < 1e-4f * pa && qb < 1e-4f * pb) line_pairs<N, TP, 2>(P, acc, a.dsharp, inv_gm1);
            else if (qa < 1e-2f * pa && qb < 1e-2f * pb) line_pairs<N, TP, 3>(P, acc, a.dsharp, inv_gm1);
            else line_pairs<N, TP, 0>(P, acc, a.dsharp, inv_gm1);
#pragma unroll
            for (int p = 0; p < N; ++p) acc[p][4] += acc[p][4];  // energy row was halved
        } else {
            line_pairs<N, TP, 0>(P, acc, a.dsharp, inv_gm1);
        }
    }

    // the stage epilogue's second operand (RK register or SSP base), loaded before the
    // barrier so its latency overlaps the other waves' line tasks
    const float* aux = (MODE == MODE_2N && !FIRST) ? a.reg
                     : ((MODE == MODE_SSP2 || MODE == MODE_SSP3) ? a.ubase : nullptr);
    float xv[NPT][5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const rsrc_t rx = make_rsrc(aux ? aux + v * Nd + n0 : a.Q, aux ? blk_bytes : 0u);
#pragma unroll
        for (int r = 0; r < NPT; ++r) xv[r][v] = bload(rx, (uint32_t)(tid + r * TPB) * 4u);
    }
    // the stage input of this thread's nodes, before its LDS region takes row sums
    float uq[NPT][5];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int j = tid + r * TPB;
#pragma unroll
        for (int v = 0; v < 5; ++v) uq[r][v] = sQ[v * NN + (j < NN ? j : 0)];
    }
    __syncthreads();  // flux inputs and node states are dead from here
    if (line_ok) store_line<N, NN>(ldir == 0 ? sP : (ldir == 1 ? sP + 5 * NN : sQ), L, acc);
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
            kv[r][v] = fmaf(s0, sP[lslot(0, v) * NN + jj],
                            fmaf(s1, sP[(5 + lslot(1, v)) * NN + jj], s2 * sQ[lslot(2, v) * NN + jj]));
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const rsrc_t ro = make_rsrc(a.out + v * Nd + n0, blk_bytes);
        const rsrc_t rr = make_rsrc(a.reg + v * Nd + n0, MODE == MODE_2N ? blk_bytes : 0u);
#pragma unroll
        for (int r = 0; r < NPT; ++r) {
            const int j = tid + r * TPB;
            const uint32_t off = (uint32_t)j * 4u;
            const float uin = uq[r][v];
            const float k = kv[r][v], xx = xv[r][v];
            float out;
            if constexpr (MODE == MODE_RHS) {
                out = k;
            } else if constexpr (MODE == MODE_2N) {
                const float t = FIRST ? a.dt * k : fmaf(a.A, xx, a.dt * k);
                bstore(rr, off, t);
                out = fmaf(a.B, t, uin);
            } else if constexpr (MODE == MODE_SSP1) {
                out = fmaf(a.dt, k, uin);
            } else if constexpr (MODE == MODE_SSP2) {
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
    const dim3 g(nblocks), b(FCfg<N>::TPB);
    switch (a.mode) {
        case MODE_RHS: hipLaunchKernelGGL((k_fast3d<N, TP, MODE_RHS, false>), g, b, 0, s, a); break;
        case MODE_2N:
            if (a.first) hipLaunchKernelGGL((k_fast3d<N, TP, MODE_2N, true>), g, b, 0, s, a);
            else hipLaunchKernelGGL((k_fast3d<N, TP, MODE_2N, false>), g, b, 0, s, a);
            break;
        case MODE_SSP1: hipLaunchKernelGGL((k_fast3d<N, TP, MODE_SSP1, false>), g, b, 0, s, a); break;
        case MODE_SSP2: hipLaunchKernelGGL((k_fast3d<N, TP, MODE_SSP2, false>), g, b, 0, s, a); break;
        case MODE_SSP3: hipLaunchKernelGGL((k_fast3d<N, TP, MODE_SSP3, false>), g, b, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
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
