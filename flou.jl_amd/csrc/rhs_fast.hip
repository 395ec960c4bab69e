// FP32 ("chosen form") fused element kernel for the 2-D and 3-D split-form Euler rhs on
// axis-aligned affine elements (Flou's CartesianMesh quads and hexahedra), GLL nodes,
// LxF(StdAverage) or StdAverage surface flux, with the RK stage epilogue.
//
// Same mathematics as SplitDivOperator + interface_fluxes! + the FR lift
// (OpDivergence.jl:200-299, Interfaces.jl:111-136), organised for gfx950:
//   * EPB elements per block, one thread per (element, direction, line) -- every thread
//     does the same work.  Each thread first loads the face tables and then the two
//     exterior traces of its line (neighbour nodes or boundary states, into registers),
//     issued after the block's own nodes so that the staging waits for those only; a
//     neighbour inside the block is read from LDS after staging instead;
//   * staging: every thread loads all NV = ND + 2 variables of its nodes (coalesced, one
//     uniform base per variable) and writes the node states and the two-point flux
//     inputs to LDS;
//   * line phase: the numerical flux on the two face nodes where the line ends, lifted
//     into the two end rows (GLL: ∂g lives on the ends), then all N(N-1)/2 pairs of the
//     line, each evaluated once (F♯ is symmetric), both rows accumulated in registers
//     (F♯ never touches memory).  Line-frame velocity slots are consecutive in LDS
//     (u, v, w, u, v in 3-D; u, v, u in 2-D), so one per-lane offset serves every
//     direction;
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
#include "fast_util.hpp"

namespace flouhip {
namespace fast {

// Elements per block: 3-D, the most that keep four waves per SIMD (2 at N = 5); 2-D, a
// count whose line tasks fill whole waves (or nearly: 126 of 128 lanes at N = 7).
template <int ND, int N>
constexpr int fast_epb() {
    return ND == 3 ? (N == 4 ? 4 : (N == 5 ? 3 : (N == 6 ? 2 : 1)))
                   : (N == 4 ? 16 : (N == 5 ? 32 : (N == 6 ? 16 : (N == 7 ? 9 : 8))));
}

template <int ND, int N>
struct FCfg {
    static constexpr int NV = ND + 2;                                 // conserved variables
    static constexpr int NFC = 2 * ND;                                // faces per element
    static constexpr int NPE = ipow(N, ND);
    static constexpr int NPF = ipow(N, ND - 1);
    static constexpr int NL = ND * NPF;                               // line tasks per element
    static constexpr int EPB = fast_epb<ND, N>();
    static constexpr int TPB = ((EPB * NL + 63) / 64) * 64;           // threads per block
    static constexpr int NN = EPB * NPE;                              // nodes per block
    static constexpr int NPT = (NN + TPB - 1) / TPB;                  // nodes per thread
    // 2-D: slot arrays hold block node j at pnode(j) = j + j/N, one float of padding after
    // every row of N nodes (plane stride NNP), so that the lines a wave reads together, N
    // nodes apart, spread over the LDS banks (2-D N=8: -14 % per stage; neutral in 3-D,
    // which keeps the unpadded layout).
    static constexpr bool PAD = ND == 2;
    static constexpr int NNP = PAD ? NN + NN / N : NN;
    // LDS layout in floats, arrays indexed by padded block node.  Q: [NV][NNP] node states.
    // P: two-point flux inputs -- slot 0 density, slots 1 .. 2ND-1 the velocities in cyclic
    // order (u, v, w, u, v in 3-D; u, v, u in 2-D), so the ND line-frame velocity slots of a
    // line in any direction are consecutive, then three more scalars (PIN slots).  Once
    // every line is done the row sums (line-frame variable order) of directions
    // 0 .. ND-2 go to P, those of direction ND-1 to Q.  G: [EPB][4] Ja_dd/J per element;
    // F: [EPB][2ND] float2 face sign and lift coefficient.
    static constexpr int PIN = 2 * ND + 3;
    static constexpr int PSL = PIN > (ND - 1) * NV ? PIN : (ND - 1) * NV;
    static constexpr int OFF_Q = 0;
    static constexpr int OFF_P = r4(NV * NNP);
    static constexpr int OFF_G = r4(OFF_P + PSL * NNP);
    static constexpr int OFF_F = OFF_G + 4 * EPB;
    static constexpr int LDS = OFF_F + 2 * NFC * EPB;
};

// Node inputs (the velocity repeated cyclically as above; a last one, H = 1/(4β) =
// p/(2ρ), for the Chandrasekhar flux):
//   Chandrasekhar:  ρ/2, u/2 .. | β = ρ/(2p), -(1/6)/ρ | H
//   Kennedy-Gruber: ρ/2, u/2 .. | p/2, e/2
//   StdAverage:     ρ,   u   .. | p,   ρe
// In a line task value 1 is the velocity along the line, values 2 .. ND the others
// (cyclic order: LDS slots 1+d .. ND+d for direction d); values ND+1, ND+2, ND+3 the
// three scalars.  The line-frame input array of a node: Pl[0] density, Pl[1..ND]
// velocities, Pl[ND+1..ND+3] scalars.  Chandrasekhar returns the flux factored as
// (m, F_1, -, .., h/2): F = (m, F_1, m·u_t.., m·h/2) with the energy component halved
// (line_pairs uses the node velocities for the transverse momenta).
template <int ND, int TP, int PATH>
__device__ __forceinline__ void twopoint(const float* Pi, const float* Pj, float inv_gm1, float* F) {
    constexpr int NV = ND + 2, SB = ND + 1;  // first scalar slot
    if constexpr (TP == TP_CHANDRASEKHAR) {
        const float hs = Pi[0] + Pj[0], dA = Pi[0] - Pj[0];
        const float db = Pi[SB] - Pj[SB];
        const float rb = PATH == 2 ? 0.0f : rcp(Pi[SB] + Pj[SB]);
        const float fb = db * rb;
        const float ub = fb * fb;
        float rho, X;  // X = (F_β/(γ-1) + F_ρ)/2
        float dot = Pi[ND] * Pj[ND];
#pragma unroll
        for (int k = ND - 1; k >= 1; --k) dot = fmaf(Pi[k], Pj[k], dot);
        const float un = Pi[1] + Pj[1];
        if constexpr (PATH == 2) {
            // Both log-means to first order in u (u < 1e-4: F = 1 + u/3, truncation < 2e-9),
            // without a reciprocal per pair:
            //   1/(β1+β2) = (H1+H2)(1 - f_β²) exactly, and f_β = (β1-β2)/(β1+β2) ≈
            //   (β1-β2)(H1+H2) inside the O(u) terms (relative error 2f⁴ < 2e-8);
            //   the second scalar holds -(1/6)/ρ, so fr = dA·(s_1 + s_2) = -f_ρ/3·(1 + O(f²)),
            //   ρ̂ = ρ̄(1 - f_ρ²/3) = hs + dA·fr and the F_ρ term f_ρ²/6 of X is 3/2·fr².
            const float sh = Pi[SB + 2] + Pj[SB + 2];
            const float fh = db * sh;
            const float uh = fh * fh;
            const float rbh = fmaf(-sh, uh, sh);
            const float fr = dA * (Pi[SB + 1] + Pj[SB + 1]);
            X = fmaf(uh, inv_gm1 * (1.0f / 6.0f), fmaf(fr * fr, 1.5f, 0.5f * (inv_gm1 + 1.0f)));
            rho = fmaf(dA, fr, hs);
            const float p = hs * rbh;
            const float m = rho * un;
            F[0] = m;
            F[1] = fmaf(m, un, p);
            F[NV - 1] = fmaf(rbh, X, dot);
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
                else Fb = 0.5f * (lg(Pi[SB]) - lg(Pj[SB])) * rcp(fb);
            }
            X = fmaf(Fb, 0.5f * inv_gm1, 0.5f * Fr);
            rho = hs * rcp(Fr);
        }
        const float p = hs * rb;
        const float m = rho * un;
        // factored: F = (m, F[1], m·u_t.., m·F[NV-1])
        F[0] = m;
        F[1] = fmaf(m, un, p);
        F[NV - 1] = fmaf(rb, X, dot);
    } else if constexpr (TP == TP_KENNEDY_GRUBER) {
        const float rho = Pi[0] + Pj[0];
        const float un = Pi[1] + Pj[1];
        const float p = Pi[SB] + Pj[SB];
        const float e = Pi[SB + 1] + Pj[SB + 1];
        const float m = rho * un;
        F[0] = m;
        F[1] = fmaf(m, un, p);
#pragma unroll
        for (int k = 2; k <= ND; ++k) F[k] = m * (Pi[k] + Pj[k]);
        F[NV - 1] = fmaf(m, e, p * un);
    } else {
        const float mi = Pi[0] * Pi[1], mj = Pj[0] * Pj[1];
        F[0] = 0.5f * (mi + mj);
        F[1] = 0.5f * (fmaf(mi, Pi[1], Pi[SB]) + fmaf(mj, Pj[1], Pj[SB]));
#pragma unroll
        for (int k = 2; k <= ND; ++k) F[k] = 0.5f * (mi * Pi[k] + mj * Pj[k]);
        F[NV - 1] = 0.5f * ((Pi[SB + 1] + Pi[SB]) * Pi[1] + (Pj[SB + 1] + Pj[SB]) * Pj[1]);
    }
}

// Rows i and j of the line accumulate -D♯_ij F and -D♯_ji F.  Rows 0 and N-1 already
// hold the lifted face fluxes; the pairs (0, j) carry the first contribution to every
// other row, so they set those accumulators.  Chandrasekhar: the transverse momentum
// components m (u_i + u_j)/2 are split, -Σ_j D♯_ij m_ij u_i = u_i × (the row's mass sum),
// which saves the velocity sums of every pair.
template <int ND, int N, int TP, int PATH>
__device__ __forceinline__ void line_pairs(const float (&P)[N][ND + 4], float (&acc)[N][ND + 2],
                                           const float* __restrict__ Dsh, float inv_gm1) {
    constexpr int NV = ND + 2;
    const float f0 = acc[0][0], f1 = acc[N - 1][0];  // face terms preset in rows 0, N-1
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = i + 1; j < N; ++j) {
            float F[NV];
            twopoint<ND, TP, PATH>(P[i], P[j], inv_gm1, F);
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
                for (int v = 2; v <= ND; ++v) {
                    acc[i][v] = fmaf(-mi, P[j][v], acc[i][v]);
                    acc[j][v] = setj ? -mj * P[i][v] : fmaf(-mj, P[i][v], acc[j][v]);
                }
                acc[i][NV - 1] = fmaf(-mi, F[NV - 1], acc[i][NV - 1]);
                acc[j][NV - 1] = setj ? -mj * F[NV - 1] : fmaf(-mj, F[NV - 1], acc[j][NV - 1]);
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) {
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
#pragma unroll
            for (int v = 2; v <= ND; ++v) acc[p][v] = fmaf(P[p][v], m, acc[p][v]);
        }
    }
}

// Line addressing.  Node p of a line is block node base + p*stride; the momentum along the
// line and the others (line-frame slots 1 .. ND) are physical variables m[0 .. ND-1] --
// per-lane offsets, so one instruction stream serves lines of every direction.  vel is
// the per-lane offset (1 + direction)*NNP of the line's first velocity slot in P.
template <int ND>
struct LineAddr {
    int base, stride, vel;
    int m[ND];
};

// flux inputs of the N nodes of a line, line-frame order (slot arrays of plane stride NN)
template <int ND, int N, int NN, bool H>
__device__ __forceinline__ void load_line(const float* __restrict__ sP, const LineAddr<ND>& L,
                                          float (&P)[N][ND + 4]) {
#pragma unroll
    for (int p = 0; p < N; ++p) {
        const float* q = sP + L.base + p * L.stride;
        const float* qv = q + L.vel;
        P[p][0] = q[0];
#pragma unroll
        for (int k = 0; k < ND; ++k) P[p][1 + k] = qv[k * NN];
        P[p][ND + 1] = q[(2 * ND) * NN];
        P[p][ND + 2] = q[(2 * ND + 1) * NN];
        P[p][ND + 3] = H ? q[(2 * ND + 2) * NN] : 0.0f;
    }
}

// the NV variables in line-frame order from an array with variable stride VS
template <int ND, int VS>
__device__ __forceinline__ void load_vars(const float* __restrict__ s, const LineAddr<ND>& L, float (&o)[ND + 2]) {
    o[0] = s[0];
#pragma unroll
    for (int k = 0; k < ND; ++k) o[1 + k] = s[L.m[k] * VS];
    o[ND + 1] = s[(ND + 1) * VS];
}

// row sums of a line in line-frame variable order (the node phase undoes the permutation)
template <int ND, int N, int NN>
__device__ __forceinline__ void store_line(float* __restrict__ R, const LineAddr<ND>& L,
                                           const float (&acc)[N][ND + 2]) {
#pragma unroll
    for (int p = 0; p < N; ++p) {
        float* r = R + L.base + p * L.stride;
#pragma unroll
        for (int c = 0; c < ND + 2; ++c) r[c * NN] = acc[p][c];
    }
}


// line-frame slot of physical variable v in a line of direction d (momenta cycle)
template <int ND>
__host__ __device__ constexpr int lslot(int d, int v) {
    return (v == 0 || v == ND + 1) ? v : ((v - 1 - d + ND) % ND) + 1;
}

// Numerical flux on the face where a line ends (row R of the line) and its lift, which
// starts row R: acc[R] = -coef·[s(f_own + f_ext) + λ(Q_own - Q_ext)], in the line frame
// (slot 1 = momentum along the line); the energy entry is scaled by ehalf.  With the
// Chandrasekhar inputs (PO: ρ/2, u_n/2, H = p/(2ρ) of the own node in Po) the own side
// takes u_n, p and c² from them instead of dividing by ρ; the own node's positivity is
// already checked at staging.
template <int ND, int N, int R, bool PO>
__device__ __forceinline__ void face_init(const float (&o)[ND + 2], const float (&x)[ND + 2], const float* Po,
                                          float s, float coef, float gm1, float gamma, float diss, float ehalf,
                                          float (&acc)[N][ND + 2], bool& bad) {
    constexpr int NV = ND + 2;
    float fo[NV], fx[NV], ao, ax;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const float* Q = side ? x : o;
        float* f = side ? fx : fo;
        float p, un, c2;
        if (PO && !side) {
            un = Po[1] + Po[1];
            p = 4.0f * Po[0] * Po[ND + 3];
            c2 = (2.0f * gamma) * Po[ND + 3];
        } else {
            const float ir = rcp(Q[0]);
            float m2 = Q[ND] * Q[ND];
#pragma unroll
            for (int k = ND - 1; k >= 1; --k) m2 = fmaf(Q[k], Q[k], m2);
            p = gm1 * fmaf(-0.5f * m2, ir, Q[NV - 1]);
            un = Q[1] * ir;
            c2 = gamma * p * ir;
            bad |= !(c2 >= 0.0f);
        }
        const float a = fabsf(un) + sq(c2);
        f[0] = Q[1];
        f[1] = fmaf(Q[1], un, p);
#pragma unroll
        for (int k = 2; k <= ND; ++k) f[k] = Q[k] * un;
        f[NV - 1] = (Q[NV - 1] + p) * un;
        if (side) ax = a; else ao = a;
    }
    const float cs = coef * s, cl = coef * fmaxf(ao, ax) * diss;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const float t = fmaf(-cs, fo[v] + fx[v], -cl * (o[v] - x[v]));
        acc[R][v] = v == NV - 1 ? ehalf * t : t;
    }
}

// FIRST: stage 0 of a 2N scheme (the register is written, not read)
template <int ND, int N, int TP, int MODE, bool FIRST>
__global__ void __launch_bounds__((FCfg<ND, N>::TPB)) k_fast(const FastArgs a) {
    using C = FCfg<ND, N>;
    constexpr int NV = C::NV, NFC = C::NFC, NPE = C::NPE, NPF = C::NPF, NL = C::NL, EPB = C::EPB;
    constexpr int TPB = C::TPB, NN = C::NN, NNP = C::NNP, NPT = C::NPT;
    constexpr int FG = 3 * ND + 1;  // face geometry entry: frame (n, t, b) and surface J
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
    // Direction-major: consecutive lanes run the lines of one direction over the block's
    // elements, so most waves read LDS with one stride (C3: -2 %, C2: -5 % per stage
    // against element-major lanes, whose waves mixed two directions' strides).  When one
    // direction's lines fit in a wave and the block has one wave per direction (3-D N=7,
    // 2-D N=7), wave d runs direction d and its spare lanes repeat lines of it (C5: -2.5 %).
    constexpr bool WDIR = EPB * NPF <= 64 && TPB == ND * 64;
    static_assert(!WDIR || 2 * EPB * NPF >= 64, "spare lanes repeat lines of their wave's direction");
    // (wave-uniform wave index: the line direction of a WDIR wave lives in a scalar register)
    const int wk = tid & 63, wd = WDIR ? __builtin_amdgcn_readfirstlane(tid >> 6) : 0;
    const bool task = WDIR ? wk < EPB * NPF : tid < EPB * NL;
    const int tl = WDIR ? wd * (EPB * NPF) + (task ? wk : wk - EPB * NPF) : (task ? tid : tid - EPB * NL);
    const int ldir = WDIR ? wd : tl / (EPB * NPF), lr = tl - ldir * (EPB * NPF);
    const int lle = lr / NPF, lf = lr - lle * NPF;
    const bool line_in = lle < nel;
    const bool line_ok = task && line_in;

    // face tables of both line ends (faces 2*ldir and 2*ldir+1), first: they feed the
    // exterior addresses
    int xnb[2], xinf[2];
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const int slot = (e0 + lle) * NFC + 2 * ldir + sd;
        xnb[sd] = line_ok ? a.nb[slot] : -2;
        xinf[sd] = line_ok ? a.finfo[slot] : 0;
    }

    // per face (lanes tid < NFC*EPB): the face-geometry index and the side, loaded with the
    // face tables so that the dependent geometry loads below wait for these only
    const int fle = tid / NFC, fk = tid - (tid / NFC) * NFC;
    const bool face_lane = tid < NFC * EPB && fle < nel;
    int fgi = 0, fside = 0;
    if (face_lane) {
        fgi = a.fgi[(e0 + fle) * NFC + fk];
        fside = finfo_side(a.finfo[(e0 + fle) * NFC + fk]);
    }

    // ---- staging ---------------------------------------------------------------------
    // element nodes, all variables per node (lanes past the block's last element read 0)
    const uint32_t blk_bytes = (uint32_t)(nel * NPE) * 4u;
    float q[NPT][NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const rsrc_t rq = make_rsrc(a.Q + v * Nd + n0, blk_bytes);
#pragma unroll
        for (int r = 0; r < NPT; ++r) q[r][v] = bload(rq, (uint32_t)(tid + r * TPB) * 4u);
    }
    // own outward normal sign along the axis and the lift coefficient of each face,
    // issued before the exterior traces so that waiting for them does not wait for those
    float fsn = 0.0f, fjac = 0.0f, feg = 0.0f, fed = 0.0f;
    if (face_lane) {
        const float* eg = a.egeom + (int64_t)(e0 + fle) * a.egeom_stride;
        fsn = a.fgeom[(int64_t)fgi * FG + (fk >> 1)];
        fjac = a.fgeom[(int64_t)fgi * FG + 3 * ND];
        feg = eg[4 + (fk >> 1)];
        if (fk < ND) fed = eg[fk];
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
                     ? ((uint32_t)xnb[sd] * NPE + line_node<ND, N>(kp >> 1, lf, (kp & 1) ? N - 1 : 0)) * 4u
                     : 0xffffffffu;
    }
    float xr[2][NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const rsrc_t rv = make_rsrc(a.Q + v * Nd, (uint32_t)Nd * 4u);
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) xr[sd][v] = bload(rv, xo[sd]);
    }
    // flux inputs and LDS writes
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int j = tid + r * TPB;
        if (j >= NN) continue;
        const int pj = pnode<N, C::PAD>(j);
        const float rho = q[r][0];
        const float ir = rcp(rho);
        float in[ND + 4], u[ND], hir = 0.0f;
        const float uscale = TP == TP_CHANDRASEKHAR ? 0.5f * ir : ir;
        if constexpr (TP == TP_CHANDRASEKHAR) hir = uscale;  // halved velocities from 1/(2ρ)
#pragma unroll
        for (int k = 0; k < ND; ++k) u[k] = q[r][1 + k] * uscale;
        float uu = u[ND - 1] * u[ND - 1];
#pragma unroll
        for (int k = ND - 2; k >= 0; --k) uu = fmaf(u[k], u[k], uu);
        // 0.5ρ|u|² = 2ρ|u/2|²
        const float p = gm1 * fmaf((TP == TP_CHANDRASEKHAR ? -2.0f : -0.5f) * rho, uu, q[r][NV - 1]);
        if (j < nel * NPE && (!(rho > 0.0f) || !(p > 0.0f))) err |= ERR_LOG_NEG;
#pragma unroll
        for (int c = 0; c < NV; ++c) sQ[c * NNP + pj] = q[r][c];
        if constexpr (TP == TP_CHANDRASEKHAR) {
            in[0] = 0.5f * rho;
            in[ND + 1] = in[0] * rcp(p);
            in[ND + 2] = (-1.0f / 6.0f) * ir;
            in[ND + 3] = p * hir;
        } else if constexpr (TP == TP_KENNEDY_GRUBER) {
            in[0] = 0.5f * rho;
#pragma unroll
            for (int k = 0; k < ND; ++k) u[k] *= 0.5f;
            in[ND + 1] = 0.5f * p;
            in[ND + 2] = 0.5f * q[r][NV - 1] * ir;
        } else {
            in[0] = rho;
            in[ND + 1] = p;
            in[ND + 2] = q[r][NV - 1];
        }
        sP[pj] = in[0];
#pragma unroll
        for (int c = 0; c < 2 * ND - 1; ++c) sP[(1 + c) * NNP + pj] = u[c % ND];
        sP[(2 * ND) * NNP + pj] = in[ND + 1];
        sP[(2 * ND + 1) * NNP + pj] = in[ND + 2];
        if constexpr (TP == TP_CHANDRASEKHAR) sP[(2 * ND + 2) * NNP + pj] = in[ND + 3];
    }
    if (face_lane) {
        const float lift = (fk & 1) ? a.lift_p : a.lift_m;
        sF[tid] = make_float2(fside ? -fsn : fsn, 0.5f * feg * lift * fjac);
        if (fk < ND) sG[fle * 4 + fk] = fed;
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
                src = a.bcstate + ((int64_t)a.fid[(e0 + lle) * NFC + k] * NPF + lf) * NV;
            } else if (type == BC_INFLOW) {
                src = a.bcconst + bd * NV;
            } else {
                src = a.Q + n0 + lle * NPE + line_node<ND, N>(ldir, lf, sd ? N - 1 : 0);
                vs = Nd;
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) xr[sd][v] = src[v * vs];
        }
    }
    __syncthreads();
    // neighbours in this block: their node states are already in LDS
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        if (inblk[sd]) {
            const int kp = finfo_nbpos(xinf[sd]);
            const int node = pnode<N, C::PAD>((xnb[sd] - e0) * NPE + line_node<ND, N>(kp >> 1, lf, (kp & 1) ? N - 1 : 0));
#pragma unroll
            for (int v = 0; v < NV; ++v) xr[sd][v] = sQ[v * NNP + node];
        }
    }
    // ---- line tasks ------------------------------------------------------------------
    LineAddr<ND> L;
    L.base = pnode<N, C::PAD>(lle * NPE + line_node<ND, N>(ldir, lf, 0));
    L.stride = ldir == 0 ? 1 : (ldir == 1 ? (C::PAD ? N + 1 : N) : (C::PAD ? N * (N + 1) : NPF));
    L.vel = (1 + ldir) * NNP;
#pragma unroll
    for (int k = 0; k < ND; ++k) L.m[k] = 1 + (ldir + k) % ND;
    // exterior traces in the line frame; a slip wall reflects the normal momentum
    float xl[2][NV];
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const bool slip = xnb[sd] == -1 && a.bctype[finfo_bd(xinf[sd])] == BC_SLIP;
        xl[sd][0] = xr[sd][0];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            // momentum 1 + (ldir + k) % ND, selected without a per-lane array index
            float m = xr[sd][1 + k];
#pragma unroll
            for (int d = 1; d < ND; ++d)
                if (ldir == d) m = xr[sd][1 + (d + k) % ND];
            xl[sd][1 + k] = (k == 0 && slip) ? -m : m;
        }
        xl[sd][NV - 1] = xr[sd][NV - 1];
    }
    float acc[N][NV];
    {
        // face fluxes at the two line ends start rows 0 and N-1
        float o0[NV], o1[NV];
        load_vars<ND, NNP>(sQ + L.base, L, o0);
        load_vars<ND, NNP>(sQ + L.base + (N - 1) * L.stride, L, o1);
        const float2 g0 = sF[lle * NFC + 2 * ldir], g1 = sF[lle * NFC + 2 * ldir + 1];
        const float ehalf = TP == TP_CHANDRASEKHAR ? 0.5f : 1.0f;  // energy row accumulated halved
        constexpr bool PO = TP == TP_CHANDRASEKHAR;
        float P[N][ND + 4];
        load_line<ND, N, NNP, PO>(sP, L, P);
        bool bad = false;
        face_init<ND, N, 0, PO>(o0, xl[0], P[0], g0.x, g0.y, gm1, gamma, a.diss, ehalf, acc, bad);
        face_init<ND, N, N - 1, PO>(o1, xl[1], P[N - 1], g1.x, g1.y, gm1, gamma, a.diss, ehalf, acc, bad);
        if (bad && line_ok) err |= ERR_SQRT_NEG;

        if constexpr (TP == TP_CHANDRASEKHAR) {
            constexpr int SB = ND + 1;
            float amax = P[0][0], amin = P[0][0], bmax = P[0][SB], bmin = P[0][SB];
#pragma unroll
            for (int p = 1; p < N; ++p) {
                amax = fmaxf(amax, P[p][0]); amin = fminf(amin, P[p][0]);
                bmax = fmaxf(bmax, P[p][SB]); bmin = fminf(bmin, P[p][SB]);
            }
            // u = ((max - min)/(max + min))² of ρ and β along the line, compared without a
            // division: (max - min)² < u0 (max + min)² (a margin of 1e-4 relative for the
            // rounding of the squares)
            const float da = amax - amin, sa = amax + amin, db = bmax - bmin, sb = bmax + bmin;
            const float qa = da * da * 1.0001f, pa = sa * sa, qb = db * db * 1.0001f, pb = sb * sb;
            if (qa < 1e-4f * pa && qb < 1e-4f * pb) line_pairs<ND, N, TP, 2>(P, acc, a.dsharp, inv_gm1);
            else if (qa < 1e-2f * pa && qb < 1e-2f * pb) line_pairs<ND, N, TP, 3>(P, acc, a.dsharp, inv_gm1);
            else line_pairs<ND, N, TP, 0>(P, acc, a.dsharp, inv_gm1);
#pragma unroll
            for (int p = 0; p < N; ++p) acc[p][NV - 1] += acc[p][NV - 1];  // energy row was halved
        } else {
            line_pairs<ND, N, TP, 0>(P, acc, a.dsharp, inv_gm1);
        }
    }

    // the stage epilogue's second operand (RK register or SSP base), loaded before the
    // barrier so its latency overlaps the other waves' line tasks -- when it is a few
    // registers per thread (3-D); with more nodes per thread (2-D) after the barriers
    const float* aux = (MODE == MODE_2N && !FIRST) ? a.reg
                     : ((MODE == MODE_SSP2 || MODE == MODE_SSP3) ? a.ubase : nullptr);
    constexpr bool EARLY_AUX = NPT * NV <= 12;
    float xv[NPT][NV];
    auto load_aux = [&]() {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const rsrc_t rx = make_rsrc(aux ? aux + v * Nd + n0 : a.Q, aux ? blk_bytes : 0u);
#pragma unroll
            for (int r = 0; r < NPT; ++r) xv[r][v] = bload(rx, (uint32_t)(tid + r * TPB) * 4u);
        }
    };
    if constexpr (EARLY_AUX) load_aux();
    // the stage input of this thread's nodes, before its LDS region takes row sums
    float uq[NPT][NV];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int j = tid + r * TPB;
#pragma unroll
        for (int v = 0; v < NV; ++v) uq[r][v] = sQ[v * NNP + pnode<N, C::PAD>(j < NN ? j : 0)];
    }
    __syncthreads();  // flux inputs and node states are dead from here
    if (line_ok) store_line<ND, N, NNP>(ldir == ND - 1 ? sQ : sP + ldir * NV * NNP, L, acc);
    if constexpr (!EARLY_AUX) load_aux();
    __syncthreads();

    // ---- per node: Σ_d (Ja_dd/J) row sums and the stage epilogue -----------------------
    // (stores of lanes past the block's last element are dropped by the range check)
    float kv[NPT][NV];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int j = tid + r * TPB;
        const int jj = j < NN ? j : 0;
        const int le = jj / NPE, pj = pnode<N, C::PAD>(jj);
        float s[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) s[d] = sG[le * 4 + d];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            float t = s[ND - 1] * sQ[lslot<ND>(ND - 1, v) * NNP + pj];
#pragma unroll
            for (int d = ND - 2; d >= 0; --d) t = fmaf(s[d], sP[(d * NV + lslot<ND>(d, v)) * NNP + pj], t);
            kv[r][v] = t;
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
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

// ======================================================================================
// Curved elements (per-node metric and J): the same block organisation, with the metric
// inside the two-point flux, F♯(i, j) = f♯(Q_i, Q_j)·(Ja^d_i + Ja^d_j)/2
// (Equations/Euler.jl:385-536, PhysicalRegions.jl:1001-1033), the numerical flux along
// the face-node normal n with the surface Jacobian (interface_fluxes!, Interfaces.jl:111-136),
// and 1/J per node (apply_massmatrix!).  Rows are kept in physical variable order.

// Elements per block: the line tasks fill most of the block's waves.
template <int ND, int N>
constexpr int fastc_epb() {
    return ND == 3 ? (N == 4 ? 4 : (N == 5 ? 3 : (N == 6 ? 2 : 1)))
                   : (N == 4 ? 16 : (N == 5 ? 32 : (N == 6 ? 16 : (N == 7 ? 9 : 8))));
}

template <int ND, int N>
struct CCfg {
    static constexpr int NV = ND + 2, NFC = 2 * ND;
    static constexpr int NPE = ipow(N, ND), NPF = ipow(N, ND - 1), NL = ND * NPF;
    static constexpr int EPB = fastc_epb<ND, N>();
    static constexpr int TPB = ((EPB * NL + 63) / 64) * 64;
    static constexpr int NN = EPB * NPE;
    static constexpr int NPT = (NN + TPB - 1) / TPB;
    // LDS slot arrays [slot][NN]: Q (NV), P (density, ND velocities, 3 scalars), M (the
    // ND contravariant metric vectors Ja^d/2, ND*ND), and 1/J.  The row sums of the ND
    // directions (ND*NV slots) go over Q, P and M once the lines are done; 1/J stays.
    static constexpr int PIN = ND + 4;
    static constexpr int SLOTS = NV + PIN + ND * ND + 1;
    static_assert(ND * NV <= SLOTS - 1, "row sums must fit below the 1/J slot");
    static constexpr int OFF_Q = 0;
    static constexpr int OFF_P = NV * NN;
    static constexpr int OFF_M = OFF_P + PIN * NN;
    static constexpr int OFF_IJ = OFF_M + ND * ND * NN;
    static constexpr int LDS = r4(SLOTS * NN);
};

// Two-point flux along n̄ = Mi + Mj (M = Ja^d/2) in physical variables.  Chandrasekhar
// returns (m, F_m.., h/2) with m = ρ̂ ū·n̄ and the energy factored, F_E = m·h (the
// energy row accumulated halved); Kennedy-Gruber and StdAverage return the flux.
template <int ND, int TP, int PATH>
__device__ __forceinline__ void twopoint_c(const float* Pi, const float* Pj, const float* Mi, const float* Mj,
                                           float inv_gm1, float* F) {
    constexpr int NV = ND + 2, SB = ND + 1;
    float n[ND], ub[ND];
#pragma unroll
    for (int k = 0; k < ND; ++k) { n[k] = Mi[k] + Mj[k]; ub[k] = Pi[1 + k] + Pj[1 + k]; }
    if constexpr (TP == TP_CHANDRASEKHAR) {
        float un = ub[0] * n[0];
#pragma unroll
        for (int k = 1; k < ND; ++k) un = fmaf(ub[k], n[k], un);
        float dot = Pi[ND] * Pj[ND];
#pragma unroll
        for (int k = ND - 1; k >= 1; --k) dot = fmaf(Pi[k], Pj[k], dot);
        const float hs = Pi[0] + Pj[0], dA = Pi[0] - Pj[0];
        const float db = Pi[SB] - Pj[SB];
        float rho, X, rb;
        if constexpr (PATH == 2) {
            // first order in u on both log-means, no reciprocal per pair (see twopoint)
            const float sh = Pi[SB + 2] + Pj[SB + 2];
            const float fh = db * sh;
            const float uh = fh * fh;
            rb = fmaf(-sh, uh, sh);
            const float fr = dA * (Pi[SB + 1] + Pj[SB + 1]);
            X = fmaf(uh, inv_gm1 * (1.0f / 6.0f), fmaf(fr * fr, 1.5f, 0.5f * (inv_gm1 + 1.0f)));
            rho = fmaf(dA, fr, hs);
        } else {
            rb = rcp(Pi[SB] + Pj[SB]);
            const float fb = db * rb, ub2 = fb * fb;
            const float fr = dA * rcp(hs);
            const float ur = fr * fr;
            float Fr, Fb;
            if constexpr (PATH == 3) {
                Fr = fmaf(ur, fmaf(ur, fmaf(ur, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
                Fb = fmaf(ub2, fmaf(ub2, fmaf(ub2, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
            } else {
                if (ur < 0.01f) Fr = fmaf(ur, fmaf(ur, fmaf(ur, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
                else Fr = 0.5f * (lg(Pi[0]) - lg(Pj[0])) * rcp(fr);
                if (ub2 < 0.01f) Fb = fmaf(ub2, fmaf(ub2, fmaf(ub2, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
                else Fb = 0.5f * (lg(Pi[SB]) - lg(Pj[SB])) * rcp(fb);
            }
            X = fmaf(Fb, 0.5f * inv_gm1, 0.5f * Fr);
            rho = hs * rcp(Fr);
        }
        const float p = hs * rb;
        const float m = rho * un;
        F[0] = m;
#pragma unroll
        for (int k = 0; k < ND; ++k) F[1 + k] = fmaf(m, ub[k], p * n[k]);
        F[NV - 1] = fmaf(rb, X, dot);
    } else if constexpr (TP == TP_KENNEDY_GRUBER) {
        float un = ub[0] * n[0];
#pragma unroll
        for (int k = 1; k < ND; ++k) un = fmaf(ub[k], n[k], un);
        const float rho = Pi[0] + Pj[0];
        const float p = Pi[SB] + Pj[SB];
        const float e = Pi[SB + 1] + Pj[SB + 1];
        const float m = rho * un;
        F[0] = m;
#pragma unroll
        for (int k = 0; k < ND; ++k) F[1 + k] = fmaf(m, ub[k], p * n[k]);
        F[NV - 1] = fmaf(m, e, p * un);
    } else {
        // ½(f(Q_i) + f(Q_j))·n̄ with full velocities; Pl = (ρ, u.., p, ρe)
        float uni = Pi[1] * n[0], unj = Pj[1] * n[0];
#pragma unroll
        for (int k = 1; k < ND; ++k) { uni = fmaf(Pi[1 + k], n[k], uni); unj = fmaf(Pj[1 + k], n[k], unj); }
        const float mi = Pi[0] * uni, mj = Pj[0] * unj;
        F[0] = 0.5f * (mi + mj);
#pragma unroll
        for (int k = 0; k < ND; ++k)
            F[1 + k] = 0.5f * (fmaf(mi, Pi[1 + k], Pi[SB] * n[k]) + fmaf(mj, Pj[1 + k], Pj[SB] * n[k]));
        F[NV - 1] = 0.5f * ((Pi[SB + 1] + Pi[SB]) * uni + (Pj[SB + 1] + Pj[SB]) * unj);
    }
}

template <int ND, int N, int TP, int PATH>
__device__ __forceinline__ void line_pairs_c(const float (&P)[N][ND + 4], const float (&M)[N][ND],
                                             float (&acc)[N][ND + 2], const float* __restrict__ Dsh, float inv_gm1) {
    constexpr int NV = ND + 2;
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = i + 1; j < N; ++j) {
            float F[NV];
            twopoint_c<ND, TP, PATH>(P[i], P[j], M[i], M[j], inv_gm1, F);
            const float cij = Dsh[i * N + j], cji = Dsh[j * N + i];
            const bool setj = i == 0 && j < N - 1;  // first contribution to row j
            if constexpr (TP == TP_CHANDRASEKHAR) {
                const float mi = cij * F[0], mj = cji * F[0];
                acc[i][0] -= mi;
                acc[j][0] = setj ? -mj : acc[j][0] - mj;
#pragma unroll
                for (int v = 1; v <= ND; ++v) {
                    acc[i][v] = fmaf(-cij, F[v], acc[i][v]);
                    acc[j][v] = setj ? -cji * F[v] : fmaf(-cji, F[v], acc[j][v]);
                }
                acc[i][NV - 1] = fmaf(-mi, F[NV - 1], acc[i][NV - 1]);
                acc[j][NV - 1] = setj ? -mj * F[NV - 1] : fmaf(-mj, F[NV - 1], acc[j][NV - 1]);
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    acc[i][v] = fmaf(-cij, F[v], acc[i][v]);
                    acc[j][v] = setj ? -cji * F[v] : fmaf(-cji, F[v], acc[j][v]);
                }
            }
        }
    }
}

// Numerical flux at a line end in physical variables, as the own element sees it:
// acc[R] = -coef·jac·[(f(Q_own) + f(Q_ext))·n_own + λ·diss·(Q_own - Q_ext)], with n_own the
// own outward unit normal (the master side's normal, negated on the slave side) and
// λ = max(|u·n| + c).  Both elements of a face evaluate the same sums with opposite n, so
// their values are exactly opposite.  The energy entry is scaled by ehalf.
template <int ND, int N, int R>
__device__ __forceinline__ void face_init_c(const float (&o)[ND + 2], const float (&x)[ND + 2], const float* n,
                                            float coef, float gm1, float gamma, float diss, float ehalf,
                                            float (&acc)[N][ND + 2], bool& bad) {
    constexpr int NV = ND + 2;
    float fs[NV], a2[2];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const float* Q = side ? x : o;
        const float ir = rcp(Q[0]);
        float m2 = Q[ND] * Q[ND], mn = Q[ND] * n[ND - 1];
#pragma unroll
        for (int k = ND - 1; k >= 1; --k) { m2 = fmaf(Q[k], Q[k], m2); mn = fmaf(Q[k], n[k - 1], mn); }
        const float p = gm1 * fmaf(-0.5f * m2, ir, Q[NV - 1]);
        const float un = mn * ir;
        const float c2 = gamma * p * ir;
        bad |= !(c2 >= 0.0f);
        a2[side] = fabsf(un) + sq(c2);
        if (side == 0) {
            fs[0] = mn;
#pragma unroll
            for (int k = 1; k <= ND; ++k) fs[k] = fmaf(Q[k], un, p * n[k - 1]);
            fs[NV - 1] = (Q[NV - 1] + p) * un;
        } else {
            fs[0] += mn;
#pragma unroll
            for (int k = 1; k <= ND; ++k) fs[k] += fmaf(Q[k], un, p * n[k - 1]);
            fs[NV - 1] += (Q[NV - 1] + p) * un;
        }
    }
    const float cl = coef * fmaxf(a2[0], a2[1]) * diss;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const float t = fmaf(-coef, fs[v], -cl * (o[v] - x[v]));
        acc[R][v] = v == NV - 1 ? ehalf * t : t;
    }
}

template <int ND, int N, int TP, int MODE, bool FIRST>
__global__ void __launch_bounds__((CCfg<ND, N>::TPB)) k_fastc(const FastArgs a) {
    using C = CCfg<ND, N>;
    constexpr int NV = C::NV, NFC = C::NFC, NPE = C::NPE, NPF = C::NPF, NL = C::NL, EPB = C::EPB;
    constexpr int TPB = C::TPB, NN = C::NN, NPT = C::NPT;
    __shared__ float4 smem[(C::LDS + 3) / 4];
    float* const sm = reinterpret_cast<float*>(smem);
    float* const sQ = sm + C::OFF_Q;
    float* const sP = sm + C::OFF_P;
    float* const sM = sm + C::OFF_M;
    float* const sIJ = sm + C::OFF_IJ;

    const int tid = threadIdx.x;
    const int e0 = xcd_remap(blockIdx.x, gridDim.x) * EPB;
    const int64_t Nd = a.ndofs;
    const float gamma = a.gamma, gm1 = a.gamma - 1.0f, inv_gm1 = a.inv_gm1;
    const int nel = a.ne - e0 < EPB ? a.ne - e0 : EPB;
    const uint32_t n0 = (uint32_t)e0 * NPE;
    int err = 0;

    // ---- this thread's line task (lanes past the tasks repeat a valid line) ----------
    // Direction-major: consecutive lanes run the lines of one direction over the block's
    // elements, so most waves read LDS with one stride (C3: -2 %, C2: -5 % per stage
    // against element-major lanes, whose waves mixed two directions' strides).  When one
    // direction's lines fit in a wave and the block has one wave per direction (3-D N=7,
    // 2-D N=7), wave d runs direction d and its spare lanes repeat lines of it (C5: -2.5 %).
    constexpr bool WDIR = EPB * NPF <= 64 && TPB == ND * 64;
    static_assert(!WDIR || 2 * EPB * NPF >= 64, "spare lanes repeat lines of their wave's direction");
    // (wave-uniform wave index: the line direction of a WDIR wave lives in a scalar register)
    const int wk = tid & 63, wd = WDIR ? __builtin_amdgcn_readfirstlane(tid >> 6) : 0;
    const bool task = WDIR ? wk < EPB * NPF : tid < EPB * NL;
    const int tl = WDIR ? wd * (EPB * NPF) + (task ? wk : wk - EPB * NPF) : (task ? tid : tid - EPB * NL);
    const int ldir = WDIR ? wd : tl / (EPB * NPF), lr = tl - ldir * (EPB * NPF);
    const int lle = lr / NPF, lf = lr - lle * NPF;
    const bool line_in = lle < nel;
    const bool line_ok = task && line_in;
    int xnb[2], xinf[2], xfg[2];
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const int slot = (e0 + lle) * NFC + 2 * ldir + sd;
        xnb[sd] = line_ok ? a.nb[slot] : -2;
        xinf[sd] = line_ok ? a.finfo[slot] : 0;
        xfg[sd] = line_ok ? a.fgi[slot] : 0;
    }

    // ---- staging: node states, per-node metric and J -----------------------------------
    const uint32_t blk_bytes = (uint32_t)(nel * NPE) * 4u;
    float q[NPT][NV], g[NPT][ND * ND + 1];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const rsrc_t rq = make_rsrc(a.Q + v * Nd + n0, blk_bytes);
#pragma unroll
        for (int r = 0; r < NPT; ++r) q[r][v] = bload(rq, (uint32_t)(tid + r * TPB) * 4u);
    }
#pragma unroll
    for (int c = 0; c < ND * ND + 1; ++c) {
        const rsrc_t rg = make_rsrc(a.ngeom + c * Nd + n0, blk_bytes);
#pragma unroll
        for (int r = 0; r < NPT; ++r) g[r][c] = bload(rg, (uint32_t)(tid + r * TPB) * 4u);
    }
    // face normals and surface Jacobians of both line ends (per face node, master side),
    // one 16-byte load each
    float fn[2][ND], fj[2];
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const float4 q = line_ok ? a.fnj[(int64_t)xfg[sd] * NPF + lf] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        fn[sd][0] = q.x;
        if (ND > 1) fn[sd][1] = q.y;
        if (ND > 2) fn[sd][ND - 1] = q.z;
        fj[sd] = q.w;
    }
    // exterior traces
    uint32_t xo[2];
    bool inblk[2];
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const int kp = finfo_nbpos(xinf[sd]);
        inblk[sd] = xnb[sd] >= e0 && xnb[sd] < e0 + nel;
        xo[sd] = xnb[sd] >= 0 && !inblk[sd]
                     ? ((uint32_t)xnb[sd] * NPE + line_node<ND, N>(kp >> 1, lf, (kp & 1) ? N - 1 : 0)) * 4u
                     : 0xffffffffu;
    }
    float xr[2][NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const rsrc_t rv = make_rsrc(a.Q + v * Nd, (uint32_t)Nd * 4u);
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) xr[sd][v] = bload(rv, xo[sd]);
    }
    // flux inputs, metric and 1/J to LDS
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int j = tid + r * TPB;
        if (j >= NN) continue;
        const float rho = q[r][0];
        const float ir = rcp(rho);
        float u[ND];
        const float uscale = TP == TP_STDAVG ? ir : 0.5f * ir;
#pragma unroll
        for (int k = 0; k < ND; ++k) u[k] = q[r][1 + k] * uscale;
        float uu = u[ND - 1] * u[ND - 1];
#pragma unroll
        for (int k = ND - 2; k >= 0; --k) uu = fmaf(u[k], u[k], uu);
        const float p = gm1 * fmaf((TP == TP_STDAVG ? -0.5f : -2.0f) * rho, uu, q[r][NV - 1]);
        if (j < nel * NPE && (!(rho > 0.0f) || !(p > 0.0f) || !(g[r][ND * ND] > 0.0f))) err |= ERR_LOG_NEG;
#pragma unroll
        for (int c = 0; c < NV; ++c) sQ[c * NN + j] = q[r][c];
        sP[j] = TP == TP_STDAVG ? rho : 0.5f * rho;
#pragma unroll
        for (int k = 0; k < ND; ++k) sP[(1 + k) * NN + j] = u[k];
        if constexpr (TP == TP_CHANDRASEKHAR) {
            sP[(ND + 1) * NN + j] = 0.5f * rho * rcp(p);
            sP[(ND + 2) * NN + j] = (-1.0f / 6.0f) * ir;
            sP[(ND + 3) * NN + j] = p * (0.5f * ir);
        } else if constexpr (TP == TP_KENNEDY_GRUBER) {
            sP[(ND + 1) * NN + j] = 0.5f * p;
            sP[(ND + 2) * NN + j] = 0.5f * q[r][NV - 1] * ir;
        } else {
            sP[(ND + 1) * NN + j] = p;
            sP[(ND + 2) * NN + j] = q[r][NV - 1];
        }
        // metric: column d of the node's matrix is Ja^d; stored halved per direction
#pragma unroll
        for (int c = 0; c < ND * ND; ++c) sM[c * NN + j] = 0.5f * g[r][c];
        sIJ[j] = rcp(g[r][ND * ND]);
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
                src = a.bcstate + ((int64_t)a.fid[(e0 + lle) * NFC + k] * NPF + lf) * NV;
            } else if (type == BC_INFLOW) {
                src = a.bcconst + bd * NV;
            } else {
                src = a.Q + n0 + lle * NPE + line_node<ND, N>(ldir, lf, sd ? N - 1 : 0);
                vs = Nd;
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) xr[sd][v] = src[v * vs];
        }
    }
    __syncthreads();
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        if (inblk[sd]) {
            const int kp = finfo_nbpos(xinf[sd]);
            const int node = (xnb[sd] - e0) * NPE + line_node<ND, N>(kp >> 1, lf, (kp & 1) ? N - 1 : 0);
#pragma unroll
            for (int v = 0; v < NV; ++v) xr[sd][v] = sQ[v * NN + node];
        }
    }
    // a slip wall reflects the normal momentum: m - 2(m·n)n
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        if (xnb[sd] == -1 && a.bctype[finfo_bd(xinf[sd])] == BC_SLIP) {
            float mn = xr[sd][1] * fn[sd][0];
#pragma unroll
            for (int k = 1; k < ND; ++k) mn = fmaf(xr[sd][1 + k], fn[sd][k], mn);
#pragma unroll
            for (int k = 0; k < ND; ++k) xr[sd][1 + k] = fmaf(-2.0f * mn, fn[sd][k], xr[sd][1 + k]);
        }
    }

    // ---- line tasks --------------------------------------------------------------------
    const int base = lle * NPE + line_node<ND, N>(ldir, lf, 0);
    const int stride = ldir == 0 ? 1 : (ldir == 1 ? N : NPF);
    float acc[N][NV];
    {
        float P[N][ND + 4], M[N][ND];
#pragma unroll
        for (int p = 0; p < N; ++p) {
            const int nd = base + p * stride;
#pragma unroll
            for (int c = 0; c < ND + 3; ++c) P[p][c] = sP[c * NN + nd];
            P[p][ND + 3] = TP == TP_CHANDRASEKHAR ? sP[(ND + 3) * NN + nd] : 0.0f;
#pragma unroll
            for (int k = 0; k < ND; ++k) M[p][k] = sM[(k + ND * ldir) * NN + nd];
        }
        float o0[NV], o1[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            o0[v] = sQ[v * NN + base];
            o1[v] = sQ[v * NN + base + (N - 1) * stride];
        }
        const float ehalf = TP == TP_CHANDRASEKHAR ? 0.5f : 1.0f;
        bool bad = false;
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {
            // own outward normal: the master's normal, negated on the slave side
            const float s = finfo_side(xinf[sd]) ? -1.0f : 1.0f;
            float nown[ND];
#pragma unroll
            for (int k = 0; k < ND; ++k) nown[k] = s * fn[sd][k];
            const float coef = 0.5f * fj[sd] * (sd ? a.lift_p : a.lift_m);
            if (sd == 0) face_init_c<ND, N, 0>(o0, xr[0], nown, coef, gm1, gamma, a.diss, ehalf, acc, bad);
            else face_init_c<ND, N, N - 1>(o1, xr[1], nown, coef, gm1, gamma, a.diss, ehalf, acc, bad);
        }
        if (bad && line_ok) err |= ERR_SQRT_NEG;
        if constexpr (TP == TP_CHANDRASEKHAR) {
            constexpr int SB = ND + 1;
            float amax = P[0][0], amin = P[0][0], bmax = P[0][SB], bmin = P[0][SB];
#pragma unroll
            for (int p = 1; p < N; ++p) {
                amax = fmaxf(amax, P[p][0]); amin = fminf(amin, P[p][0]);
                bmax = fmaxf(bmax, P[p][SB]); bmin = fminf(bmin, P[p][SB]);
            }
            const float da = amax - amin, sa = amax + amin, db = bmax - bmin, sb = bmax + bmin;
            const float qa = da * da * 1.0001f, pa = sa * sa, qb = db * db * 1.0001f, pb = sb * sb;
            if (qa < 1e-4f * pa && qb < 1e-4f * pb) line_pairs_c<ND, N, TP, 2>(P, M, acc, a.dsharp, inv_gm1);
            else if (qa < 1e-2f * pa && qb < 1e-2f * pb) line_pairs_c<ND, N, TP, 3>(P, M, acc, a.dsharp, inv_gm1);
            else line_pairs_c<ND, N, TP, 0>(P, M, acc, a.dsharp, inv_gm1);
#pragma unroll
            for (int p = 0; p < N; ++p) acc[p][NV - 1] += acc[p][NV - 1];  // energy row was halved
        } else {
            line_pairs_c<ND, N, TP, 0>(P, M, acc, a.dsharp, inv_gm1);
        }
    }

    const float* aux = (MODE == MODE_2N && !FIRST) ? a.reg
                     : ((MODE == MODE_SSP2 || MODE == MODE_SSP3) ? a.ubase : nullptr);
    float xv[NPT][NV], uq[NPT][NV];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int j = tid + r * TPB;
#pragma unroll
        for (int v = 0; v < NV; ++v) uq[r][v] = sQ[v * NN + (j < NN ? j : 0)];
    }
    __syncthreads();  // inputs are dead from here; 1/J stays
    if (line_ok) {
        // row sums of direction d in slots [d*NV, d*NV + NV)
        float* R = sm + ldir * NV * NN;
#pragma unroll
        for (int p = 0; p < N; ++p)
#pragma unroll
            for (int c = 0; c < NV; ++c) R[c * NN + base + p * stride] = acc[p][c];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const rsrc_t rx = make_rsrc(aux ? aux + v * Nd + n0 : a.Q, aux ? blk_bytes : 0u);
#pragma unroll
        for (int r = 0; r < NPT; ++r) xv[r][v] = bload(rx, (uint32_t)(tid + r * TPB) * 4u);
    }
    __syncthreads();

    // ---- per node: (Σ_d row sums)/J and the stage epilogue ----------------------------
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const rsrc_t ro = make_rsrc(a.out + v * Nd + n0, blk_bytes);
        const rsrc_t rr = make_rsrc(a.reg + v * Nd + n0, MODE == MODE_2N ? blk_bytes : 0u);
#pragma unroll
        for (int r = 0; r < NPT; ++r) {
            const int j = tid + r * TPB;
            const int jj = j < NN ? j : 0;
            float t = sm[v * NN + jj];
#pragma unroll
            for (int d = 1; d < ND; ++d) t += sm[(d * NV + v) * NN + jj];
            const float k = t * sIJ[jj];
            const uint32_t off = (uint32_t)j * 4u;
            const float uin = uq[r][v], xx = xv[r][v];
            float out;
            if constexpr (MODE == MODE_RHS) {
                out = k;
            } else if constexpr (MODE == MODE_2N) {
                const float tt = FIRST ? a.dt * k : fmaf(a.A, xx, a.dt * k);
                bstore(rr, off, tt);
                out = fmaf(a.B, tt, uin);
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

template <int ND, int N, int TP>
static hipError_t launch_c(const FastArgs& a, hipStream_t s) {
    using C = CCfg<ND, N>;
    const int nblocks = (a.ne + C::EPB - 1) / C::EPB;
    const dim3 g(nblocks), b(C::TPB);
    switch (a.mode) {
        case MODE_RHS: hipLaunchKernelGGL((k_fastc<ND, N, TP, MODE_RHS, false>), g, b, 0, s, a); break;
        case MODE_2N:
            if (a.first) hipLaunchKernelGGL((k_fastc<ND, N, TP, MODE_2N, true>), g, b, 0, s, a);
            else hipLaunchKernelGGL((k_fastc<ND, N, TP, MODE_2N, false>), g, b, 0, s, a);
            break;
        case MODE_SSP1: hipLaunchKernelGGL((k_fastc<ND, N, TP, MODE_SSP1, false>), g, b, 0, s, a); break;
        case MODE_SSP2: hipLaunchKernelGGL((k_fastc<ND, N, TP, MODE_SSP2, false>), g, b, 0, s, a); break;
        case MODE_SSP3: hipLaunchKernelGGL((k_fastc<ND, N, TP, MODE_SSP3, false>), g, b, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int ND, int N, int TP>
static hipError_t launch(const FastArgs& a, hipStream_t s) {
    using C = FCfg<ND, N>;
    const int nblocks = (a.ne + C::EPB - 1) / C::EPB;
    const dim3 g(nblocks), b(C::TPB);
    switch (a.mode) {
        case MODE_RHS: hipLaunchKernelGGL((k_fast<ND, N, TP, MODE_RHS, false>), g, b, 0, s, a); break;
        case MODE_2N:
            if (a.first) hipLaunchKernelGGL((k_fast<ND, N, TP, MODE_2N, true>), g, b, 0, s, a);
            else hipLaunchKernelGGL((k_fast<ND, N, TP, MODE_2N, false>), g, b, 0, s, a);
            break;
        case MODE_SSP1: hipLaunchKernelGGL((k_fast<ND, N, TP, MODE_SSP1, false>), g, b, 0, s, a); break;
        case MODE_SSP2: hipLaunchKernelGGL((k_fast<ND, N, TP, MODE_SSP2, false>), g, b, 0, s, a); break;
        case MODE_SSP3: hipLaunchKernelGGL((k_fast<ND, N, TP, MODE_SSP3, false>), g, b, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int ND, int N>
static FastLauncher pick_tp(int tp) {
    switch (tp) {
        case TP_CHANDRASEKHAR: return &launch<ND, N, TP_CHANDRASEKHAR>;
        case TP_KENNEDY_GRUBER: return &launch<ND, N, TP_KENNEDY_GRUBER>;
        case TP_STDAVG: return &launch<ND, N, TP_STDAVG>;
        default: return nullptr;
    }
}

template <int ND, int N, bool CURV>
static FastLauncher pick_tp_any(int tp) {
    if constexpr (!CURV) {
        return pick_tp<ND, N>(tp);
    } else {
        switch (tp) {
            case TP_CHANDRASEKHAR: return &launch_c<ND, N, TP_CHANDRASEKHAR>;
            case TP_KENNEDY_GRUBER: return &launch_c<ND, N, TP_KENNEDY_GRUBER>;
            case TP_STDAVG: return &launch_c<ND, N, TP_STDAVG>;
            default: return nullptr;
        }
    }
}

template <int ND, bool CURV>
static FastLauncher pick_n(int n, int tp) {
    switch (n) {
        case 4: return pick_tp_any<ND, 4, CURV>(tp);
        case 5: return pick_tp_any<ND, 5, CURV>(tp);
        case 6: return pick_tp_any<ND, 6, CURV>(tp);
        case 7: return pick_tp_any<ND, 7, CURV>(tp);
        case 8: return pick_tp_any<ND, 8, CURV>(tp);
        default: return nullptr;
    }
}

}  // namespace fast

FastLauncher fast_launcher(int nd, int n, int tp, bool curved) {
    if (nd == 2) return curved ? fast::pick_n<2, true>(n, tp) : fast::pick_n<2, false>(n, tp);
    if (nd == 3) return curved ? fast::pick_n<3, true>(n, tp) : fast::pick_n<3, false>(n, tp);
    return nullptr;
}

}  // namespace flouhip
