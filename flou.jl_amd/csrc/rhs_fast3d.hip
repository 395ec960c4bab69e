// FP32 ("chosen form") fused element kernel for the 3-D split-form Euler rhs on
// axis-aligned affine hexahedra (Flou's CartesianMesh), GLL nodes, LxF(StdAverage) or
// StdAverage surface flux, with the RK stage epilogue.
//
// Same mathematics as SplitDivOperator + interface_fluxes! + the FR lift
// (OpDivergence.jl:200-299, Interfaces.jl:111-136), organised for gfx950:
//   * EPB elements per block; the nodal values of all variables are staged in LDS with
//     coalesced loads and the two-point flux inputs are formed once per node;
//   * one thread per (element, direction, line) -- every thread does the same work:
//       - all N(N-1)/2 pairs of its line, each evaluated once (F♯ is symmetric), both rows
//         accumulated in registers (F♯ never touches memory);
//       - the numerical flux on the two face nodes where the line ends (Rusanov, master
//         side as left state, so both elements of a face get exactly opposite values) and
//         the FR lift of those fluxes into the two end rows (GLL: ∂g lives on the ends);
//     then writes its row sums to LDS;
//   * per node: Σ_d (Ja_dd/J)·(row sum of direction d) and the RK stage epilogue.
// FP32 reformulations (the FP64 plain kernel keeps the reference's operation order):
//   - node inputs are pre-halved (ρ/2, u/2) so a mean is one add;
//   - log-mean ρ̂ = (a+b)/(2F), F = atanh(f)/f, f = (a-b)/(a+b), u = f²
//     (Utilities.jl:34-44).  Each line picks a branch-free path from the spread of ρ and
//     β along it: u < 1e-4: F = 1 + u/3, 1/F = 1 - u/3 (truncation < 2e-9);
//     u < 1e-2: the reference's three-term series; otherwise per pair the series or
//     F = (ln a - ln b)/(2f);
//   - 1/(2β̂) = F_β/(β1+β2) and p̂/ρ̂ = F_ρ/(β1+β2) (algebraically identical);
//   - the diagonal metric and 1/J are applied once per node, the lift coefficients
//     ∂g/Ja_dd are folded so the lift can be added to the direction's row sums.
#include "common.hpp"
#include "fast.hpp"

namespace flouhip {
namespace fast3d {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float lg(float x) { return __logf(x); }
__device__ __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); }

template <int N>
struct FCfg {
    static constexpr int NPE = N * N * N;
    static constexpr int NPF = N * N;
    static constexpr int NL = 3 * N * N;                        // line tasks per element
    static constexpr int EPB = N <= 6 ? 2 : 1;                  // elements per block
    static constexpr int TPB = ((EPB * NL + 63) / 64) * 64;     // threads per block
};

// Node inputs (8 floats): Chandrasekhar: ρ/2, u/2, v/2, w/2, β, |u|²/4, -, -
//                         Kennedy-Gruber: ρ/2, u/2, v/2, w/2, p/2, e/2, -, -
//                         StdAverage:     ρ, u, v, w, p, ρe, -, -
// In a line task slot 1 holds the velocity along the line, slots 2, 3 the other two.
template <int TP, int PATH>
__device__ __forceinline__ void twopoint(const float* Pi, const float* Pj, float inv_gm1, float* F) {
    if constexpr (TP == TP_CHANDRASEKHAR) {
        const float hs = Pi[0] + Pj[0], dA = Pi[0] - Pj[0];
        const float sb = Pi[4] + Pj[4], db = Pi[4] - Pj[4];
        const float rs = rcp(hs), rb = rcp(sb);
        const float fr = dA * rs, fb = db * rb;
        const float ur = fr * fr, ub = fb * fb;
        float Fr, Fb, iFr;
        if constexpr (PATH == 1) {
            Fr = fmaf(ur, 1.0f / 3.0f, 1.0f);
            Fb = fmaf(ub, 1.0f / 3.0f, 1.0f);
            iFr = fmaf(ur, -1.0f / 3.0f, 1.0f);
        } else if constexpr (PATH == 3) {
            Fr = fmaf(ur, fmaf(ur, fmaf(ur, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
            Fb = fmaf(ub, fmaf(ub, fmaf(ub, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
            iFr = rcp(Fr);
        } else {
            if (ur < 0.01f) Fr = fmaf(ur, fmaf(ur, fmaf(ur, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
            else Fr = 0.5f * (lg(Pi[0]) - lg(Pj[0])) * rcp(fr);
            if (ub < 0.01f) Fb = fmaf(ub, fmaf(ub, fmaf(ub, 1.0f / 7.0f, 1.0f / 5.0f), 1.0f / 3.0f), 1.0f);
            else Fb = 0.5f * (lg(Pi[4]) - lg(Pj[4])) * rcp(fb);
            iFr = rcp(Fr);
        }
        const float rho = hs * iFr;
        const float p = hs * rb;
        const float un = Pi[1] + Pj[1];
        const float ut1 = Pi[2] + Pj[2];
        const float ut2 = Pi[3] + Pj[3];
        const float u2 = fmaf(un, un, fmaf(ut1, ut1, ut2 * ut2));
        const float h = fmaf(rb, fmaf(Fb, inv_gm1, Fr), u2 - (Pi[5] + Pj[5]));
        const float m = rho * un;
        F[0] = m;
        F[1] = fmaf(m, un, p);
        F[2] = m * ut1;
        F[3] = m * ut2;
        F[4] = m * h;
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
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                acc[i][v] = fmaf(-cij, F[v], acc[i][v]);
                acc[j][v] = fmaf(-cji, F[v], acc[j][v]);
            }
        }
    }
}

// Rusanov flux through the face normal to axis d between QL (master side) and QR, along
// the master's normal s·e_d; returns F·n (not yet scaled by the face Jacobian).
__device__ __forceinline__ void rusanov_axis(const float* QL, const float* QR, int d, float s,
                                             float gm1, float gamma, float diss, float* F, int* err) {
    float fL[5], fR[5], uL, uR, cL, cR;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const float* Q = side ? QR : QL;
        float* f = side ? fR : fL;
        const float ir = rcp(Q[0]);
        const float m2 = fmaf(Q[1], Q[1], fmaf(Q[2], Q[2], Q[3] * Q[3]));
        const float p = gm1 * fmaf(-0.5f * m2, ir, Q[4]);
        const float md = d == 0 ? Q[1] : (d == 1 ? Q[2] : Q[3]);
        const float un = md * ir;
        const float c2 = gamma * p * ir;
        if (!(c2 >= 0.0f)) atomicOr(err, ERR_SQRT_NEG);
        const float c = sq(c2);
        f[0] = md;
        f[1] = fmaf(Q[1], un, d == 0 ? p : 0.0f);
        f[2] = fmaf(Q[2], un, d == 1 ? p : 0.0f);
        f[3] = fmaf(Q[3], un, d == 2 ? p : 0.0f);
        f[4] = (Q[4] + p) * un;
        if (side) { uR = un; cR = c; } else { uL = un; cL = c; }
    }
    const float lam = fmaxf(fabsf(uL) + cL, fabsf(uR) + cR) * diss;
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = 0.5f * fmaf(s, fL[v] + fR[v], lam * (QL[v] - QR[v]));
}

// Where the exterior trace of a face node comes from: a neighbour node, a Dirichlet or
// inflow state, or (outflow / slip walls) the element's own node.  Returns base pointer and
// stride so that the five loads are issued without branches around them.
template <int N>
__device__ __forceinline__ const float* exterior_src(const FastArgs& a, int e, int k, int f,
                                                     int nbe, int info, int64_t& stride) {
    constexpr int NPE = N * N * N, NPF = N * N;
    if (nbe >= 0) {
        const int kp = finfo_nbpos(info);
        stride = a.ndofs;
        return a.Q + (int64_t)nbe * NPE + line_node<3, N>(kp >> 1, f, (kp & 1) ? N - 1 : 0);
    }
    const int bd = finfo_bd(info);
    const int type = a.bctype[bd];
    if (type == BC_DIRICHLET) {
        stride = 1;
        return a.bcstate + ((int64_t)a.fid[e * 6 + k] * NPF + f) * 5;
    }
    if (type == BC_INFLOW) {
        stride = 1;
        return a.bcconst + bd * 5;
    }
    stride = a.ndofs;  // outflow / slip: the own trace (reflected below for slip)
    return a.Q + (int64_t)e * NPE + line_node<3, N>(k >> 1, f, (k & 1) ? N - 1 : 0);
}

// Numerical flux on the face where a line ends (row R of the line) and its lift into that
// row: acc[R] -= (∂g / Ja_dd) · (±surface J) · F, F in the line's frame.
template <int N, int R>
__device__ __forceinline__ void face_lift(const float (&own)[5], float (&ext)[5], int info, bool slip,
                                          float s, float jac, int dir, float coef, float gm1,
                                          float gamma, float diss, float (&acc)[N][5], int* err) {
    if (slip) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (c == dir) ext[1 + c] = -ext[1 + c];
    }
    const bool slave = finfo_side(info);
    float QL[5], QR[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        QL[v] = slave ? ext[v] : own[v];
        QR[v] = slave ? own[v] : ext[v];
    }
    float F[5];
    rusanov_axis(QL, QR, dir, s, gm1, gamma, diss, F, err);
    const float m1 = dir == 0 ? F[1] : (dir == 1 ? F[2] : F[3]);
    const float m2 = dir == 0 ? F[2] : (dir == 1 ? F[3] : F[1]);
    const float m3 = dir == 0 ? F[3] : (dir == 1 ? F[1] : F[2]);
    F[1] = m1;
    F[2] = m2;
    F[3] = m3;
    const float c = coef * (slave ? -jac : jac);
#pragma unroll
    for (int v = 0; v < 5; ++v) acc[R][v] = fmaf(-c, F[v], acc[R][v]);
}

template <int N, int TP>
__global__ void __launch_bounds__((FCfg<N>::TPB)) k_fast3d(const FastArgs a) {
    using C = FCfg<N>;
    constexpr int NPE = C::NPE, NPF = C::NPF, NL = C::NL, EPB = C::EPB, TPB = C::TPB;
    constexpr int NQ = 5 * EPB * NPE;              // staged values per block
    constexpr int QPT = (NQ + TPB - 1) / TPB;      // staged values per thread
    constexpr int NPT = (EPB * NPE + TPB - 1) / TPB;  // nodes per thread (epilogue)
    __shared__ float sQ[EPB][5][NPE];
    __shared__ float4 sP[EPB][NPE][2];
    __shared__ float sPart[EPB][3][5][NPE];
    __shared__ float sG[EPB][8];

    const int tid = threadIdx.x;
    const int e0 = xcd_remap(blockIdx.x, gridDim.x) * EPB;
    const int64_t Nd = a.ndofs;
    const float gamma = a.gamma, gm1 = a.gamma - 1.0f, inv_gm1 = 1.0f / (a.gamma - 1.0f);
    const int nel = a.ne - e0 < EPB ? a.ne - e0 : EPB;

    // line task of this thread and its face-table entries (issued first: they feed the
    // neighbour-trace addresses)
    const bool is_line = tid < EPB * NL;
    const int lle = tid / NL, lt = tid - (tid / NL) * NL;
    const int ldir = lt / NPF, lf = lt - (lt / NPF) * NPF;
    const bool line_ok = is_line && lle < nel;
    const int le_e = e0 + lle;
    int info0 = 0, info1 = 0, nb0 = -1, nb1 = -1, gi0 = 0, gi1 = 0;
    if (line_ok) {
        const int s0 = le_e * 6 + 2 * ldir;
        info0 = a.finfo[s0]; info1 = a.finfo[s0 + 1];
        nb0 = a.nb[s0]; nb1 = a.nb[s0 + 1];
        gi0 = a.fgi[s0]; gi1 = a.fgi[s0 + 1];
    }

    // stage the element block: all loads in flight before the first LDS write
    {
        float buf[QPT];
#pragma unroll
        for (int r = 0; r < QPT; ++r) {
            const int i = tid + r * TPB;
            const int v = i / (EPB * NPE), j = i - v * (EPB * NPE);
            buf[r] = (i < NQ && j / NPE < nel) ? a.Q[v * Nd + (int64_t)e0 * NPE + j] : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < QPT; ++r) {
            const int i = tid + r * TPB;
            if (i < NQ) {
                const int v = i / (EPB * NPE), j = i - v * (EPB * NPE);
                sQ[j / NPE][v][j % NPE] = buf[r];
            }
        }
        if (tid < 8 * EPB) {
            const int le = tid >> 3;
            if (le < nel) sG[le][tid & 7] = a.egeom[(int64_t)(e0 + le) * a.egeom_stride + (tid & 7)];
        }
    }
    __syncthreads();

    // per-node inputs of the two-point flux
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int i = tid + r * TPB;
        const int le = i / NPE, nd = i - le * NPE;
        if (i >= EPB * NPE || le >= nel) continue;
        const float rho = sQ[le][0][nd];
        const float ir = rcp(rho);
        const float u = sQ[le][1][nd] * ir, v = sQ[le][2][nd] * ir, w = sQ[le][3][nd] * ir;
        const float q = fmaf(u, u, fmaf(v, v, w * w));
        const float p = gm1 * fmaf(-0.5f * rho, q, sQ[le][4][nd]);
        if (!(rho > 0.0f) || !(p > 0.0f)) atomicOr(a.errflag, ERR_LOG_NEG);
        float4 p0, p1;
        if constexpr (TP == TP_CHANDRASEKHAR) {
            p0 = make_float4(0.5f * rho, 0.5f * u, 0.5f * v, 0.5f * w);
            p1 = make_float4(0.5f * rho * rcp(p), 0.25f * q, 0.0f, 0.0f);
        } else if constexpr (TP == TP_KENNEDY_GRUBER) {
            p0 = make_float4(0.5f * rho, 0.5f * u, 0.5f * v, 0.5f * w);
            p1 = make_float4(0.5f * p, 0.5f * sQ[le][4][nd] * ir, 0.0f, 0.0f);
        } else {
            p0 = make_float4(rho, u, v, w);
            p1 = make_float4(p, sQ[le][4][nd], 0.0f, 0.0f);
        }
        sP[le][nd][0] = p0;
        sP[le][nd][1] = p1;
    }
    __syncthreads();

    if (line_ok) {
        const int le = lle, dir = ldir, f = lf, e = le_e;
        // neighbour traces of the two faces this line ends on, loads issued before the pairs
        float ext0[5], ext1[5];
        bool slip0 = false, slip1 = false;
        {
            int64_t st0, st1;
            const float* src0 = exterior_src<N>(a, e, 2 * dir, f, nb0, info0, st0);
            const float* src1 = exterior_src<N>(a, e, 2 * dir + 1, f, nb1, info1, st1);
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                ext0[v] = src0[v * st0];
                ext1[v] = src1[v * st1];
            }
            if (nb0 < 0) slip0 = a.bctype[finfo_bd(info0)] == BC_SLIP;
            if (nb1 < 0) slip1 = a.bctype[finfo_bd(info1)] == BC_SLIP;
        }
        const float sgn0 = a.fgeom[(int64_t)gi0 * 10 + dir], jac0 = a.fgeom[(int64_t)gi0 * 10 + 9];
        const float sgn1 = a.fgeom[(int64_t)gi1 * 10 + dir], jac1 = a.fgeom[(int64_t)gi1 * 10 + 9];

        float P[N][6];
#pragma unroll
        for (int p = 0; p < N; ++p) {
            const int nd = line_node<3, N>(dir, f, p);
            const float4 q0 = sP[le][nd][0], q1 = sP[le][nd][1];
            P[p][0] = q0.x;
            P[p][1] = dir == 0 ? q0.y : (dir == 1 ? q0.z : q0.w);
            P[p][2] = dir == 0 ? q0.z : (dir == 1 ? q0.w : q0.y);
            P[p][3] = dir == 0 ? q0.w : (dir == 1 ? q0.y : q0.z);
            P[p][4] = q1.x;
            P[p][5] = q1.y;
        }
        float acc[N][5];
#pragma unroll
        for (int p = 0; p < N; ++p)
#pragma unroll
            for (int v = 0; v < 5; ++v) acc[p][v] = 0.0f;
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
        } else {
            line_pairs<N, TP, 0>(P, acc, a.dsharp, inv_gm1);
        }
        // face fluxes at the two line ends and their lift (∂g / Ja_dd folded)
        const int n0 = line_node<3, N>(dir, f, 0), n1 = line_node<3, N>(dir, f, N - 1);
        float own0[5], own1[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) { own0[v] = sQ[le][v][n0]; own1[v] = sQ[le][v][n1]; }
        const float ija = sG[le][4 + dir];
        face_lift<N, 0>(own0, ext0, info0, slip0, sgn0, jac0, dir, ija * a.lift_m, gm1, gamma, a.diss, acc, a.errflag);
        face_lift<N, N - 1>(own1, ext1, info1, slip1, sgn1, jac1, dir, ija * a.lift_p, gm1, gamma, a.diss, acc, a.errflag);
#pragma unroll
        for (int p = 0; p < N; ++p) {
            const int nd = line_node<3, N>(dir, f, p);
            const float a1 = acc[p][1], a2 = acc[p][2], a3 = acc[p][3];
            sPart[le][dir][0][nd] = acc[p][0];
            sPart[le][dir][1][nd] = dir == 0 ? a1 : (dir == 1 ? a3 : a2);
            sPart[le][dir][2][nd] = dir == 0 ? a2 : (dir == 1 ? a1 : a3);
            sPart[le][dir][3][nd] = dir == 0 ? a3 : (dir == 1 ? a2 : a1);
            sPart[le][dir][4][nd] = acc[p][4];
        }
    }

    // the stage epilogue's second operand (RK register or SSP base), loaded before the
    // barrier so its latency overlaps the other waves' line tasks
    const float* aux = (a.mode == MODE_2N && !a.first) ? a.reg
                     : ((a.mode == MODE_SSP2 || a.mode == MODE_SSP3) ? a.ubase : nullptr);
    float xv[NPT][5];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int i = tid + r * TPB;
        const int le = i / NPE, node = i - le * NPE;
        const bool ok = i < EPB * NPE && le < nel && aux;
#pragma unroll
        for (int v = 0; v < 5; ++v)
            xv[r][v] = ok ? aux[v * Nd + (int64_t)(e0 + le) * NPE + node] : 0.0f;
    }
    __syncthreads();

    // per node: Σ_d (Ja_dd/J) row sums and the stage epilogue
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int i = tid + r * TPB;
        const int le = i / NPE, node = i - le * NPE;
        if (i >= EPB * NPE || le >= nel) continue;
        const float s0 = sG[le][0], s1 = sG[le][1], s2 = sG[le][2];
        const int64_t gi = (int64_t)(e0 + le) * NPE + node;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            const float kv = fmaf(s0, sPart[le][0][v][node], fmaf(s1, sPart[le][1][v][node], s2 * sPart[le][2][v][node]));
            const int64_t idx = v * Nd + gi;
            const float uin = sQ[le][v][node];
            const float x = xv[r][v];
            float out;
            if (a.mode == MODE_RHS) {
                out = kv;
            } else if (a.mode == MODE_2N) {
                const float t = a.first ? a.dt * kv : fmaf(a.A, x, a.dt * kv);
                a.reg[idx] = t;
                out = fmaf(a.B, t, uin);
            } else if (a.mode == MODE_SSP1) {
                out = fmaf(a.dt, kv, uin);
            } else if (a.mode == MODE_SSP2) {
                out = 0.25f * (3.0f * x + uin + a.dt * kv);
            } else {
                out = (1.0f / 3.0f) * (x + 2.0f * uin + 2.0f * a.dt * kv);
            }
            a.out[idx] = out;
        }
    }
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
