// FP32 ("chosen form") fused element kernel for the 3-D split-form Euler rhs on
// axis-aligned affine hexahedra (Flou's CartesianMesh), GLL nodes, LxF(StdAverage) or
// StdAverage surface flux, with the RK stage epilogue.
//
// Same mathematics as SplitDivOperator + interface_fluxes! + the FR lift
// (OpDivergence.jl:200-299, Interfaces.jl:111-136), re-organised for gfx950:
//   * one element per block; the 216 (N=6) nodal values of all variables are staged in LDS
//     with coalesced loads, and per node the two-point flux inputs are computed once
//     (Chandrasekhar: ρ, u, v, w, β = ρ/2p, ln ρ, ln β, |u|²) instead of once per pair;
//   * line tasks: one thread per (direction, line) evaluates every pair of its line once
//     (F♯ is symmetric) and accumulates both rows in registers, F♯ entries never touch
//     memory; only the per-direction row sums go to LDS;
//   * face tasks run concurrently on the other waves: the Rusanov flux on every face node,
//     evaluated with the master side as left state so both elements of a face receive
//     exactly opposite values;
//   * the last phase adds the three directions scaled by Ja_dd/J, the lift of the face
//     fluxes (boundary nodes only, GLL), and applies the RK stage combination.
// Reformulations relative to the reference (FP32 only; the FP64 plain kernel keeps the
// reference's operation order):
//   - log-mean: ρ̂ = (a+b)/(2F), F = 1+u/3+u²/5+u³/7 (u = f², f = (a-b)/(a+b)) for u < 0.01
//     (same series as Utilities.jl:34-44, with f formed from a-b and a+b instead of a/b),
//     F = (ln a - ln b)/(2f) otherwise, with ln cached per node;
//   - 1/(2β̂) = F_β/(β1+β2) and p̂/ρ̂ = F_ρ/(β1+β2) (algebraically identical), so a pair
//     needs three reciprocals (hardware v_rcp_f32) and no division or log;
//   - the metric (diagonal, constant per element) and 1/J are applied once per node.
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
    static constexpr int NL = 3 * N * N;                // line tasks
    static constexpr int FT0 = ((NL + 63) / 64) * 64;   // first face-task thread
    static constexpr int TPB = ((NPE + 63) / 64) * 64 > FT0 + 64 ? ((NPE + 63) / 64) * 64 : FT0 + 64;
    static constexpr int NFT = TPB - FT0;               // face-task threads
    static_assert(NFT > 0, "no threads left for face tasks");
};

// Two-point flux in the frame of the line: P[1] is the velocity along the line, P[2], P[3]
// the other two.  Returns the line-direction component of the physical two-point flux,
// momentum components in the same permuted order.
template <int TP>
__device__ __forceinline__ void twopoint(const float* Pi, const float* Pj, float inv_gm1, float* F) {
    if constexpr (TP == TP_CHANDRASEKHAR) {
        // P = (ρ, u_n, u_t1, u_t2, β, ln ρ, ln β, |u|²)
        const float sr = Pi[0] + Pj[0], dr = Pi[0] - Pj[0];
        const float sb = Pi[4] + Pj[4], db = Pi[4] - Pj[4];
        const float rs = rcp(sr), rb = rcp(sb);
        const float fr = dr * rs, fb = db * rb;
        const float ur = fr * fr, ub = fb * fb;
        float Fr, Fb;
        if (ur < 0.01f) Fr = 1.0f + ur * (1.0f / 3.0f + ur * (1.0f / 5.0f + ur * (1.0f / 7.0f)));
        else Fr = 0.5f * (Pi[5] - Pj[5]) * rcp(fr);
        if (ub < 0.01f) Fb = 1.0f + ub * (1.0f / 3.0f + ub * (1.0f / 5.0f + ub * (1.0f / 7.0f)));
        else Fb = 0.5f * (Pi[6] - Pj[6]) * rcp(fb);
        const float hs = 0.5f * sr;
        const float rho = hs * rcp(Fr);
        const float p = hs * rb;
        const float un = 0.5f * (Pi[1] + Pj[1]);
        const float ut1 = 0.5f * (Pi[2] + Pj[2]);
        const float ut2 = 0.5f * (Pi[3] + Pj[3]);
        const float h = rb * (Fb * inv_gm1 + Fr) - 0.25f * (Pi[7] + Pj[7]) + (un * un + ut1 * ut1 + ut2 * ut2);
        const float m = rho * un;
        F[0] = m;
        F[1] = m * un + p;
        F[2] = m * ut1;
        F[3] = m * ut2;
        F[4] = m * h;
    } else if constexpr (TP == TP_KENNEDY_GRUBER) {
        // P = (ρ, u_n, u_t1, u_t2, p, e = ρe/ρ)
        const float rho = 0.5f * (Pi[0] + Pj[0]);
        const float un = 0.5f * (Pi[1] + Pj[1]);
        const float ut1 = 0.5f * (Pi[2] + Pj[2]);
        const float ut2 = 0.5f * (Pi[3] + Pj[3]);
        const float p = 0.5f * (Pi[4] + Pj[4]);
        const float e = 0.5f * (Pi[5] + Pj[5]);
        const float m = rho * un;
        F[0] = m;
        F[1] = m * un + p;
        F[2] = m * ut1;
        F[3] = m * ut2;
        F[4] = m * e + p * un;
    } else {
        // StdAverage: ½(f_n(Q_i) + f_n(Q_j)), P = (ρ, u_n, u_t1, u_t2, p, ρe)
        const float mi = Pi[0] * Pi[1], mj = Pj[0] * Pj[1];
        F[0] = 0.5f * (mi + mj);
        F[1] = 0.5f * ((mi * Pi[1] + Pi[4]) + (mj * Pj[1] + Pj[4]));
        F[2] = 0.5f * (mi * Pi[2] + mj * Pj[2]);
        F[3] = 0.5f * (mi * Pi[3] + mj * Pj[3]);
        F[4] = 0.5f * ((Pi[5] + Pi[4]) * Pi[1] + (Pj[5] + Pj[4]) * Pj[1]);
    }
}

// Physical Euler flux along axis d (momentum in x, y, z order) and sound speed.
__device__ __forceinline__ void euler_flux_axis(const float* Q, int d, float gm1, float gamma,
                                                float* f, float& un, float& c, int* err) {
    const float ir = rcp(Q[0]);
    const float m2 = Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3];
    const float p = gm1 * (Q[4] - 0.5f * m2 * ir);
    const float md = d == 0 ? Q[1] : (d == 1 ? Q[2] : Q[3]);
    un = md * ir;
    const float c2 = gamma * p * ir;
    if (!(c2 >= 0.0f)) atomicOr(err, ERR_SQRT_NEG);
    c = sq(c2);
    f[0] = md;
    f[1] = Q[1] * un + (d == 0 ? p : 0.0f);
    f[2] = Q[2] * un + (d == 1 ? p : 0.0f);
    f[3] = Q[3] * un + (d == 2 ? p : 0.0f);
    f[4] = (Q[4] + p) * un;
}

template <int N, int TP>
__global__ void __launch_bounds__((FCfg<N>::TPB)) k_fast3d(const FastArgs a) {
    using C = FCfg<N>;
    constexpr int NPE = C::NPE, NPF = C::NPF, NL = C::NL, TPB = C::TPB, FT0 = C::FT0, NFT = C::NFT;
    __shared__ float sQ[5][NPE];
    __shared__ float sP[8][NPE];
    __shared__ float sPart[3][5][NPE];
    __shared__ float sFn[6][5][NPF];
    __shared__ float sG[8];

    const int tid = threadIdx.x;
    const int e = xcd_remap(blockIdx.x, gridDim.x);
    if (e >= a.ne) return;  // uniform over the block
    const int64_t Nd = a.ndofs;
    const float gamma = a.gamma, gm1 = a.gamma - 1.0f, inv_gm1 = 1.0f / (a.gamma - 1.0f);
    {
        const float* Qe = a.Q + (int64_t)e * NPE;
        for (int i = tid; i < 5 * NPE; i += TPB) {
            const int v = i / NPE, nd = i - v * NPE;
            sQ[v][nd] = Qe[v * Nd + nd];
        }
        if (tid < 8) sG[tid] = a.egeom[(int64_t)e * a.egeom_stride + tid];
    }
    __syncthreads();

    // Per-node inputs of the two-point flux
    if (tid < NPE) {
        const float rho = sQ[0][tid];
        const float ir = rcp(rho);
        const float u = sQ[1][tid] * ir, v = sQ[2][tid] * ir, w = sQ[3][tid] * ir;
        const float q = u * u + v * v + w * w;
        const float p = gm1 * (sQ[4][tid] - 0.5f * rho * q);
        if (!(rho > 0.0f) || !(p > 0.0f)) atomicOr(a.errflag, ERR_LOG_NEG);
        sP[0][tid] = rho;
        sP[1][tid] = u;
        sP[2][tid] = v;
        sP[3][tid] = w;
        if constexpr (TP == TP_CHANDRASEKHAR) {
            const float beta = 0.5f * rho * rcp(p);
            sP[4][tid] = beta;
            sP[5][tid] = lg(rho);
            sP[6][tid] = lg(beta);
            sP[7][tid] = q;
        } else if constexpr (TP == TP_KENNEDY_GRUBER) {
            sP[4][tid] = p;
            sP[5][tid] = sQ[4][tid] * ir;
        } else {
            sP[4][tid] = p;
            sP[5][tid] = sQ[4][tid];
        }
    }
    __syncthreads();

    if (tid < NL) {
        // ---- line task: all pairs of one line, both rows accumulated in registers
        const int dir = tid / NPF, f = tid - dir * NPF;
        float P[N][8];
#pragma unroll
        for (int p = 0; p < N; ++p) {
            const int nd = line_node<3, N>(dir, f, p);
#pragma unroll
            for (int k = 0; k < 8; ++k) P[p][k] = sP[k][nd];
            // permute velocities so that P[p][1] is along the line
            const float u = P[p][1], v = P[p][2], w = P[p][3];
            P[p][1] = dir == 0 ? u : (dir == 1 ? v : w);
            P[p][2] = dir == 0 ? v : (dir == 1 ? w : u);
            P[p][3] = dir == 0 ? w : (dir == 1 ? u : v);
        }
        float acc[N][5];
#pragma unroll
        for (int p = 0; p < N; ++p)
#pragma unroll
            for (int v = 0; v < 5; ++v) acc[p][v] = 0.0f;
        const float* Dsh = a.dsharp;
#pragma unroll
        for (int i = 0; i < N; ++i) {
#pragma unroll
            for (int j = i + 1; j < N; ++j) {
                float F[5];
                twopoint<TP>(P[i], P[j], inv_gm1, F);
                const float cij = Dsh[i * N + j], cji = Dsh[j * N + i];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    acc[i][v] = fmaf(-cij, F[v], acc[i][v]);
                    acc[j][v] = fmaf(-cji, F[v], acc[j][v]);
                }
            }
        }
#pragma unroll
        for (int p = 0; p < N; ++p) {
            const int nd = line_node<3, N>(dir, f, p);
            // undo the velocity permutation
            const float a1 = acc[p][1], a2 = acc[p][2], a3 = acc[p][3];
            sPart[dir][0][nd] = acc[p][0];
            sPart[dir][1][nd] = dir == 0 ? a1 : (dir == 1 ? a3 : a2);
            sPart[dir][2][nd] = dir == 0 ? a2 : (dir == 1 ? a1 : a3);
            sPart[dir][3][nd] = dir == 0 ? a3 : (dir == 1 ? a2 : a1);
            sPart[dir][4][nd] = acc[p][4];
        }
    } else if (tid >= FT0) {
        // ---- face tasks: Rusanov flux on every face node, master side as the left state
        for (int it = tid - FT0; it < 6 * NPF; it += NFT) {
            const int k = it / NPF, f = it - k * NPF;
            const int d = k >> 1, lr = k & 1;
            float own[5], ext[5];
            const int on = line_node<3, N>(d, f, lr ? N - 1 : 0);
#pragma unroll
            for (int v = 0; v < 5; ++v) own[v] = sQ[v][on];
            const int slot = e * 6 + k;
            const int info = a.finfo[slot];
            const int nbe = a.nb[slot];
            const float* g = a.fgeom + (int64_t)a.fgi[slot] * 10;
            if (nbe >= 0) {
                const int kp = finfo_nbpos(info);
                const int nn = line_node<3, N>(kp >> 1, f, (kp & 1) ? N - 1 : 0);
                const float* Qn = a.Q + (int64_t)nbe * NPE + nn;
#pragma unroll
                for (int v = 0; v < 5; ++v) ext[v] = Qn[v * Nd];
            } else {
                const int bd = finfo_bd(info);
                const int type = a.bctype[bd];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    if (type == BC_DIRICHLET) ext[v] = a.bcstate[((int64_t)a.fid[slot] * NPF + f) * 5 + v];
                    else if (type == BC_INFLOW) ext[v] = a.bcconst[bd * 5 + v];
                    else ext[v] = own[v];
                }
                if (type == BC_SLIP) ext[1 + d] = -ext[1 + d];
            }
            const bool slave = finfo_side(info);
            const float* QL = slave ? ext : own;
            const float* QR = slave ? own : ext;
            const float s = g[d];  // master normal = s e_d
            float fL[5], fR[5], uL, uR, cL, cR;
            euler_flux_axis(QL, d, gm1, gamma, fL, uL, cL, a.errflag);
            euler_flux_axis(QR, d, gm1, gamma, fR, uR, cR, a.errflag);
            const float lam = fmaxf(fabsf(uL) + cL, fabsf(uR) + cR) * a.diss;
            const float jac = g[9];
            const float sj = slave ? -jac : jac;
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const float F = 0.5f * (s * (fL[v] + fR[v]) + lam * (QL[v] - QR[v]));
                sFn[k][v][f] = F * sj;
            }
        }
    }
    __syncthreads();
    if (tid >= NPE) return;

    // ---- per node: Σ_d (Ja_dd/J) row sums, the FR lift, the stage epilogue
    const int node = tid;
    const int ix = node % N, iy = (node / N) % N, iz = node / (N * N);
    const float s0 = sG[0], s1 = sG[1], s2 = sG[2], invJ = sG[3];
    const float gmi = a.lift_m * invJ, gpi = a.lift_p * invJ;
    float val[5];
#pragma unroll
    for (int v = 0; v < 5; ++v)
        val[v] = s0 * sPart[0][v][node] + s1 * sPart[1][v][node] + s2 * sPart[2][v][node];
    // lift at boundary nodes (GLL: ∂g vanishes at interior nodes)
    {
        const int fxi = iy + N * iz, fyi = ix + N * iz, fzi = ix + N * iy;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            if (ix == 0) val[v] -= gmi * sFn[0][v][fxi];
            if (ix == N - 1) val[v] -= gpi * sFn[1][v][fxi];
            if (iy == 0) val[v] -= gmi * sFn[2][v][fyi];
            if (iy == N - 1) val[v] -= gpi * sFn[3][v][fyi];
            if (iz == 0) val[v] -= gmi * sFn[4][v][fzi];
            if (iz == N - 1) val[v] -= gpi * sFn[5][v][fzi];
        }
    }
    const int64_t gi = (int64_t)e * NPE + node;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const int64_t idx = v * Nd + gi;
        const float kv = val[v];
        const float uin = sQ[v][node];
        switch (a.mode) {
            case MODE_RHS:
                a.out[idx] = kv;
                break;
            case MODE_2N: {
                const float t = a.first ? a.dt * kv : fmaf(a.A, a.reg[idx], a.dt * kv);
                a.reg[idx] = t;
                a.out[idx] = fmaf(a.B, t, uin);
                break;
            }
            case MODE_SSP1:
                a.out[idx] = fmaf(a.dt, kv, uin);
                break;
            case MODE_SSP2:
                a.out[idx] = 0.25f * (3.0f * a.ubase[idx] + uin + a.dt * kv);
                break;
            default:
                a.out[idx] = (1.0f / 3.0f) * (a.ubase[idx] + 2.0f * uin + 2.0f * a.dt * kv);
                break;
        }
    }
}

template <int N, int TP>
static hipError_t launch(const FastArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((k_fast3d<N, TP>), dim3(a.ne), dim3(FCfg<N>::TPB), 0, s, a);
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
