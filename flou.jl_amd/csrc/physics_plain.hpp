// Point-wise physics of the plain element kernel, in the reference's evaluation order.
//
// With T = double and -ffp-contract=off these reproduce Flou.jl's Float64 arithmetic term
// by term: left folds for n-ary sums, `a / 2b` as a / (2*b), `x^2` as x*x, `x^3` as
// (x*x)*x, divisions by 3, 5, 7 kept as divisions.  With T = float the same structure runs
// in single precision.
#pragma once

#include "common.hpp"

namespace flouhip {
namespace plain {

template <typename T> __device__ inline T dsqrt(T x, int* err) {
    if (x < T(0)) atomicOr(err, ERR_SQRT_NEG);
    return sqrt(x);
}
template <typename T> __device__ inline T dlog(T x, int* err) {
    if (x < T(0)) atomicOr(err, ERR_LOG_NEG);
    return log(x);
}

// pressure (FlouCommon/Euler.jl:142-155)
template <typename T, int ND>
__device__ inline T pressure(const T* Q, T gamma) {
    T s = Q[1] * Q[1];
#pragma unroll
    for (int k = 1; k < ND; ++k) s = s + Q[1 + k] * Q[1 + k];
    return (gamma - T(1)) * (Q[ND + 1] - s / (T(2) * Q[0]));
}

// vars_cons2prim (FlouCommon/Euler.jl:237-253)
template <typename T, int ND>
__device__ inline void cons2prim(const T* Q, T gamma, T* P) {
    const T p = pressure<T, ND>(Q, gamma);
    P[0] = Q[0];
#pragma unroll
    for (int k = 0; k < ND; ++k) P[1 + k] = Q[1 + k] / Q[0];
    P[ND + 1] = p;
}

// logarithmic_mean (FlouCommon/Utilities.jl:34-44)
template <typename T>
__device__ inline T logarithmic_mean(T al, T ar, int* err) {
    const T xi = al / ar;
    const T f = (xi - T(1)) / (xi + T(1));
    const T u = f * f;
    T F;
    if (u < T(0.01)) {
        F = ((T(1) + u / T(3)) + (u * u) / T(5)) + ((u * u) * u) / T(7);
    } else {
        F = dlog(xi, err) / (T(2) * f);
    }
    return (al + ar) / (T(2) * F);
}

// Contravariant physical volume flux in direction dir (OpDivergence.jl:28-37,
// FlouCommon/Euler.jl:54-114, LinearAdvection.jl:42-44, PhysicalRegions.jl:1001-1033).
// M: metric, column-major nd x nd.
template <typename T, int ND, int EQ>
__device__ inline void contravariant_flux(const T* Q, const T* M, int dir, T gamma,
                                          const T* avel, T* Ft) {
    constexpr int NV = EQ == EQ_ADVECTION ? 1 : ND + 2;
    T F[ND][NV];
    if constexpr (EQ == EQ_ADVECTION) {
#pragma unroll
        for (int k = 0; k < ND; ++k) F[k][0] = avel[k] * Q[0];
    } else {
        T u[ND];
#pragma unroll
        for (int m = 0; m < ND; ++m) u[m] = Q[1 + m] / Q[0];
        const T p = pressure<T, ND>(Q, gamma);
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            F[k][0] = Q[1 + k];
#pragma unroll
            for (int m = 0; m < ND; ++m) {
                F[k][1 + m] = Q[1 + k] * u[m];
                if (m == k) F[k][1 + m] = F[k][1 + m] + p;
            }
            F[k][ND + 1] = (Q[ND + 1] + p) * u[k];
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        T s = F[0][v] * M[0 + ND * dir];
#pragma unroll
        for (int k = 1; k < ND; ++k) s = s + F[k][v] * M[k + ND * dir];
        Ft[v] = s;
    }
}

// Two-point volume flux (Equations/Euler.jl:385-536; Equations/LinearAdvection.jl:44-47).
// P1/P2: cached primitives (ρ, u.., p); b1/b2: ρ/(2p).  Ja1/Ja2: metric column `dir`.
template <typename T, int ND, int EQ, int TP>
__device__ inline void twopointflux(const T* Q1, const T* Q2, const T* P1, const T* P2, T b1,
                                    T b2, const T* Ja1, const T* Ja2, T gamma, const T* avel,
                                    int* err, T* F) {
    constexpr int NV = EQ == EQ_ADVECTION ? 1 : ND + 2;
    T nrm[ND];
#pragma unroll
    for (int k = 0; k < ND; ++k) nrm[k] = (Ja1[k] + Ja2[k]) / T(2);
    if constexpr (EQ == EQ_ADVECTION) {
        T an = avel[0] * nrm[0];
#pragma unroll
        for (int k = 1; k < ND; ++k) an = an + avel[k] * nrm[k];
        F[0] = (an * (Q1[0] + Q2[0])) / T(2);
    } else {
    T f[NV][ND];
    if constexpr (TP == TP_STDAVG) {
        const T p1 = P1[ND + 1], p2 = P2[ND + 1];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            f[0][k] = (Q1[1 + k] + Q2[1 + k]) / T(2);
#pragma unroll
            for (int c = 0; c < ND; ++c) {
                T s = Q1[1 + c] * P1[1 + k];
                if (c == k) s = s + p1;
                s = s + Q2[1 + c] * P2[1 + k];
                if (c == k) s = s + p2;
                f[1 + c][k] = s / T(2);
            }
            f[ND + 1][k] = ((Q1[ND + 1] + p1) * P1[1 + k] + (Q2[ND + 1] + p2) * P2[1 + k]) / T(2);
        }
    } else if constexpr (TP == TP_CHANDRASEKHAR) {
        const T rho1 = P1[0], rho2 = P2[0];
        T um[ND];
#pragma unroll
        for (int k = 0; k < ND; ++k) um[k] = (P1[1 + k] + P2[1 + k]) / T(2);
        const T rho = logarithmic_mean(rho1, rho2, err);
        const T p = (rho1 + rho2) / (T(2) * (b1 + b2));
        const T beta = logarithmic_mean(b1, b2, err);
        T sq = P1[1] * P1[1];
#pragma unroll
        for (int k = 1; k < ND; ++k) sq = sq + P1[1 + k] * P1[1 + k];
#pragma unroll
        for (int k = 0; k < ND; ++k) sq = sq + P2[1 + k] * P2[1 + k];
        T h = T(1) / ((T(2) * beta) * (gamma - T(1))) - sq / T(4);
        h = h + p / rho;
#pragma unroll
        for (int k = 0; k < ND; ++k) h = h + um[k] * um[k];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            f[0][k] = rho * um[k];
#pragma unroll
            for (int c = 0; c < ND; ++c) {
                if (c == k) f[1 + c][k] = rho * (um[k] * um[k]) + p;
                else f[1 + c][k] = (rho * um[c < k ? c : k]) * um[c < k ? k : c];
            }
            f[ND + 1][k] = (rho * um[k]) * h;
        }
    } else {  // Kennedy & Gruber, JCP 227 (2008); not present in the reference
        const T p1 = P1[ND + 1], p2 = P2[ND + 1];
        const T rho = (P1[0] + P2[0]) / T(2);
        T um[ND];
#pragma unroll
        for (int k = 0; k < ND; ++k) um[k] = (P1[1 + k] + P2[1 + k]) / T(2);
        const T p = (p1 + p2) / T(2);
        const T e = (Q1[ND + 1] / Q1[0] + Q2[ND + 1] / Q2[0]) / T(2);
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            const T m = rho * um[k];
            f[0][k] = m;
#pragma unroll
            for (int c = 0; c < ND; ++c) {
                f[1 + c][k] = m * um[c];
                if (c == k) f[1 + c][k] = f[1 + c][k] + p;
            }
            f[ND + 1][k] = m * e + p * um[k];
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        T s = f[v][0] * nrm[0];
#pragma unroll
        for (int k = 1; k < ND; ++k) s = s + f[v][k] * nrm[k];
        F[v] = s;
    }
    }
}

// rotate2face / rotate2phys (Equations/Euler.jl:16-64).  fr = (n, t, b), nd each.
template <typename T, int ND>
__device__ inline void rotate2face(const T* Q, const T* fr, T* R) {
    const T* n = fr;
    const T* t = fr + ND;
    const T* b = fr + 2 * ND;
    R[0] = Q[0];
    if constexpr (ND == 1) {
        R[1] = Q[1] * n[0];
    } else if constexpr (ND == 2) {
        R[1] = Q[1] * n[0] + Q[2] * n[1];
        R[2] = Q[1] * t[0] + Q[2] * t[1];
    } else {
        R[1] = (Q[1] * n[0] + Q[2] * n[1]) + Q[3] * n[ND - 1];
        R[2] = (Q[1] * t[0] + Q[2] * t[1]) + Q[3] * t[ND - 1];
        R[3] = (Q[1] * b[0] + Q[2] * b[1]) + Q[3] * b[ND - 1];
    }
    R[ND + 1] = Q[ND + 1];
}

template <typename T, int ND>
__device__ inline void rotate2phys(const T* R, const T* fr, T* Q) {
    const T* n = fr;
    const T* t = fr + ND;
    const T* b = fr + 2 * ND;
    Q[0] = R[0];
    if constexpr (ND == 1) {
        Q[1] = R[1] * n[0];
    } else if constexpr (ND == 2) {
        Q[1] = R[1] * n[0] + R[2] * t[0];
        Q[2] = R[1] * n[1] + R[2] * t[1];
    } else {
        Q[1] = (R[1] * n[0] + R[2] * t[0]) + R[ND] * b[0];
        Q[2] = (R[1] * n[1] + R[2] * t[1]) + R[ND] * b[1];
        Q[ND] = (R[1] * n[ND - 1] + R[2] * t[ND - 1]) + R[ND] * b[ND - 1];
    }
    Q[ND + 1] = R[ND + 1];
}

// vars_cons2entropy (FlouCommon/Euler.jl:273-307)
template <typename T, int ND>
__device__ inline void cons2entropy(const T* Q, T gamma, int* err, T* W) {
    const T p = pressure<T, ND>(Q, gamma);
    const T s = dlog(p, err) - gamma * dlog(Q[0], err);
    T m2 = Q[1] * Q[1];
#pragma unroll
    for (int k = 1; k < ND; ++k) m2 = m2 + Q[1 + k] * Q[1 + k];
    W[0] = (gamma - s) / (gamma - T(1)) - (m2 / Q[0]) / (T(2) * p);
#pragma unroll
    for (int k = 0; k < ND; ++k) W[1 + k] = Q[1 + k] / p;
    W[ND + 1] = -Q[0] / p;
}

// vars_entropy2prim then vars_prim2cons (FlouCommon/Euler.jl:309-339, 255-271).
template <typename T, int ND>
__device__ inline void entropy2cons(const T* W, T gamma, int* err, T* Q) {
    const T wl = W[ND + 1];
    T u[ND];
#pragma unroll
    for (int k = 0; k < ND; ++k) u[k] = -W[1 + k] / wl;
    T q2 = u[0] * u[0];
#pragma unroll
    for (int k = 1; k < ND; ++k) q2 = q2 + u[k] * u[k];
    const T s = gamma - (gamma - T(1)) * (W[0] - (wl * q2) / T(2));
    if (-wl < T(0)) atomicOr(err, ERR_POW_NEG);
    const T base = pow(-wl, gamma) * exp(s);
    if (base < T(0)) atomicOr(err, ERR_POW_NEG);
    const T p = pow(base, T(1) / (T(1) - gamma));
    const T rho = -p * wl;
    Q[0] = rho;
#pragma unroll
    for (int k = 0; k < ND; ++k) Q[1 + k] = rho * u[k];
    Q[ND + 1] = p / (gamma - T(1)) + (rho * q2) / T(2);
}

// Euler numerical flux in the face frame (Equations/Euler.jl:99-380).
template <typename T, int ND>
__device__ inline void euler_numericalflux(const T* Ql, const T* Qr, int savg, int sdiss,
                                           int typo, T intensity, T g, int* err, T* Fn) {
    constexpr int NV = ND + 2;
    T Pl[NV], Pr[NV];
    cons2prim<T, ND>(Ql, g, Pl);
    cons2prim<T, ND>(Qr, g, Pr);
    const T pl = Pl[ND + 1], pr = Pr[ND + 1];
    const T ul = Pl[1], ur = Pr[1];
    if (savg == AVG_STD) {
        Fn[0] = (Ql[1] + Qr[1]) / T(2);
        Fn[1] = (((Ql[1] * ul + pl) + Qr[1] * ur) + pr) / T(2);
#pragma unroll
        for (int c = 1; c < ND; ++c) Fn[1 + c] = (Ql[1 + c] * ul + Qr[1 + c] * ur) / T(2);
        Fn[ND + 1] = ((Ql[ND + 1] + pl) * ul + (Qr[ND + 1] + pr) * ur) / T(2);
    } else {
        // ChandrasekharAverage numerical flux (Euler.jl:169-226).  `typo` reproduces
        // Euler.jl:216, which squares wl where wr is meant (3-D only).
        const T rl = Pl[0], rr = Pr[0];
        T um[ND];
#pragma unroll
        for (int k = 0; k < ND; ++k) um[k] = (Pl[1 + k] + Pr[1 + k]) / T(2);
        const T bl = rl / (T(2) * pl), br = rr / (T(2) * pr);
        const T rho = logarithmic_mean(rl, rr, err);
        const T p = (rl + rr) / (T(2) * (bl + br));
        const T beta = logarithmic_mean(bl, br, err);
        T sq = Pl[1] * Pl[1];
#pragma unroll
        for (int k = 1; k < ND; ++k) sq = sq + Pl[1 + k] * Pl[1 + k];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            const T c = (ND == 3 && k == 2 && typo) ? Pl[ND] : Pr[1 + k];
            sq = sq + c * c;
        }
        T h = T(1) / ((T(2) * beta) * (g - T(1))) - sq / T(4);
        h = h + p / rho;
#pragma unroll
        for (int k = 0; k < ND; ++k) h = h + um[k] * um[k];
        Fn[0] = rho * um[0];
        Fn[1] = rho * (um[0] * um[0]) + p;
#pragma unroll
        for (int c = 1; c < ND; ++c) Fn[1 + c] = (rho * um[0]) * um[c];
        Fn[ND + 1] = (rho * um[0]) * h;
    }
    if (sdiss == DISS_LXF) {
        const T al = dsqrt(g * pl / Pl[0], err);
        const T ar = dsqrt(g * pr / Pr[0], err);
        const T la = fabs(ul) + al, lb = fabs(ur) + ar;
        const T lam = (la < lb) ? lb : la;
#pragma unroll
        for (int v = 0; v < NV; ++v) Fn[v] = Fn[v] + ((lam * (Ql[v] - Qr[v])) / T(2)) * intensity;
    } else if (sdiss == DISS_SCALAR) {
        // ScalarDissipation (Euler.jl:228-301); `typo` reproduces the 3-D left momenta
        // taken from Qr (Euler.jl:257)
        const T rl = Pl[0], rr = Pr[0];
        T um[ND];
#pragma unroll
        for (int k = 0; k < ND; ++k) um[k] = (Pl[1 + k] + Pr[1 + k]) / T(2);
        const T rho = (rl + rr) / T(2);
        const T bl = rl / (T(2) * pl), br = rr / (T(2) * pr);
        const T beta = logarithmic_mean(bl, br, err);
        const T al = dsqrt(g * pl / rl, err);
        const T ar = dsqrt(g * pr / rr, err);
        const T la = fabs(ul) + al, lb = fabs(ur) + ar;
        const T lam = (la < lb) ? lb : la;
        T D[NV];
        D[0] = rr - rl;
#pragma unroll
        for (int k = 0; k < ND; ++k) D[1 + k] = Qr[1 + k] - ((ND == 3 && typo) ? Qr[1 + k] : Ql[1 + k]);
        T s = (T(1) / beta) / (g - T(1)) + Pl[1] * Pr[1];
#pragma unroll
        for (int k = 1; k < ND; ++k) s = s + Pl[1 + k] * Pr[1 + k];
        T t = um[0] * (Pr[1] - Pl[1]);
#pragma unroll
        for (int k = 1; k < ND; ++k) t = t + um[k] * (Pr[1 + k] - Pl[1 + k]);
        t = t + (T(1) / br - T(1) / bl) / (T(2) * (g - T(1)));
        D[ND + 1] = (s * (rr - rl)) / T(2) + rho * t;
#pragma unroll
        for (int v = 0; v < NV; ++v) Fn[v] = Fn[v] - ((lam / T(2)) * D[v]) * intensity;
    } else if (sdiss == DISS_MATRIX) {
        const T rl = Pl[0], rr = Pr[0];
        T um[ND];
#pragma unroll
        for (int k = 0; k < ND; ++k) um[k] = (Pl[1 + k] + Pr[1 + k]) / T(2);
        T m2 = um[0] * um[0];
#pragma unroll
        for (int k = 1; k < ND; ++k) m2 = m2 + um[k] * um[k];
        T q2 = Pl[1] * Pl[1];
#pragma unroll
        for (int k = 1; k < ND; ++k) q2 = q2 + Pl[1 + k] * Pl[1 + k];
#pragma unroll
        for (int k = 0; k < ND; ++k) q2 = q2 + Pr[1 + k] * Pr[1 + k];
        const T v2 = T(2) * m2 - q2 / T(2);
        const T bl = rl / (T(2) * pl), br = rr / (T(2) * pr);
        const T rho = logarithmic_mean(rl, rr, err);
        const T p = (rl + rr) / (T(2) * (bl + br));
        const T beta = logarithmic_mean(bl, br, err);
        const T a = dsqrt(g * p / rho, err);
        const T h = ((g / (T(2) * beta)) / (g - T(1))) + v2 / T(2);
        T Wl[NV], Wr[NV], dW[NV];
        cons2entropy<T, ND>(Ql, g, err, Wl);
        cons2entropy<T, ND>(Qr, g, err, Wr);
#pragma unroll
        for (int v = 0; v < NV; ++v) dW[v] = Wl[v] - Wr[v];
        const T u = um[0];
        T Lam[NV], Tm[NV], R[NV][NV];
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int j = 0; j < NV; ++j) R[i][j] = T(0);
        Lam[0] = fabs(u - a);
        Lam[NV - 1] = fabs(u + a);
#pragma unroll
        for (int c = 1; c < NV - 1; ++c) Lam[c] = fabs(u);
        Tm[0] = rho / (T(2) * g);
        Tm[1] = ((g - T(1)) * rho) / g;
#pragma unroll
        for (int c = 2; c < NV - 1; ++c) Tm[c] = p;
        Tm[NV - 1] = rho / (T(2) * g);
        R[0][0] = T(1); R[1][0] = u - a;
#pragma unroll
        for (int k = 1; k < ND; ++k) R[1 + k][0] = um[k];
        R[NV - 1][0] = h - u * a;
        R[0][1] = T(1); R[1][1] = u;
#pragma unroll
        for (int k = 1; k < ND; ++k) R[1 + k][1] = um[k];
        R[NV - 1][1] = v2 / T(2);
#pragma unroll
        for (int k = 1; k < ND; ++k) {
            R[1 + k][1 + k] = T(1);
            R[NV - 1][1 + k] = um[k];
        }
        R[0][NV - 1] = T(1); R[1][NV - 1] = u + a;
#pragma unroll
        for (int k = 1; k < ND; ++k) R[1 + k][NV - 1] = um[k];
        R[NV - 1][NV - 1] = h + u * a;
        T RL[NV][NV];
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int j = 0; j < NV; ++j) RL[i][j] = (R[i][j] * Lam[j]) * Tm[j];
        T MdW[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            T s = T(0);
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                T Mij = RL[i][0] * R[j][0];
#pragma unroll
                for (int k = 1; k < NV; ++k) Mij = Mij + RL[i][k] * R[j][k];
                s = (j == 0) ? Mij * dW[0] : s + Mij * dW[j];
            }
            MdW[i] = s;
        }
#pragma unroll
        for (int i = 0; i < NV; ++i) Fn[i] = Fn[i] + (MdW[i] / T(2)) * intensity;
    }
}


// HybridDivOperator pieces (OpDivergence.jl:441-801).
// Subcell-interface frame of direction dir on Cartesian elements (PhysicalSubgrid for
// CartesianMesh, PhysicalRegions.jl:75-145): n along the axis, (t, b) in the reference's
// order, then the surface Jacobian.
template <typename T, int ND>
__device__ inline void subgrid_frame(int dir, T sjac, T* g) {
#pragma unroll
    for (int i = 0; i < 3 * ND; ++i) g[i] = T(0);
    if constexpr (ND == 1) {
        g[0] = T(1);
    } else if constexpr (ND == 2) {
        if (dir == 0) { g[0] = T(1); g[2 + 1] = T(1); }
        else { g[1] = T(1); g[2 + 0] = T(-1); }
    } else {
        const int t = (dir + 1) % 3, b = (dir + 2) % 3;
        g[dir] = T(1); g[3 + t] = T(1); g[6 + b] = T(1);
    }
    g[3 * ND] = sjac;
}

// Blending factor (OpDivergence.jl:781-801): δ = max((√(b²+c) - b)/√(b²+c), 1/2).
template <typename T>
__device__ inline T hybrid_delta(T b, T c, int* err) {
    T s = dsqrt(b * b + c, err);
    s = (s - b) / s;
    return (s < T(0.5)) ? T(0.5) : s;
}

}  // namespace plain
}  // namespace flouhip
