// Reverse-mode (vector-Jacobian) products of the point-wise physics of the rhs.
//
// Every function here returns the exact derivative of the corresponding forward function
// of physics_plain.hpp (same branches: the log-mean series below u = 0.01, the side that
// wins the LxF max), accumulated into the caller's cotangent (+=).  Derivations work on
// primitives (ρ, u, p) and are pulled back to conservative variables by prim_pullback.
// No reference fixes the rounding of these products, so the hot ones (pair fluxes, face
// fluxes) multiply by a few shared reciprocals instead of dividing.
// The functions are __host__ __device__ so that they can be unit-tested on the CPU.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#define FLOU_HD __host__ __device__
#else
#define FLOU_HD
#endif

namespace flouhip {
namespace adj {

// 1/x for normal x whose reciprocal is normal: on the device the hardware reciprocal refined
// by Newton steps (two in FP64, one in FP32; within about an ulp of the division and far
// cheaper than its IEEE sequence), on the host the division.
template <typename T>
FLOU_HD inline T rcp(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 8) {
        T r = __builtin_amdgcn_rcp(x);
        r = fma(r, fma(-x, r, T(1)), r);
        return fma(r, fma(-x, r, T(1)), r);
    } else {
        const T r = __builtin_amdgcn_rcpf(x);
        return fmaf(r, fmaf(-x, r, T(1)), r);
    }
#else
    return T(1) / x;
#endif
}

// Cotangents of the primitives (ρ, u_1..u_d, p) of a state Q, pulled back to Q:
//   u = m/ρ,  p = (γ-1)(E - |m|²/(2ρ))   (FlouCommon/Euler.jl:142-155, 237-253)
template <typename T, int ND>
FLOU_HD inline void prim_pullback(const T* Q, T gamma, T rb, const T* ub, T pb, T* Qb) {
    const T ir = T(1) / Q[0];
    T uu = T(0), ubu = T(0);
    T u[ND];
    for (int k = 0; k < ND; ++k) {
        u[k] = Q[1 + k] * ir;
        uu += u[k] * u[k];
        ubu += ub[k] * u[k];
    }
    const T g1 = gamma - T(1);
    Qb[0] += rb - ubu * ir + g1 * pb * uu * T(0.5);
    for (int k = 0; k < ND; ++k) Qb[1 + k] += ub[k] * ir - g1 * pb * u[k];
    Qb[ND + 1] += g1 * pb;
}

// prim_pullback with the reciprocal 1/ρ given
template <typename T, int ND>
FLOU_HD inline void prim_pullback_ir(const T* Q, T ir, T gamma, T rb, const T* ub, T pb, T* Qb) {
    T uu = T(0), ubu = T(0);
    T u[ND];
    for (int k = 0; k < ND; ++k) {
        u[k] = Q[1 + k] * ir;
        uu += u[k] * u[k];
        ubu += ub[k] * u[k];
    }
    const T g1 = gamma - T(1);
    Qb[0] += rb - ubu * ir + g1 * pb * uu * T(0.5);
    for (int k = 0; k < ND; ++k) Qb[1 + k] += ub[k] * ir - g1 * pb * u[k];
    Qb[ND + 1] += g1 * pb;
}

// Physical Euler flux along a vector n, f_n(Q) = (m·n, (m·n)u + p n, (E+p) u·n)
// (FlouCommon/Euler.jl:54-114 contracted as in PhysicalRegions.jl:1001-1033):
// Qb += (∂f_n/∂Q)^T gb.
template <typename T, int ND>
FLOU_HD inline void flux_vjp_ir(const T* Q, T ir, const T* n, T gamma, const T* gb, T* Qb) {
    const T rho = Q[0];
    T u[ND], m2 = T(0);
    for (int k = 0; k < ND; ++k) {
        u[k] = Q[1 + k] * ir;
        m2 += Q[1 + k] * u[k];
    }
    const T p = (gamma - T(1)) * (Q[ND + 1] - m2 * T(0.5));
    T un = T(0), gmn = T(0), gmu = T(0);
    for (int k = 0; k < ND; ++k) {
        un += u[k] * n[k];
        gmn += gb[1 + k] * n[k];
        gmu += gb[1 + k] * u[k];
    }
    // f_0 = ρ un, f_{1+c} = ρ un u_c + p n_c, f_E = (E+p) un, un = u·n
    const T A = gb[0] + gmu;
    const T unb = rho * A + gb[ND + 1] * (Q[ND + 1] + p);
    T ub[ND];
    for (int k = 0; k < ND; ++k) ub[k] = rho * un * gb[1 + k] + unb * n[k];
    const T pb = gmn + gb[ND + 1] * un;
    Qb[ND + 1] += gb[ND + 1] * un;
    prim_pullback_ir<T, ND>(Q, ir, gamma, un * A, ub, pb, Qb);
}

template <typename T, int ND>
FLOU_HD inline void flux_vjp(const T* Q, const T* n, T gamma, const T* gb, T* Qb) {
    flux_vjp_ir<T, ND>(Q, T(1) / Q[0], n, gamma, gb, Qb);
}

// logarithmic_mean(a, b) (FlouCommon/Utilities.jl:34-44) and ∂/∂a, same branch as the
// forward: L = (a+b)/(2F(u)), f = (a-b)/(a+b), u = f².
template <typename T>
FLOU_HD inline T logmean_da(T a, T b, T* L) {
    const T s = a + b;
    const T f = (a - b) / s;
    const T u = f * f;
    T F, dF;
    if (u < T(0.01)) {
        F = ((T(1) + u / T(3)) + (u * u) / T(5)) + ((u * u) * u) / T(7);
        dF = T(1) / T(3) + u * (T(2) / T(5) + u * (T(3) / T(7)));
    } else {
        F = log(a / b) / (T(2) * f);
        dF = (T(1) / (T(1) - u) - F) / (T(2) * u);
    }
    *L = s / (T(2) * F);
    // dL/da = 1/(2F) - s/(2F²) F'(u) 2f 2b/s²
    return T(1) / (T(2) * F) - (T(2) * f * b * dF) / (F * F * s);
}

// Two-point volume flux along n̄ (Equations/Euler.jl:385-536; Kennedy-Gruber closed form):
// Qb1 += (∂F♯(Q1, Q2)/∂Q1)^T Gb.  F♯ is symmetric, so ∂/∂Q2 is this with Q1, Q2 swapped.
template <typename T, int ND, int TP>
FLOU_HD inline void twopoint_vjp1(const T* Q1, const T* Q2, const T* n, T gamma, const T* Gb,
                                  T* Qb1) {
    if constexpr (TP == TP_STDAVG) {
        T h[ND + 2];
        for (int v = 0; v < ND + 2; ++v) h[v] = T(0.5) * Gb[v];
        flux_vjp<T, ND>(Q1, n, gamma, h, Qb1);
        return;
    } else {
        const T g1 = gamma - T(1);
        T u1[ND], u2[ND], um[ND];
        T q1 = T(0), q2 = T(0);
        for (int k = 0; k < ND; ++k) {
            u1[k] = Q1[1 + k] / Q1[0];
            u2[k] = Q2[1 + k] / Q2[0];
            um[k] = (u1[k] + u2[k]) * T(0.5);
            q1 += Q1[1 + k] * Q1[1 + k];
            q2 += Q2[1 + k] * Q2[1 + k];
        }
        const T p1 = g1 * (Q1[ND + 1] - q1 / (T(2) * Q1[0]));
        const T p2 = g1 * (Q2[ND + 1] - q2 / (T(2) * Q2[0]));
        T umn = T(0), Gmn = T(0), Gmu = T(0);
        for (int k = 0; k < ND; ++k) {
            umn += um[k] * n[k];
            Gmn += Gb[1 + k] * n[k];
            Gmu += Gb[1 + k] * um[k];
        }
        T ub1[ND];
        T rb1 = T(0), pb1 = T(0);
        if constexpr (TP == TP_KENNEDY_GRUBER) {
            // F = (mn, mn ū + p̄ n̄, mn ē + p̄ ūn), mn = ρ̄ ūn
            const T rho = (Q1[0] + Q2[0]) * T(0.5);
            const T p = (p1 + p2) * T(0.5);
            const T e = (Q1[ND + 1] / Q1[0] + Q2[ND + 1] / Q2[0]) * T(0.5);
            const T mn = rho * umn;
            const T mnb = Gb[0] + Gmu + Gb[ND + 1] * e;
            const T pbar = Gmn + Gb[ND + 1] * umn;
            const T eb = Gb[ND + 1] * mn;
            const T rhob = mnb * umn;
            for (int k = 0; k < ND; ++k)
                ub1[k] = T(0.5) * (mn * Gb[1 + k] + (Gb[ND + 1] * p + mnb * rho) * n[k]);
            rb1 = T(0.5) * rhob;
            pb1 = T(0.5) * pbar;
            // e1 = E1/ρ1
            const T e1b = T(0.5) * eb;
            Qb1[ND + 1] += e1b / Q1[0];
            rb1 -= e1b * Q1[ND + 1] / (Q1[0] * Q1[0]);
        } else {  // TP_CHANDRASEKHAR
            const T b1 = Q1[0] / (T(2) * p1), b2 = Q2[0] / (T(2) * p2);
            T rho, beta;
            const T drho = logmean_da(Q1[0], Q2[0], &rho);
            const T dbeta = logmean_da(b1, b2, &beta);
            const T sb = b1 + b2;
            const T p = (Q1[0] + Q2[0]) / (T(2) * sb);
            T uu = T(0);
            for (int k = 0; k < ND; ++k) uu += um[k] * um[k];
            T sq = T(0);
            for (int k = 0; k < ND; ++k) sq += u1[k] * u1[k] + u2[k] * u2[k];
            const T h = T(1) / ((T(2) * beta) * g1) - sq / T(4) + p / rho + uu;
            const T mn = rho * umn;
            const T hb = Gb[ND + 1] * mn;
            const T mnb = Gb[0] + Gmu + Gb[ND + 1] * h;
            const T pbar = Gmn + hb / rho;
            const T rhob = mnb * umn - hb * p / (rho * rho);
            const T betab = -hb / (T(2) * beta * beta * g1);
            const T sqb = -hb / T(4);
            // ū̄ = mn Gm + mnb ρ n + 2 hb ū ; u1 gets half of it plus 2 sqb u1
            for (int k = 0; k < ND; ++k)
                ub1[k] = T(0.5) * (mn * Gb[1 + k] + mnb * rho * n[k] + T(2) * hb * um[k]) +
                         T(2) * sqb * u1[k];
            // p = (ρ1+ρ2)/(2(β1+β2))
            rb1 = rhob * drho + pbar / (T(2) * sb);
            T bb1 = betab * dbeta - pbar * p / sb;
            // β1 = ρ1/(2 p1)
            rb1 += bb1 / (T(2) * p1);
            pb1 = -bb1 * b1 / p1;
        }
        prim_pullback<T, ND>(Q1, gamma, rb1, ub1, pb1, Qb1);
    }
}

// logarithmic_mean(a, b) with both partial derivatives from one evaluation (the branch of
// the forward): ∂L/∂a = 1/(2F) - 2 f b F'/(F² s), ∂L/∂b = 1/(2F) + 2 f a F'/(F² s).  The
// derivative code has no reference operation order to keep, so it multiplies by
// reciprocals (two divisions on the series branch).
template <typename T>
FLOU_HD inline void logmean_dab(T a, T b, T* L, T* da, T* db) {
    const T s = a + b, is = rcp(s);
    const T f = (a - b) * is;
    const T u = f * f;
    T F, dF;
    if (u < T(0.01)) {
        F = T(1) + u * (T(1) / T(3) + u * (T(1) / T(5) + u * (T(1) / T(7))));
        dF = T(1) / T(3) + u * (T(2) / T(5) + u * (T(3) / T(7)));
    } else {
        F = log(a / b) / (T(2) * f);
        dF = (T(1) / (T(1) - u) - F) / (T(2) * u);
    }
    const T iF = rcp(F), h = T(0.5) * iF;
    *L = s * h;
    const T g = T(2) * f * dF * iF * iF * is;
    *da = h - g * b;
    *db = h + g * a;
}

// logmean_dab that also returns 1/(a+b) and F, for callers that reuse them
template <typename T>
FLOU_HD inline void logmean_dab2(T a, T b, T* L, T* da, T* db, T* is_out, T* F_out) {
    const T s = a + b, is = rcp(s);
    const T f = (a - b) * is;
    const T u = f * f;
    T F, dF;
    if (u < T(0.01)) {
        F = T(1) + u * (T(1) / T(3) + u * (T(1) / T(5) + u * (T(1) / T(7))));
        dF = T(1) / T(3) + u * (T(2) / T(5) + u * (T(3) / T(7)));
    } else {
        F = log(a / b) / (T(2) * f);
        dF = (T(1) / (T(1) - u) - F) / (T(2) * u);
    }
    const T iF = rcp(F), h = T(0.5) * iF;
    *L = s * h;
    const T g = T(2) * f * dF * iF * iF * is;
    *da = h - g * b;
    *db = h + g * a;
    *is_out = is;
    *F_out = F;
}

// prim_pullback with 1/ρ and the velocity u given
template <typename T, int ND>
FLOU_HD inline void prim_pullback_u(T ir, const T* u, T gamma, T rb, const T* ub, T pb, T* Qb) {
    T uu = T(0), ubu = T(0);
    for (int k = 0; k < ND; ++k) {
        uu += u[k] * u[k];
        ubu += ub[k] * u[k];
    }
    const T g1 = gamma - T(1);
    Qb[0] += rb - ubu * ir + g1 * pb * uu * T(0.5);
    for (int k = 0; k < ND; ++k) Qb[1 + k] += ub[k] * ir - g1 * pb * u[k];
    Qb[ND + 1] += g1 * pb;
}

// Per-node inputs of twopoint_vjp2p: R = (u_1 .. u_d, p, 1/ρ, 1/p, β = ρ/(2p)).
template <typename T, int ND>
FLOU_HD inline void node_prims(const T* Q, T gamma, T* R) {
    const T ir = rcp(Q[0]);
    T q = T(0);
    for (int k = 0; k < ND; ++k) {
        R[k] = Q[1 + k] * ir;
        q += Q[1 + k] * R[k];
    }
    const T p = (gamma - T(1)) * (Q[ND + 1] - T(0.5) * q);
    const T ip = rcp(p);
    R[ND] = p;
    R[ND + 1] = ir;
    R[ND + 2] = ip;
    R[ND + 3] = T(0.5) * Q[0] * ip;
}

// twopoint_vjp2 with the per-node primitives and reciprocals given (node_prims) and the
// pair's reciprocals shared: 1/(β1+β2) from the β log-mean, 1/ρ̂ = 2F_ρ/(ρ1+ρ2) and
// 1/β̂ = 2F_β/(β1+β2) from the log-mean factors (four divisions per pair on the series
// branch).  Equal to twopoint_vjp2 up to rounding.  PRIM (Chandrasekhar only): Qb1 and Qb2
// receive the cotangents of the primitives (ρ, u_1..u_d, p) instead, for a caller that
// sums them over all pairs of a node and applies prim_pullback_u once.
template <typename T, int ND, int TP, bool PRIM = false>
FLOU_HD inline void twopoint_vjp2p(const T* Q1, const T* R1, const T* Q2, const T* R2, const T* n, T gamma,
                                   T ig1, const T* Gb, T* Qb1, T* Qb2) {
    static_assert(!PRIM || TP == TP_CHANDRASEKHAR, "primitive cotangents: Chandrasekhar flux only");
    if constexpr (TP == TP_STDAVG) {
        T h[ND + 2];
        for (int v = 0; v < ND + 2; ++v) h[v] = T(0.5) * Gb[v];
        flux_vjp_ir<T, ND>(Q1, R1[ND + 1], n, gamma, h, Qb1);
        flux_vjp_ir<T, ND>(Q2, R2[ND + 1], n, gamma, h, Qb2);
        return;
    } else {
        const T* u1 = R1;
        const T* u2 = R2;
        const T ir1 = R1[ND + 1], ir2 = R2[ND + 1];
        const T p1 = R1[ND], p2 = R2[ND];
        T um[ND];
        T umn = T(0), Gmn = T(0), Gmu = T(0);
        for (int k = 0; k < ND; ++k) {
            um[k] = (u1[k] + u2[k]) * T(0.5);
            umn += um[k] * n[k];
            Gmn += Gb[1 + k] * n[k];
            Gmu += Gb[1 + k] * um[k];
        }
        T ub1[ND], ub2[ND];
        T rb1, rb2, pb1, pb2;
        if constexpr (TP == TP_KENNEDY_GRUBER) {
            const T rho = (Q1[0] + Q2[0]) * T(0.5);
            const T p = (p1 + p2) * T(0.5);
            const T e = (Q1[ND + 1] * ir1 + Q2[ND + 1] * ir2) * T(0.5);
            const T mn = rho * umn;
            const T mnb = Gb[0] + Gmu + Gb[ND + 1] * e;
            const T pbar = Gmn + Gb[ND + 1] * umn;
            const T eb = Gb[ND + 1] * mn;
            const T rhob = mnb * umn;
            for (int k = 0; k < ND; ++k) {
                ub1[k] = T(0.5) * (mn * Gb[1 + k] + (Gb[ND + 1] * p + mnb * rho) * n[k]);
                ub2[k] = ub1[k];
            }
            const T esb = T(0.5) * eb;
            rb1 = T(0.5) * rhob - esb * Q1[ND + 1] * ir1 * ir1;
            rb2 = T(0.5) * rhob - esb * Q2[ND + 1] * ir2 * ir2;
            pb1 = pb2 = T(0.5) * pbar;
            Qb1[ND + 1] += esb * ir1;
            Qb2[ND + 1] += esb * ir2;
        } else {  // TP_CHANDRASEKHAR
            const T ip1 = R1[ND + 2], ip2 = R2[ND + 2];
            const T b1 = R1[ND + 3], b2 = R2[ND + 3];
            T rho, beta, drho1, drho2, dbeta1, dbeta2, isr, Fr, isb, Fb;
            logmean_dab2(Q1[0], Q2[0], &rho, &drho1, &drho2, &isr, &Fr);
            logmean_dab2(b1, b2, &beta, &dbeta1, &dbeta2, &isb, &Fb);
            const T irho = T(2) * Fr * isr, ibeta = T(2) * Fb * isb;
            const T p = T(0.5) * (Q1[0] + Q2[0]) * isb;
            T uu = T(0), sq = T(0);
            for (int k = 0; k < ND; ++k) {
                uu += um[k] * um[k];
                sq += u1[k] * u1[k] + u2[k] * u2[k];
            }
            const T h = T(0.5) * ibeta * ig1 - T(0.25) * sq + p * irho + uu;
            const T mn = rho * umn;
            const T hb = Gb[ND + 1] * mn;
            const T mnb = Gb[0] + Gmu + Gb[ND + 1] * h;
            const T pbar = Gmn + hb * irho;
            const T rhob = mnb * umn - hb * p * irho * irho;
            const T betab = T(-0.5) * hb * ibeta * ibeta * ig1;
            const T sqb2 = T(-0.5) * hb;
            for (int k = 0; k < ND; ++k) {
                const T c = T(0.5) * (mn * Gb[1 + k] + mnb * rho * n[k] + T(2) * hb * um[k]);
                ub1[k] = c + sqb2 * u1[k];
                ub2[k] = c + sqb2 * u2[k];
            }
            const T rp = T(0.5) * pbar * isb, bp = pbar * p * isb;
            const T bb1 = betab * dbeta1 - bp, bb2 = betab * dbeta2 - bp;
            rb1 = rhob * drho1 + rp + T(0.5) * bb1 * ip1;
            rb2 = rhob * drho2 + rp + T(0.5) * bb2 * ip2;
            pb1 = -bb1 * b1 * ip1;
            pb2 = -bb2 * b2 * ip2;
        }
        if constexpr (PRIM) {
            Qb1[0] += rb1;
            Qb2[0] += rb2;
            for (int k = 0; k < ND; ++k) {
                Qb1[1 + k] += ub1[k];
                Qb2[1 + k] += ub2[k];
            }
            Qb1[ND + 1] += pb1;
            Qb2[ND + 1] += pb2;
        } else {
            prim_pullback_u<T, ND>(ir1, u1, gamma, rb1, ub1, pb1, Qb1);
            prim_pullback_u<T, ND>(ir2, u2, gamma, rb2, ub2, pb2, Qb2);
        }
    }
}

// Both sides of the two-point flux VJP at once: Qb1 += (∂F♯/∂Q1)^T Gb and
// Qb2 += (∂F♯/∂Q2)^T Gb, sharing the symmetric intermediates (means, log-means, p̂, ĥ, the
// mass flux and their cotangents) and multiplying by reuses of a few reciprocals; equal to
// twopoint_vjp1(Q1, Q2) and twopoint_vjp1(Q2, Q1) up to rounding.
template <typename T, int ND, int TP>
FLOU_HD inline void twopoint_vjp2(const T* Q1, const T* Q2, const T* n, T gamma, const T* Gb, T* Qb1,
                                  T* Qb2) {
    if constexpr (TP == TP_STDAVG) {
        T h[ND + 2];
        for (int v = 0; v < ND + 2; ++v) h[v] = T(0.5) * Gb[v];
        flux_vjp<T, ND>(Q1, n, gamma, h, Qb1);
        flux_vjp<T, ND>(Q2, n, gamma, h, Qb2);
        return;
    } else {
        const T g1 = gamma - T(1);
        const T ir1 = T(1) / Q1[0], ir2 = T(1) / Q2[0];
        T u1[ND], u2[ND], um[ND];
        T q1 = T(0), q2 = T(0);
        for (int k = 0; k < ND; ++k) {
            u1[k] = Q1[1 + k] * ir1;
            u2[k] = Q2[1 + k] * ir2;
            um[k] = (u1[k] + u2[k]) * T(0.5);
            q1 += Q1[1 + k] * u1[k];
            q2 += Q2[1 + k] * u2[k];
        }
        const T p1 = g1 * (Q1[ND + 1] - T(0.5) * q1);
        const T p2 = g1 * (Q2[ND + 1] - T(0.5) * q2);
        T umn = T(0), Gmn = T(0), Gmu = T(0);
        for (int k = 0; k < ND; ++k) {
            umn += um[k] * n[k];
            Gmn += Gb[1 + k] * n[k];
            Gmu += Gb[1 + k] * um[k];
        }
        T ub1[ND], ub2[ND];
        T rb1, rb2, pb1, pb2;
        if constexpr (TP == TP_KENNEDY_GRUBER) {
            const T rho = (Q1[0] + Q2[0]) * T(0.5);
            const T p = (p1 + p2) * T(0.5);
            const T e = (Q1[ND + 1] * ir1 + Q2[ND + 1] * ir2) * T(0.5);
            const T mn = rho * umn;
            const T mnb = Gb[0] + Gmu + Gb[ND + 1] * e;
            const T pbar = Gmn + Gb[ND + 1] * umn;
            const T eb = Gb[ND + 1] * mn;
            const T rhob = mnb * umn;
            for (int k = 0; k < ND; ++k) {
                ub1[k] = T(0.5) * (mn * Gb[1 + k] + (Gb[ND + 1] * p + mnb * rho) * n[k]);
                ub2[k] = ub1[k];
            }
            const T esb = T(0.5) * eb;
            rb1 = T(0.5) * rhob - esb * Q1[ND + 1] * ir1 * ir1;
            rb2 = T(0.5) * rhob - esb * Q2[ND + 1] * ir2 * ir2;
            pb1 = pb2 = T(0.5) * pbar;
            Qb1[ND + 1] += esb * ir1;
            Qb2[ND + 1] += esb * ir2;
        } else {  // TP_CHANDRASEKHAR
            const T ip1 = T(1) / p1, ip2 = T(1) / p2;
            const T b1 = T(0.5) * Q1[0] * ip1, b2 = T(0.5) * Q2[0] * ip2;
            T rho, beta, drho1, drho2, dbeta1, dbeta2;
            logmean_dab(Q1[0], Q2[0], &rho, &drho1, &drho2);
            logmean_dab(b1, b2, &beta, &dbeta1, &dbeta2);
            const T sb = b1 + b2, isb = T(1) / sb;
            const T irho = T(1) / rho, ibeta = T(1) / beta, ig1 = T(1) / g1;
            const T p = T(0.5) * (Q1[0] + Q2[0]) * isb;
            T uu = T(0), sq = T(0);
            for (int k = 0; k < ND; ++k) {
                uu += um[k] * um[k];
                sq += u1[k] * u1[k] + u2[k] * u2[k];
            }
            const T h = T(0.5) * ibeta * ig1 - T(0.25) * sq + p * irho + uu;
            const T mn = rho * umn;
            const T hb = Gb[ND + 1] * mn;
            const T mnb = Gb[0] + Gmu + Gb[ND + 1] * h;
            const T pbar = Gmn + hb * irho;
            const T rhob = mnb * umn - hb * p * irho * irho;
            const T betab = T(-0.5) * hb * ibeta * ibeta * ig1;
            const T sqb2 = T(-0.5) * hb;
            for (int k = 0; k < ND; ++k) {
                const T c = T(0.5) * (mn * Gb[1 + k] + mnb * rho * n[k] + T(2) * hb * um[k]);
                ub1[k] = c + sqb2 * u1[k];
                ub2[k] = c + sqb2 * u2[k];
            }
            const T rp = T(0.5) * pbar * isb, bp = pbar * p * isb;
            const T bb1 = betab * dbeta1 - bp, bb2 = betab * dbeta2 - bp;
            rb1 = rhob * drho1 + rp + T(0.5) * bb1 * ip1;
            rb2 = rhob * drho2 + rp + T(0.5) * bb2 * ip2;
            pb1 = -bb1 * b1 * ip1;
            pb2 = -bb2 * b2 * ip2;
        }
        prim_pullback_ir<T, ND>(Q1, ir1, gamma, rb1, ub1, pb1, Qb1);
        prim_pullback_ir<T, ND>(Q2, ir2, gamma, rb2, ub2, pb2, Qb2);
    }
}

// LxF(StdAverage) / StdAverage numerical flux in the face frame, normal = first axis
// (Equations/Euler.jl:99-165): Rlb += (∂Φ/∂Rl)^T Fb, Rrb += (∂Φ/∂Rr)^T Fb.
template <typename T, int ND>
FLOU_HD inline void numflux_vjp(const T* Rl, const T* Rr, bool lxf, T intensity, T gamma,
                                const T* Fb, T* Rlb, T* Rrb) {
    constexpr int NV = ND + 2;
    T e1[ND];
    for (int k = 0; k < ND; ++k) e1[k] = k == 0 ? T(1) : T(0);
    T h[NV];
    for (int v = 0; v < NV; ++v) h[v] = T(0.5) * Fb[v];
    const T ir[2] = {rcp(Rl[0]), rcp(Rr[0])};
    flux_vjp_ir<T, ND>(Rl, ir[0], e1, gamma, h, Rlb);
    flux_vjp_ir<T, ND>(Rr, ir[1], e1, gamma, h, Rrb);
    if (!lxf) return;
    const T g1 = gamma - T(1);
    T q[2], a[2], un[2];
    for (int s = 0; s < 2; ++s) {
        const T* R = s ? Rr : Rl;
        T m2 = T(0);
        for (int k = 0; k < ND; ++k) m2 += R[1 + k] * R[1 + k];
        const T p = g1 * (R[ND + 1] - T(0.5) * m2 * ir[s]);
        un[s] = R[1] * ir[s];
        a[s] = sqrt(gamma * p * ir[s]);
        q[s] = fabs(un[s]) + a[s];
    }
    const int w = q[0] < q[1] ? 1 : 0;  // the side the forward max keeps (left on ties)
    const T lam = q[w];
    T s = T(0);
    for (int v = 0; v < NV; ++v) {
        Rlb[v] += T(0.5) * intensity * lam * Fb[v];
        Rrb[v] -= T(0.5) * intensity * lam * Fb[v];
        s += (Rl[v] - Rr[v]) * Fb[v];
    }
    const T lb = T(0.5) * intensity * s;  // cotangent of λ
    const T* R = w ? Rr : Rl;
    T* Rb = w ? Rrb : Rlb;
    // q = |u_n| + sqrt(γ p / ρ)
    T ub[ND];
    for (int k = 0; k < ND; ++k) ub[k] = T(0);
    ub[0] = (un[w] > T(0) ? T(1) : (un[w] < T(0) ? T(-1) : T(0))) * lb;
    const T ab = lb;
    const T pb = T(0.5) * ab * gamma * ir[w] * rcp(a[w]);
    const T rb = T(-0.5) * ab * a[w] * ir[w];
    prim_pullback_ir<T, ND>(R, ir[w], gamma, rb, ub, pb, Rb);
}

}  // namespace adj
}  // namespace flouhip
