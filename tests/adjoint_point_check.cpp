// CPU check of the point-wise reverse-mode products of flou.jl_amd/csrc/adjoint.hpp
// against forward-mode dual-number JVPs of the oracle (oracle/flou_oracle.hpp): for random
// states, directions V and cotangents G, <G, J V> (oracle) must equal <J^T G, V> (adjoint)
// to round-off (the two-sided two-point VJP with both states perturbed at once).  Prints the
// worst relative mismatch per function, dimension and spread;
// built and run by tests/test_adjoint_pointwise.py.
#include <cstdio>
#include <random>
#include "flou_oracle.hpp"
namespace flouhip {
enum { TP_STDAVG = 0, TP_CHANDRASEKHAR = 1, TP_KENNEDY_GRUBER = 2 };
}
#include "adjoint.hpp"
using D1 = oracle::Dual<1>;
std::mt19937_64 rng(7);
double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
template <int ND> void rand_state(double* Q, double spread) {
    double rho = 1.0 + spread * U(-1, 1), p = 1.0 + spread * U(-1, 1);
    double u[3] = {0.3 * U(-1, 1), 0.3 * U(-1, 1), 0.3 * U(-1, 1)};
    Q[0] = rho; double k = 0;
    for (int i = 0; i < ND; ++i) { Q[1 + i] = rho * u[i]; k += u[i] * u[i]; }
    Q[ND + 1] = p / 0.4 + 0.5 * rho * k;
}
template <int ND> double run(int tp, double spread, int which) {
    constexpr int NV = ND + 2;
    oracle::Disc d; d.nd = ND; d.nv = NV; d.eq = oracle::EQ_EULER; d.gamma = 1.4; d.tpflux = tp;
    d.savg = oracle::AVG_STD; d.sdiss = oracle::DISS_LXF; d.intensity = 0.7;
    double Q1[NV], Q2[NV], V[NV], G[NV], M[9], n[3];
    rand_state<ND>(Q1, spread); rand_state<ND>(Q2, spread);
    for (int v = 0; v < NV; ++v) { V[v] = U(-1, 1); G[v] = U(-1, 1); }
    for (int i = 0; i < ND * ND; ++i) M[i] = U(-1, 1);
    int dir = ND - 1;
    for (int k = 0; k < ND; ++k) n[k] = M[k + ND * dir];
    D1 q1[NV], q2[NV], F[NV];
    for (int v = 0; v < NV; ++v) { q1[v] = D1(Q1[v]); q1[v].d[0] = V[v]; q2[v] = D1(Q2[v]); }
    double Qb[NV] = {0}, Qb2[NV] = {0};
    if (which == 0) {  // physical flux along a metric column
        oracle::contravariant_flux(d, q1, M, dir, F);
        flouhip::adj::flux_vjp<double, ND>(Q1, n, 1.4, G, Qb);
    } else if (which == 1) {  // two-point flux, derivative in Q1
        oracle::twopointflux(d, q1, q2, M + ND * dir, M + ND * dir, F);
        if (tp == 0) flouhip::adj::twopoint_vjp1<double, ND, 0>(Q1, Q2, n, 1.4, G, Qb);
        if (tp == 1) flouhip::adj::twopoint_vjp1<double, ND, 1>(Q1, Q2, n, 1.4, G, Qb);
        if (tp == 2) flouhip::adj::twopoint_vjp1<double, ND, 2>(Q1, Q2, n, 1.4, G, Qb);
    } else if (which == 4 || which == 5) {  // two-point flux, both states at once
        double V2[NV];
        for (int v = 0; v < NV; ++v) { V2[v] = U(-1, 1); q2[v].d[0] = V2[v]; }
        oracle::twopointflux(d, q1, q2, M + ND * dir, M + ND * dir, F);
        if (which == 4) {  // twopoint_vjp2
            if (tp == 0) flouhip::adj::twopoint_vjp2<double, ND, 0>(Q1, Q2, n, 1.4, G, Qb, Qb2);
            if (tp == 1) flouhip::adj::twopoint_vjp2<double, ND, 1>(Q1, Q2, n, 1.4, G, Qb, Qb2);
            if (tp == 2) flouhip::adj::twopoint_vjp2<double, ND, 2>(Q1, Q2, n, 1.4, G, Qb, Qb2);
        } else {  // twopoint_vjp2p on node_prims
            double R1[ND + 4], R2[ND + 4];
            flouhip::adj::node_prims<double, ND>(Q1, 1.4, R1);
            flouhip::adj::node_prims<double, ND>(Q2, 1.4, R2);
            const double ig1 = 1.0 / 0.4;
            if (tp == 0) flouhip::adj::twopoint_vjp2p<double, ND, 0>(Q1, R1, Q2, R2, n, 1.4, ig1, G, Qb, Qb2);
            if (tp == 1) flouhip::adj::twopoint_vjp2p<double, ND, 1>(Q1, R1, Q2, R2, n, 1.4, ig1, G, Qb, Qb2);
            if (tp == 2) flouhip::adj::twopoint_vjp2p<double, ND, 2>(Q1, R1, Q2, R2, n, 1.4, ig1, G, Qb, Qb2);
        }
        double jvp = 0, vjp = 0, sc = 0;
        for (int v = 0; v < NV; ++v) {
            jvp += G[v] * F[v].d[0];
            vjp += V[v] * Qb[v] + V2[v] * Qb2[v];
            sc += fabs(G[v] * F[v].d[0]);
        }
        return fabs(jvp - vjp) / (sc + 1e-300);
    } else if (which == 3) {  // LxF, derivative in the right state
        D1 r1[NV], r2[NV];
        for (int v = 0; v < NV; ++v) { r1[v] = D1(Q1[v]); r2[v] = D1(Q2[v]); r2[v].d[0] = V[v]; }
        oracle::euler_numericalflux(d, r1, r2, F);
        double Qa[NV] = {0};
        flouhip::adj::numflux_vjp<double, ND>(Q1, Q2, true, 0.7, 1.4, G, Qa, Qb);
    } else {  // LxF in the face frame, derivative in the left state
        D1 r2[NV];
        for (int v = 0; v < NV; ++v) { r2[v] = D1(Q2[v]); r2[v].d[0] = 0.0; }
        oracle::euler_numericalflux(d, q1, r2, F);
        flouhip::adj::numflux_vjp<double, ND>(Q1, Q2, true, 0.7, 1.4, G, Qb, Qb2);
    }
    double jvp = 0, vjp = 0, sc = 0;
    for (int v = 0; v < NV; ++v) { jvp += G[v] * F[v].d[0]; vjp += V[v] * Qb[v]; sc += fabs(G[v] * F[v].d[0]); }
    return fabs(jvp - vjp) / (sc + 1e-300);
}
int main() {
    const char* nm[] = {"flux", "twopoint", "lxf-left", "lxf-right", "twopoint-both", "twopoint-both-prims"};
    for (int which = 0; which < 6; ++which)
        for (int tp = 0; tp < 3; ++tp) {
            if (which != 1 && which < 4 && tp > 0) continue;
            for (double spread : {1e-4, 0.05, 0.5}) {
                double e1 = 0, e2 = 0, e3 = 0;
                for (int t = 0; t < 200; ++t) {
                    e1 = std::max(e1, run<1>(tp, spread, which));
                    e2 = std::max(e2, run<2>(tp, spread, which));
                    e3 = std::max(e3, run<3>(tp, spread, which));
                }
                printf("%s tp=%d spread=%g %.3e %.3e %.3e\n", nm[which], tp, spread, e1, e2, e3);
            }
        }
}
