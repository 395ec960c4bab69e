// Flou.jl DGSEM oracle -- CPU restatement of the reference algorithm.
//
// TEST INFRASTRUCTURE ONLY.  This file is what the HIP path is compared with: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
// library (libflouhip.so) never links or calls it.
//
// Everything here restates the reference operation for operation, in the reference's
// evaluation order (left folds for n-ary sums, `a / 2b` read as a / (2*b), `x^2` as x*x,
// `x^3` as (x*x)*x, no FMA contraction: build with -ffp-contract=off).  Citations are to
// files of Andres-MG/Flou.jl v0.1.0.
//
// Templated on the state scalar T so that a forward-mode dual number gives directional
// derivatives of rhs! and of whole RK runs (the ForwardDiff analogue used to pin the
// adjoint).  Geometry and operators stay double.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

// ------------------------------------------------------------------------------------
// Scalar helpers (overloaded for Dual below)
inline double val(double x) { return x; }
inline double osqrt(double x) {
    // Julia's sqrt throws DomainError on negative arguments (FlouTime.jl:39-51 catches it).
    if (x < 0.0) throw std::domain_error("sqrt of negative number");
    return std::sqrt(x);
}
inline double olog(double x) {
    if (x < 0.0) throw std::domain_error("log of negative number");
    return std::log(x);
}
inline double oabs(double x) { return std::fabs(x); }
inline double oexp(double x) { return std::exp(x); }
// Julia's x^y for a real, non-integer y throws DomainError when x < 0
inline double opow(double x, double y) {
    if (x < 0.0) throw std::domain_error("pow of negative number");
    return std::pow(x, y);
}
inline double omax(double a, double b) { return (a < b) ? b : a; }
inline double omin(double a, double b) { return (b < a) ? b : a; }

// Forward-mode dual number with K directions (ForwardDiff.Dual analogue).
template <int K>
struct Dual {
    double v;
    double d[K];
    Dual() : v(0.0) { for (int k = 0; k < K; ++k) d[k] = 0.0; }
    Dual(double x) : v(x) { for (int k = 0; k < K; ++k) d[k] = 0.0; }
};
template <int K> inline double val(const Dual<K>& x) { return x.v; }
template <int K> inline Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v + b.v; for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k]; return r; }
template <int K> inline Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v - b.v; for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k]; return r; }
template <int K> inline Dual<K> operator-(const Dual<K>& a) {
    Dual<K> r; r.v = -a.v; for (int k = 0; k < K; ++k) r.d[k] = -a.d[k]; return r; }
template <int K> inline Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v * b.v; for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r; }
template <int K> inline Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v / b.v;
    for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r; }
template <int K> inline Dual<K> operator+(const Dual<K>& a, double b) { return a + Dual<K>(b); }
template <int K> inline Dual<K> operator+(double a, const Dual<K>& b) { return Dual<K>(a) + b; }
template <int K> inline Dual<K> operator-(const Dual<K>& a, double b) { return a - Dual<K>(b); }
template <int K> inline Dual<K> operator-(double a, const Dual<K>& b) { return Dual<K>(a) - b; }
template <int K> inline Dual<K> operator*(const Dual<K>& a, double b) {
    Dual<K> r; r.v = a.v * b; for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b; return r; }
template <int K> inline Dual<K> operator*(double a, const Dual<K>& b) { return b * a; }
template <int K> inline Dual<K> operator/(const Dual<K>& a, double b) {
    Dual<K> r; r.v = a.v / b; for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / b; return r; }
template <int K> inline Dual<K> operator/(double a, const Dual<K>& b) { return Dual<K>(a) / b; }
template <int K> inline Dual<K>& operator+=(Dual<K>& a, const Dual<K>& b) { a = a + b; return a; }
template <int K> inline Dual<K>& operator-=(Dual<K>& a, const Dual<K>& b) { a = a - b; return a; }
template <int K> inline bool operator<(const Dual<K>& a, const Dual<K>& b) { return a.v < b.v; }
template <int K> inline bool operator<(const Dual<K>& a, double b) { return a.v < b; }
template <int K> inline Dual<K> osqrt(const Dual<K>& a) {
    Dual<K> r; r.v = osqrt(a.v); for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / (2.0 * r.v); return r; }
template <int K> inline Dual<K> oexp(const Dual<K>& a) {
    Dual<K> r; r.v = oexp(a.v); for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * r.v; return r; }
template <int K> inline Dual<K> opow(const Dual<K>& a, double y) {
    Dual<K> r; r.v = opow(a.v, y); const double g = y * opow(a.v, y - 1.0);
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * g; return r; }
template <int K> inline Dual<K> olog(const Dual<K>& a) {
    Dual<K> r; r.v = olog(a.v); for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / a.v; return r; }
template <int K> inline Dual<K> oabs(const Dual<K>& a) { return (a.v < 0.0) ? -a : a; }
template <int K> inline Dual<K> omax(const Dual<K>& a, const Dual<K>& b) { return (a.v < b.v) ? b : a; }
template <int K> inline Dual<K> omin(const Dual<K>& a, const Dual<K>& b) { return (b.v < a.v) ? b : a; }

// ------------------------------------------------------------------------------------
// Enumerations (shared numbering with the public header's enums)
enum { EQ_ADVECTION = 0, EQ_EULER = 1 };
enum { VOL_STRONG = 0, VOL_SPLIT = 1, VOL_HYBRID = 2 };
enum { TP_STDAVG = 0, TP_CHANDRASEKHAR = 1, TP_KENNEDY_GRUBER = 2 };
enum { AVG_STD = 0, AVG_CHANDRASEKHAR = 1 };
enum { DISS_NONE = 0, DISS_LXF = 1, DISS_MATRIX = 2, DISS_SCALAR = 3 };
enum { BC_DIRICHLET = 0, BC_INFLOW = 1, BC_OUTFLOW = 2, BC_SLIP = 3 };
enum { NODES_GLL = 0, NODES_GL = 1 };
enum { RK_SSPRK33 = 0, RK_CK2N54 = 1, RK_ORK256 = 2 };

// ------------------------------------------------------------------------------------
// 1-D operators.  Row-major n x n matrices: M[i*n + j].
struct Ops {
    int n = 0;
    std::vector<double> xi, w, D, Ds, Dsh, lm, lp, gm, gp;
    std::vector<long double> xl, bw;  // nodes and barycentric weights in long double
};

// Nodes, weights, Lagrange derivative and face interpolation, computed in long double
// (barycentric, exact up to the final rounding) -- the reference builds them from monomial
// fits (StdRegions/ApproximationBases.jl:154-186); the reconstruction vectors and the
// Ds / D# combinations are then formed in double exactly as StdSegment.jl:76-89 and
// Reconstruction.jl:125-134 do.
Ops build_ops(int nodetype, int n);

// ------------------------------------------------------------------------------------
// Output to equispaced points (src/FlouSpatial/IO.jl:78-121)

// A[i][j] = l_j(xe[i]) (m x n, row-major): interp_matrix(ξe, basis)
// (StdRegions/ApproximationBases.jl:85-87, StdSegment.jl:56-58), barycentric in long double.
void interp_matrix(const Ops& o, int m, const double* xe, double* A);

// node2eq = A ⊗ ... ⊗ A as kron builds it (StdQuad.jl:53, StdHex.jl:61): entry
// (i, j) = (A[i3][j3] * A[i2][j2]) * A[i1][j1] with index 1 fastest; then per element and
// column Qe = node2eq * Q_e with j ascending (StdRegions.jl:232-238, IO.jl:86-93).
void equispaced(int nd, int n, int m, const double* A, int64_t ne, int ncols, const double* Q,
                int64_t ldq, double* Qe, int64_t lde);

// vtk_connectivities(std): the reference's listing of a cell's points in VTK Lagrange order
// -- corners, edges, faces, interior -- as 0-based lattice indices (x fastest) of an
// n^nd lattice (StdSegment.jl:171-175, StdQuad.jl:185-196, StdHex.jl:174-199).
std::vector<int64_t> vtk_connectivities(int nd, int n);

// ------------------------------------------------------------------------------------
// Discretisation: Cartesian mesh (CartesianMesh.jl) + periodic rewiring (Mesh.jl:236-322)
// + Cartesian geometry (PhysicalRegions.jl:370-408, 541-696).
struct Disc {
    int nd = 0, n = 0, nv = 0;
    long ne = 0, nf = 0;
    int npe = 0, npf = 0;
    int nel[3] = {1, 1, 1};
    double start[3] = {0, 0, 0}, finish[3] = {0, 0, 0}, dx[3] = {0, 0, 0};
    Ops ops;
    // 0-based connectivity
    std::vector<long> e_face;   // ne * 2nd
    std::vector<int> e_side;    // ne * 2nd : 0 = master side, 1 = slave side
    std::vector<long> f_elem;   // nf * 2  : -1 on the exterior side of a boundary face
    std::vector<int> f_pos;     // nf * 2  : 0-based face position inside the element
    std::vector<int> f_orient;  // nf
    std::vector<int> f_bd;      // nf : boundary index (0-based, after periodic removal) or -1
    std::vector<std::string> bdnames;
    // geometry
    std::vector<double> vtx;     // mesh vertex coordinates, nvtx * nd
    std::vector<double> e_jac;   // ne
    std::vector<double> e_metric;// ne * nd * nd, column-major: m[r + nd*c] = metric[r, c]
    std::vector<double> e_vol;   // ne
    std::vector<double> f_frame; // nf * 3 * nd : (n, t, b)
    std::vector<double> f_jac;   // nf
    // curved (per-node) geometry, used instead of the per-element/per-face arrays above
    bool curved = false;
    std::vector<double> n_jac;     // ndofs
    std::vector<double> n_metric;  // ndofs * nd * nd, column-major per node
    std::vector<double> fn_frame;  // nf * npf * 3 * nd, per face node of the master side
    std::vector<double> fn_jac;    // nf * npf
    const double* frame(long face, int fnode) const {
        return curved ? &fn_frame[((size_t)face * npf + fnode) * 3 * nd] : &f_frame[(size_t)face * 3 * nd];
    }
    double facejac(long face, int fnode) const { return curved ? fn_jac[(size_t)face * npf + fnode] : f_jac[face]; }
    const double* metric(long e, int i) const {
        return curved ? &n_metric[((size_t)e * npe + i) * nd * nd] : &e_metric[(size_t)e * nd * nd];
    }
    double jac(long e, int i) const { return curved ? n_jac[(size_t)e * npe + i] : e_jac[e]; }
    // stage limiter (Zhang & Shu) applied by rk_run after every stage
    bool limiter_on = false;
    double limiter_minval = 0.0;
    std::vector<double> coords;  // ndofs * nd node coordinates
    std::vector<double> f_coords;// nf * npf * nd face node coordinates
    // equation and operators
    int eq = EQ_EULER;
    double gamma = 1.4;
    double a[3] = {0, 0, 0};
    int volop = VOL_STRONG, tpflux = TP_STDAVG, savg = AVG_STD, sdiss = DISS_LXF;
    double intensity = 1.0;
    double blend = 1.0;                  // HybridDivOperator blending constant c
    int chandrasekhar_surface_typo = 1;  // reproduce Equations/Euler.jl:216 (3-D only)
    // boundary conditions per remaining boundary index
    std::vector<int> bctype;
    std::vector<std::vector<double>> bcdata;
    std::vector<double> f_bcstate;  // nf * npf * nv exterior states of Dirichlet faces

    long ndofs() const { return ne * npe; }
};

// Build a Cartesian mesh.  periodic[d] != 0 pairs boundaries (2d+1, 2d+2) (1-based ids),
// applied in the order d = 0, 1, 2.
void build_cartesian(Disc& disc, int nd, const int* nel, const double* start,
                     const double* finish, const int* periodic);

// Tensor-product helpers: node index of position `a` on line `f` in direction `dir`.
inline int line_node(int nd, int n, int dir, int f, int a) {
    if (nd == 1) return a;
    if (nd == 2) return (dir == 0) ? (a + n * f) : (f + n * a);
    int fi = f % n, fj = f / n;
    if (dir == 0) return a + n * (fi + n * fj);
    if (dir == 1) return fi + n * (a + n * fj);
    return fi + n * (fj + n * a);
}

// Face-node maps between the sides of a face with orientation `o` (0-based indices; quad
// faces numbered with the first index fastest, as lineardofs/cartesiandofs do):
// master2slave and slave2master of StdSegment.jl:153-167 and StdQuad.jl:138-181.
inline int quad_li(int n, int x, int y) { return (x - 1) + n * (y - 1); }  // 1-based (x, y)
inline int master2slave(int nd, int n, int i, int o) {
    if (o == 0 || nd == 1) return i;
    if (nd == 2) return n - 1 - i;  // StdSegment: ndofs - (i - 1), 1-based
    const int m1 = i % n + 1, m2 = i / n + 1;
    switch (o) {
        case 1: return quad_li(n, m2, n - m1 + 1);
        case 2: return quad_li(n, n - m1 + 1, n - m2 + 1);
        case 3: return quad_li(n, n - m2 + 1, m1);
        case 4: return quad_li(n, m2, m1);
        case 5: return quad_li(n, n - m1 + 1, m2);
        case 6: return quad_li(n, n - m2 + 1, n - m1 + 1);
        default: return quad_li(n, m1, n - m2 + 1);
    }
}
inline int slave2master(int nd, int n, int i, int o) {
    if (o == 0 || nd == 1) return i;
    if (nd == 2) return n - 1 - i;
    const int s1 = i % n + 1, s2 = i / n + 1;
    switch (o) {
        case 1: return quad_li(n, n - s2 + 1, s1);
        case 2: return quad_li(n, n - s1 + 1, n - s2 + 1);
        case 3: return quad_li(n, s2, n - s1 + 1);
        case 4: return quad_li(n, s2, s1);
        case 5: return quad_li(n, n - s1 + 1, s2);
        case 6: return quad_li(n, n - s2 + 1, n - s1 + 1);
        default: return quad_li(n, s1, n - s2 + 1);
    }
}

// ------------------------------------------------------------------------------------
// Point-wise physics (FlouCommon/Euler.jl, FlouCommon/LinearAdvection.jl)
template <class T>
inline T pressure(const T* Q, int nd, double gamma) {
    // (γ-1) * (ρe - (ρu^2 + ρv^2 + ρw^2) / 2ρ)      Euler.jl:142-155
    T s = Q[1] * Q[1];
    for (int k = 1; k < nd; ++k) s = s + Q[1 + k] * Q[1 + k];
    return (gamma - 1.0) * (Q[nd + 1] - s / (2.0 * Q[0]));
}

template <class T>
inline void cons2prim(const T* Q, int nd, double gamma, T* P) {
    // (ρ, ρu/ρ, ..., p)      Euler.jl:237-253
    T p = pressure(Q, nd, gamma);
    P[0] = Q[0];
    for (int k = 0; k < nd; ++k) P[1 + k] = Q[1 + k] / Q[0];
    P[nd + 1] = p;
}

template <class T>
inline T soundvelocity(const T& rho, const T& p, double gamma) {
    return osqrt(gamma * p / rho);  // Euler.jl:225-235
}

template <class T>
inline T logarithmic_mean(const T& al, const T& ar) {
    // FlouCommon/Utilities.jl:34-44
    T xi = al / ar;
    T f = (xi - 1.0) / (xi + 1.0);
    T u = f * f;
    T F;
    if (u < 0.01) {
        F = ((1.0 + u / 3.0) + (u * u) / 5.0) + ((u * u) * u) / 7.0;
    } else {
        F = olog(xi) / (2.0 * f);
    }
    return (al + ar) / (2.0 * F);
}

// Physical volume flux in direction k (Euler.jl:54-114; LinearAdvection.jl:42-44)
template <class T>
inline void volumeflux(const Disc& d, const T* Q, int k, T* F) {
    if (d.eq == EQ_ADVECTION) {
        F[0] = d.a[k] * Q[0];
        return;
    }
    const int nd = d.nd;
    T u[3];
    for (int m = 0; m < nd; ++m) u[m] = Q[1 + m] / Q[0];
    T p = pressure(Q, nd, d.gamma);
    F[0] = Q[1 + k];
    for (int m = 0; m < nd; ++m) {
        F[1 + m] = Q[1 + k] * u[m];
        if (m == k) F[1 + m] = F[1 + m] + p;
    }
    F[nd + 1] = (Q[nd + 1] + p) * u[k];
}

// contravariant(F, metric, dir) = Σ_k F_k * metric[k, dir]   (PhysicalRegions.jl:1001-1033)
template <class T>
inline void contravariant_flux(const Disc& d, const T* Q, const double* M, int dir, T* Ft) {
    const int nd = d.nd, nv = d.nv;
    T F[3][5];
    for (int k = 0; k < nd; ++k) volumeflux(d, Q, k, F[k]);
    for (int v = 0; v < nv; ++v) {
        T s = F[0][v] * M[0 + nd * dir];
        for (int k = 1; k < nd; ++k) s = s + F[k][v] * M[k + nd * dir];
        Ft[v] = s;
    }
}

// Two-point volume fluxes (Equations/Euler.jl:385-536; Equations/LinearAdvection.jl:44-47)
template <class T>
inline void twopointflux(const Disc& d, const T* Q1, const T* Q2, const double* Ja1,
                         const double* Ja2, T* F) {
    const int nd = d.nd, nv = d.nv;
    double nrm[3];
    for (int k = 0; k < nd; ++k) nrm[k] = (Ja1[k] + Ja2[k]) / 2.0;
    if (d.eq == EQ_ADVECTION) {
        // numericalflux(StdAverage): an * (Q1 + Q2) / 2, an = dot(a, n)
        double an = d.a[0] * nrm[0];
        for (int k = 1; k < nd; ++k) an = an + d.a[k] * nrm[k];
        F[0] = (an * (Q1[0] + Q2[0])) / 2.0;
        return;
    }
    T P1[5], P2[5];
    cons2prim(Q1, nd, d.gamma, P1);
    cons2prim(Q2, nd, d.gamma, P2);
    T f[5][3];  // f[v][k]: component v in physical direction k
    if (d.tpflux == TP_STDAVG) {
        const T p1 = P1[nd + 1], p2 = P2[nd + 1];
        for (int k = 0; k < nd; ++k) {
            f[0][k] = (Q1[1 + k] + Q2[1 + k]) / 2.0;
            for (int c = 0; c < nd; ++c) {
                T s = Q1[1 + c] * P1[1 + k];
                if (c == k) s = s + p1;
                s = s + Q2[1 + c] * P2[1 + k];
                if (c == k) s = s + p2;
                f[1 + c][k] = s / 2.0;
            }
            f[nd + 1][k] = ((Q1[nd + 1] + p1) * P1[1 + k] + (Q2[nd + 1] + p2) * P2[1 + k]) / 2.0;
        }
    } else if (d.tpflux == TP_CHANDRASEKHAR) {
        const T rho1 = P1[0], rho2 = P2[0], p1 = P1[nd + 1], p2 = P2[nd + 1];
        T um[3];
        for (int k = 0; k < nd; ++k) um[k] = (P1[1 + k] + P2[1 + k]) / 2.0;
        T b1 = rho1 / (2.0 * p1), b2 = rho2 / (2.0 * p2);
        T rho = logarithmic_mean(rho1, rho2);
        T p = (rho1 + rho2) / (2.0 * (b1 + b2));
        T beta = logarithmic_mean(b1, b2);
        T sq = P1[1] * P1[1];
        for (int k = 1; k < nd; ++k) sq = sq + P1[1 + k] * P1[1 + k];
        for (int k = 0; k < nd; ++k) sq = sq + P2[1 + k] * P2[1 + k];
        T h = 1.0 / ((2.0 * beta) * (d.gamma - 1.0)) - sq / 4.0;
        h = h + p / rho;
        for (int k = 0; k < nd; ++k) h = h + um[k] * um[k];
        for (int k = 0; k < nd; ++k) {
            f[0][k] = rho * um[k];
            for (int c = 0; c < nd; ++c) {
                if (c == k) f[1 + c][k] = rho * (um[k] * um[k]) + p;
                else f[1 + c][k] = (rho * um[std::min(c, k)]) * um[std::max(c, k)];
            }
            f[nd + 1][k] = (rho * um[k]) * h;
        }
    } else {  // TP_KENNEDY_GRUBER (new: Kennedy & Gruber, JCP 227 (2008) 1676-1700)
        const T p1 = P1[nd + 1], p2 = P2[nd + 1];
        T rho = (P1[0] + P2[0]) / 2.0;
        T um[3];
        for (int k = 0; k < nd; ++k) um[k] = (P1[1 + k] + P2[1 + k]) / 2.0;
        T p = (p1 + p2) / 2.0;
        T e = (Q1[nd + 1] / Q1[0] + Q2[nd + 1] / Q2[0]) / 2.0;
        for (int k = 0; k < nd; ++k) {
            T m = rho * um[k];
            f[0][k] = m;
            for (int c = 0; c < nd; ++c) {
                f[1 + c][k] = m * um[c];
                if (c == k) f[1 + c][k] = f[1 + c][k] + p;
            }
            f[nd + 1][k] = m * e + p * um[k];
        }
    }
    for (int v = 0; v < nv; ++v) {
        T s = f[v][0] * nrm[0];
        for (int k = 1; k < nd; ++k) s = s + f[v][k] * nrm[k];
        F[v] = s;
    }
}

// Rotations to/from the face frame (Equations/Euler.jl:16-64)
template <class T>
inline void rotate2face(int nd, const T* Q, const double* fr, T* R) {
    const double* n = fr;
    const double* t = fr + nd;
    const double* b = fr + 2 * nd;
    R[0] = Q[0];
    if (nd == 1) {
        R[1] = Q[1] * n[0];
    } else if (nd == 2) {
        R[1] = Q[1] * n[0] + Q[2] * n[1];
        R[2] = Q[1] * t[0] + Q[2] * t[1];
    } else {
        R[1] = (Q[1] * n[0] + Q[2] * n[1]) + Q[3] * n[2];
        R[2] = (Q[1] * t[0] + Q[2] * t[1]) + Q[3] * t[2];
        R[3] = (Q[1] * b[0] + Q[2] * b[1]) + Q[3] * b[2];
    }
    R[nd + 1] = Q[nd + 1];
}

template <class T>
inline void rotate2phys(int nd, const T* R, const double* fr, T* Q) {
    const double* n = fr;
    const double* t = fr + nd;
    const double* b = fr + 2 * nd;
    Q[0] = R[0];
    if (nd == 1) {
        Q[1] = R[1] * n[0];
    } else if (nd == 2) {
        Q[1] = R[1] * n[0] + R[2] * t[0];
        Q[2] = R[1] * n[1] + R[2] * t[1];
    } else {
        Q[1] = (R[1] * n[0] + R[2] * t[0]) + R[3] * b[0];
        Q[2] = (R[1] * n[1] + R[2] * t[1]) + R[3] * b[1];
        Q[3] = (R[1] * n[2] + R[2] * t[2]) + R[3] * b[2];
    }
    Q[nd + 1] = R[nd + 1];
}

// Entropy variables (FlouCommon/Euler.jl:202-217, 273-307)
template <class T>
inline void cons2entropy(const T* Q, int nd, double gamma, T* W) {
    T p = pressure(Q, nd, gamma);
    T s = olog(p) - gamma * olog(Q[0]);
    T m2 = Q[1] * Q[1];
    for (int k = 1; k < nd; ++k) m2 = m2 + Q[1 + k] * Q[1 + k];
    W[0] = (gamma - s) / (gamma - 1.0) - (m2 / Q[0]) / (2.0 * p);
    for (int k = 0; k < nd; ++k) W[1 + k] = Q[1 + k] / p;
    W[nd + 1] = -Q[0] / p;
}

// Entropy variables back to conservative ones: vars_entropy2prim then vars_prim2cons
// (FlouCommon/Euler.jl:309-339, 255-271).
template <class T>
inline void entropy2cons(const T* W, int nd, double gamma, T* Q) {
    const T wl = W[nd + 1];
    T u[3];
    for (int k = 0; k < nd; ++k) u[k] = -W[1 + k] / wl;
    T q2 = u[0] * u[0];
    for (int k = 1; k < nd; ++k) q2 = q2 + u[k] * u[k];
    const T s = gamma - (gamma - 1.0) * (W[0] - (wl * q2) / 2.0);
    const T p = opow(opow(-wl, gamma) * oexp(s), 1.0 / (1.0 - gamma));
    const T rho = -p * wl;
    Q[0] = rho;
    for (int k = 0; k < nd; ++k) Q[1 + k] = rho * u[k];
    Q[nd + 1] = p / (gamma - 1.0) + (rho * q2) / 2.0;
}

// Euler surface flux in the rotated frame (Equations/Euler.jl:99-380)
template <class T>
inline void euler_numericalflux(const Disc& d, const T* Ql, const T* Qr, T* Fn) {
    const int nd = d.nd, nv = d.nv;
    const double g = d.gamma;
    T Pl[5], Pr[5];
    cons2prim(Ql, nd, g, Pl);
    cons2prim(Qr, nd, g, Pr);
    const T pl = Pl[nd + 1], pr = Pr[nd + 1];
    const T ul = Pl[1], ur = Pr[1];
    // averaging part
    if (d.savg == AVG_STD) {
        Fn[0] = (Ql[1] + Qr[1]) / 2.0;
        Fn[1] = (((Ql[1] * ul + pl) + Qr[1] * ur) + pr) / 2.0;
        for (int c = 1; c < nd; ++c) Fn[1 + c] = (Ql[1 + c] * ul + Qr[1 + c] * ur) / 2.0;
        Fn[nd + 1] = ((Ql[nd + 1] + pl) * ul + (Qr[nd + 1] + pr) * ur) / 2.0;
    } else {  // ChandrasekharAverage (Euler.jl:169-226)
        const T rl = Pl[0], rr = Pr[0];
        T um[3];
        for (int k = 0; k < nd; ++k) um[k] = (Pl[1 + k] + Pr[1 + k]) / 2.0;
        T bl = rl / (2.0 * pl), br = rr / (2.0 * pr);
        T rho = logarithmic_mean(rl, rr);
        T p = (rl + rr) / (2.0 * (bl + br));
        T beta = logarithmic_mean(bl, br);
        T sq = Pl[1] * Pl[1];
        for (int k = 1; k < nd; ++k) sq = sq + Pl[1 + k] * Pl[1 + k];
        for (int k = 0; k < nd; ++k) {
            // Euler.jl:216 squares wl instead of wr in 3-D
            const T& c = (nd == 3 && k == 2 && d.chandrasekhar_surface_typo) ? Pl[3] : Pr[1 + k];
            sq = sq + c * c;
        }
        T h = 1.0 / ((2.0 * beta) * (g - 1.0)) - sq / 4.0;
        h = h + p / rho;
        for (int k = 0; k < nd; ++k) h = h + um[k] * um[k];
        Fn[0] = rho * um[0];
        Fn[1] = rho * (um[0] * um[0]) + p;
        for (int c = 1; c < nd; ++c) Fn[1 + c] = (rho * um[0]) * um[c];
        Fn[nd + 1] = (rho * um[0]) * h;
    }
    if (d.sdiss == DISS_LXF) {
        // Euler.jl:138-165
        T al = soundvelocity(Pl[0], pl, g);
        T ar = soundvelocity(Pr[0], pr, g);
        T lam = omax(oabs(ul) + al, oabs(ur) + ar);
        for (int v = 0; v < nv; ++v)
            Fn[v] = Fn[v] + ((lam * (Ql[v] - Qr[v])) / 2.0) * d.intensity;
    } else if (d.sdiss == DISS_SCALAR) {
        // ScalarDissipation (Euler.jl:228-301).  In 3-D the reference takes the left
        // momenta from Qr (Euler.jl:257), which zeroes the momentum dissipation; reproduced
        // when chandrasekhar_surface_typo (the "reference typos" switch) is set.
        const T rl = Pl[0], rr = Pr[0];
        const bool typo = nd == 3 && d.chandrasekhar_surface_typo;
        T um[3];
        for (int k = 0; k < nd; ++k) um[k] = (Pl[1 + k] + Pr[1 + k]) / 2.0;
        const T rho = (rl + rr) / 2.0;
        const T bl = rl / (2.0 * pl), br = rr / (2.0 * pr);
        const T beta = logarithmic_mean(bl, br);
        const T al = soundvelocity(rl, pl, g);
        const T ar = soundvelocity(rr, pr, g);
        const T lam = omax(oabs(ul) + al, oabs(ur) + ar);
        T D[5];
        D[0] = rr - rl;
        for (int k = 0; k < nd; ++k) D[1 + k] = Qr[1 + k] - (typo ? Qr[1 + k] : Ql[1 + k]);
        T s = (1.0 / beta) / (g - 1.0) + Pl[1] * Pr[1];
        for (int k = 1; k < nd; ++k) s = s + Pl[1 + k] * Pr[1 + k];
        T t = um[0] * (Pr[1] - Pl[1]);
        for (int k = 1; k < nd; ++k) t = t + um[k] * (Pr[1 + k] - Pl[1 + k]);
        t = t + (1.0 / br - 1.0 / bl) / (2.0 * (g - 1.0));
        D[nd + 1] = (s * (rr - rl)) / 2.0 + rho * t;
        for (int v = 0; v < nv; ++v) Fn[v] = Fn[v] - ((lam / 2.0) * D[v]) * d.intensity;
    } else if (d.sdiss == DISS_MATRIX) {
        // Euler.jl:308-380
        const T rl = Pl[0], rr = Pr[0];
        T um[3];
        for (int k = 0; k < nd; ++k) um[k] = (Pl[1 + k] + Pr[1 + k]) / 2.0;
        T m2 = um[0] * um[0];
        for (int k = 1; k < nd; ++k) m2 = m2 + um[k] * um[k];
        T q2 = Pl[1] * Pl[1];
        for (int k = 1; k < nd; ++k) q2 = q2 + Pl[1 + k] * Pl[1 + k];
        for (int k = 0; k < nd; ++k) q2 = q2 + Pr[1 + k] * Pr[1 + k];
        T v2 = 2.0 * m2 - q2 / 2.0;
        T bl = rl / (2.0 * pl), br = rr / (2.0 * pr);
        T rho = logarithmic_mean(rl, rr);
        T p = (rl + rr) / (2.0 * (bl + br));
        T beta = logarithmic_mean(bl, br);
        T a = soundvelocity(rho, p, g);
        T h = ((g / (2.0 * beta)) / (g - 1.0)) + v2 / 2.0;
        T Wl[5], Wr[5], dW[5];
        cons2entropy(Ql, nd, g, Wl);
        cons2entropy(Qr, nd, g, Wr);
        for (int v = 0; v < nv; ++v) dW[v] = Wl[v] - Wr[v];
        T u = um[0];
        T Lam[5], Tm[5], R[5][5];  // R[row][col]
        for (int i = 0; i < nv; ++i) for (int j = 0; j < nv; ++j) R[i][j] = T(0.0);
        // columns: (1, u-a, v, w, h-ua), (1, u, v, w, v2/2), e_v-columns, (1, u+a, v, w, h+ua)
        Lam[0] = oabs(u - a);
        Lam[nv - 1] = oabs(u + a);
        for (int c = 1; c < nv - 1; ++c) Lam[c] = oabs(u);
        Tm[0] = rho / (2.0 * g);
        Tm[1] = ((g - 1.0) * rho) / g;
        for (int c = 2; c < nv - 1; ++c) Tm[c] = p;
        Tm[nv - 1] = rho / (2.0 * g);
        // column 0
        R[0][0] = T(1.0); R[1][0] = u - a;
        for (int k = 1; k < nd; ++k) R[1 + k][0] = um[k];
        R[nv - 1][0] = h - u * a;
        // column 1
        R[0][1] = T(1.0); R[1][1] = u;
        for (int k = 1; k < nd; ++k) R[1 + k][1] = um[k];
        R[nv - 1][1] = v2 / 2.0;
        // tangential columns
        for (int k = 1; k < nd; ++k) {
            R[1 + k][1 + k] = T(1.0);
            R[nv - 1][1 + k] = um[k];
        }
        // last column
        R[0][nv - 1] = T(1.0); R[1][nv - 1] = u + a;
        for (int k = 1; k < nd; ++k) R[1 + k][nv - 1] = um[k];
        R[nv - 1][nv - 1] = h + u * a;
        // M = ((R*Λ)*T)*R'  then M*(Wl-Wr)
        T RL[5][5], M[5][5];
        for (int i = 0; i < nv; ++i)
            for (int j = 0; j < nv; ++j) RL[i][j] = (R[i][j] * Lam[j]) * Tm[j];
        for (int i = 0; i < nv; ++i)
            for (int j = 0; j < nv; ++j) {
                T s = RL[i][0] * R[j][0];
                for (int k = 1; k < nv; ++k) s = s + RL[i][k] * R[j][k];
                M[i][j] = s;
            }
        for (int i = 0; i < nv; ++i) {
            T s = M[i][0] * dW[0];
            for (int k = 1; k < nv; ++k) s = s + M[i][k] * dW[k];
            Fn[i] = Fn[i] + (s / 2.0) * d.intensity;
        }
    }
}

// Subcell-interface frame and surface Jacobian of a Cartesian element, direction dir
// (PhysicalSubgrid for CartesianMesh, PhysicalRegions.jl:75-145): n along the axis,
// (t, b) as the reference orders them; jac = product of the other half-widths.
inline void subgrid_frame(const Disc& d, int dir, double* fr, double& jac) {
    const int nd = d.nd;
    for (int i = 0; i < 3 * nd; ++i) fr[i] = 0.0;
    if (nd == 1) {
        fr[0] = 1.0;
        jac = 1.0;
    } else if (nd == 2) {
        if (dir == 0) { fr[0] = 1.0; fr[2 + 1] = 1.0; jac = d.dx[1] / 2.0; }
        else { fr[1] = 1.0; fr[2 + 0] = -1.0; jac = d.dx[0] / 2.0; }
    } else {
        const int t = (dir + 1) % 3, b = (dir + 2) % 3;
        fr[dir] = 1.0; fr[3 + t] = 1.0; fr[6 + b] = 1.0;
        jac = dir == 0 ? d.dx[1] * d.dx[2] / 4.0 : (dir == 1 ? d.dx[0] * d.dx[2] / 4.0 : d.dx[0] * d.dx[1] / 4.0);
    }
}

// Blending factor of the hybrid operator (OpDivergence.jl:781-801, the active "Fisher"
// form limited from below by 1/2): δ = max((√(b²+c) - b)/√(b²+c), 1/2).
template <class T>
inline T hybrid_delta(const T& b, double c) {
    T s = osqrt(b * b + c);
    s = (s - b) / s;
    return omax(s, T(0.5));
}

// Subcell interface fluxes F̄[1..n-1] of one line (OpDivergence.jl:576-618; index s is the
// interface between line positions s-1 and s; F̄[0] = F̄[n] = 0 stay outside).
template <class T>
void hybrid_line_fluxes(const Disc& d, const T* Qe, const T* We, long e, int dir, int f, T* Fb) {
    const int nd = d.nd, n = d.n, nv = d.nv;
    const Ops& o = d.ops;
    double fr[9], sjac;
    subgrid_frame(d, dir, fr, sjac);
    for (int s = 1; s < n; ++s) {
        T* F = Fb + (size_t)s * nv;
        for (int v = 0; v < nv; ++v) F[v] = T(0.0);
        for (int ik = s; ik < n; ++ik) {
            const int nk = line_node(nd, n, dir, f, ik);
            for (int il = 0; il < s; ++il) {
                const int nl = line_node(nd, n, dir, f, il);
                T Ft[5];
                twopointflux(d, &Qe[nl * nv], &Qe[nk * nv], d.metric(e, nl) + nd * dir,
                             d.metric(e, nk) + nd * dir, Ft);
                const double c = (2.0 * o.w[il]) * o.D[il * n + ik];
                for (int v = 0; v < nv; ++v) F[v] = F[v] + c * Ft[v];
            }
        }
        // FV flux between positions s-1 and s on the subcell interface frame
        const int nl = line_node(nd, n, dir, f, s - 1), nr = line_node(nd, n, dir, f, s);
        T Rl[5], Rr[5], Fr[5], Fv[5];
        rotate2face(nd, &Qe[nl * nv], fr, Rl);
        rotate2face(nd, &Qe[nr * nv], fr, Rr);
        euler_numericalflux(d, Rl, Rr, Fr);
        rotate2phys(nd, Fr, fr, Fv);
        for (int v = 0; v < nv; ++v) Fv[v] = Fv[v] * sjac;
        // blending on the entropy production of the high-order flux
        const T* Wl = &We[nl * nv];
        const T* Wr = &We[nr * nv];
        T b = (Wr[0] - Wl[0]) * (F[0] - Fv[0]);
        for (int v = 1; v < nv; ++v) b = b + (Wr[v] - Wl[v]) * (F[v] - Fv[v]);
        const T dl = hybrid_delta(b, d.blend);
        for (int v = 0; v < nv; ++v) F[v] = (1.0 - dl) * Fv[v] + dl * F[v];
    }
}

// HybridDivOperator on Gauss nodes (_hybrid_nb_surface_contribution!,
// OpDivergence.jl:633-779), added to dQ.  Per element, direction and line:
//   F♯ table of the line (diagonal = contravariant flux, pairs with the lower position
//   first); entropy-projected end states and Fl_i, Fr_i as in the split form, less
//   (l·Fl + Fn_l) and (r·Fr - Fn_r);
//   F̄[0] = -Fn_l, F̄[n] = Fn_r and, along the line,
//     F̄[ii+1] = ((F̄[ii] + (D♯[ii,:]·F♯[:,i]) ω[ii]) - l[ii] Fl_i) + r[ii] Fr_i,
//     blended with the FV flux of nodes ii, ii+1: F̄[ii+1] = (1-δ) F̄v + δ F̄[ii+1];
//   dQ[i] += (F̄[ii] - F̄[ii+1]) / ω[ii].
// `FN(face, side, f)` gives the stored numerical flux of a face side.
template <class T, class FNF>
void hybrid_nb_surface(const Disc& d, const T* Q, T* dQ, FNF FN) {
    const int nd = d.nd, n = d.n, nv = d.nv, npe = d.npe, npf = d.npf;
    const long ne = d.ne, N = d.ndofs();
    const Ops& o = d.ops;
    std::atomic<int> domain_failed(0);
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; ++e) try {
        const long base = e * npe;
        std::vector<T> Qe((size_t)npe * nv), We((size_t)npe * nv), Ft((size_t)npe * nd * nv);
        std::vector<T> Fs((size_t)n * n * nv), Fl((size_t)n * nv), Fr((size_t)n * nv), Fb((size_t)(n + 1) * nv);
        for (int i = 0; i < npe; ++i)
            for (int v = 0; v < nv; ++v) Qe[i * nv + v] = Q[(size_t)v * N + base + i];
        for (int i = 0; i < npe; ++i)
            for (int dir = 0; dir < nd; ++dir)
                contravariant_flux(d, &Qe[i * nv], d.metric(e, i), dir, &Ft[((size_t)i * nd + dir) * nv]);
        for (int i = 0; i < npe; ++i) cons2entropy(&Qe[i * nv], nd, d.gamma, &We[i * nv]);
        for (int dir = 0; dir < nd; ++dir) {
            double fr[9], sjac, nvec[3];
            subgrid_frame(d, dir, fr, sjac);
            for (int k = 0; k < nd; ++k) nvec[k] = fr[k] * sjac;
            const long facel = d.e_face[e * 2 * nd + 2 * dir];
            const int sidel = d.e_side[e * 2 * nd + 2 * dir];
            const long facer = d.e_face[e * 2 * nd + 2 * dir + 1];
            const int sider = d.e_side[e * 2 * nd + 2 * dir + 1];
            for (int f = 0; f < npf; ++f) {
                // F♯ of the line (_flux_splitdiv_tensorproduct!, OpDivergence.jl:248-271)
                for (int a = 0; a < n; ++a) {
                    const int na = line_node(nd, n, dir, f, a);
                    for (int v = 0; v < nv; ++v) Fs[((size_t)a * n + a) * nv + v] = Ft[((size_t)na * nd + dir) * nv + v];
                    for (int b = a + 1; b < n; ++b) {
                        const int nb = line_node(nd, n, dir, f, b);
                        T F[5];
                        twopointflux(d, &Qe[na * nv], &Qe[nb * nv], d.metric(e, na) + nd * dir,
                                     d.metric(e, nb) + nd * dir, F);
                        for (int v = 0; v < nv; ++v) {
                            Fs[((size_t)a * n + b) * nv + v] = F[v];
                            Fs[((size_t)b * n + a) * nv + v] = F[v];
                        }
                    }
                }
                // entropy-projected fluxes (_flux_splitdiv_nb_tensorproduct!, 382-427)
                T Wl[5], Wr[5], Ql[5], Qr[5];
                for (int v = 0; v < nv; ++v) {
                    const int n0 = line_node(nd, n, dir, f, 0);
                    T sl = o.lm[0] * We[n0 * nv + v], sr = o.lp[0] * We[n0 * nv + v];
                    for (int a = 1; a < n; ++a) {
                        const int na = line_node(nd, n, dir, f, a);
                        sl = sl + o.lm[a] * We[na * nv + v];
                        sr = sr + o.lp[a] * We[na * nv + v];
                    }
                    Wl[v] = sl;
                    Wr[v] = sr;
                }
                entropy2cons(Wl, nd, d.gamma, Ql);
                entropy2cons(Wr, nd, d.gamma, Qr);
                for (int a = 0; a < n; ++a) {
                    const int na = line_node(nd, n, dir, f, a);
                    twopointflux(d, &Qe[na * nv], Ql, d.metric(e, na) + nd * dir, nvec, &Fl[a * nv]);
                    twopointflux(d, &Qe[na * nv], Qr, d.metric(e, na) + nd * dir, nvec, &Fr[a * nv]);
                }
                const T* fnl = FN(facel, sidel, f);
                const T* fnr = FN(facer, sider, f);
                for (int v = 0; v < nv; ++v) {
                    T lF = o.lm[0] * Fl[v], rF = o.lp[0] * Fr[v];
                    for (int a = 1; a < n; ++a) {
                        lF = lF + o.lm[a] * Fl[a * nv + v];
                        rF = rF + o.lp[a] * Fr[a * nv + v];
                    }
                    for (int a = 0; a < n; ++a) {
                        Fl[a * nv + v] = Fl[a * nv + v] - (lF + fnl[v]);
                        Fr[a * nv + v] = Fr[a * nv + v] - (rF - fnr[v]);
                    }
                }
                // subcell fluxes along the line (_surf_hybrid_nb_tensorproduct!, 732-779)
                for (int v = 0; v < nv; ++v) {
                    Fb[v] = -fnl[v];
                    Fb[(size_t)n * nv + v] = fnr[v];
                }
                for (int ii = 0; ii + 1 < n; ++ii) {
                    const int ni = line_node(nd, n, dir, f, ii), nr = line_node(nd, n, dir, f, ii + 1);
                    T* F = &Fb[(size_t)(ii + 1) * nv];
                    for (int v = 0; v < nv; ++v) {
                        T sum = o.Dsh[ii * n + 0] * Fs[((size_t)0 * n + ii) * nv + v];
                        for (int a = 1; a < n; ++a) sum = sum + o.Dsh[ii * n + a] * Fs[((size_t)a * n + ii) * nv + v];
                        F[v] = ((Fb[(size_t)ii * nv + v] + sum * o.w[ii]) - o.lm[ii] * Fl[ii * nv + v]) +
                               o.lp[ii] * Fr[ii * nv + v];
                    }
                    T Rl[5], Rr[5], Fn[5], Fv[5];
                    rotate2face(nd, &Qe[ni * nv], fr, Rl);
                    rotate2face(nd, &Qe[nr * nv], fr, Rr);
                    euler_numericalflux(d, Rl, Rr, Fn);
                    rotate2phys(nd, Fn, fr, Fv);
                    for (int v = 0; v < nv; ++v) Fv[v] = Fv[v] * sjac;
                    const T* Wa = &We[ni * nv];
                    const T* Wb = &We[nr * nv];
                    T b = (Wb[0] - Wa[0]) * (F[0] - Fv[0]);
                    for (int v = 1; v < nv; ++v) b = b + (Wb[v] - Wa[v]) * (F[v] - Fv[v]);
                    const T dl = hybrid_delta(b, d.blend);
                    for (int v = 0; v < nv; ++v) F[v] = (1.0 - dl) * Fv[v] + dl * F[v];
                }
                for (int ii = 0; ii < n; ++ii) {
                    const int ni = line_node(nd, n, dir, f, ii);
                    for (int v = 0; v < nv; ++v) {
                        T& x = dQ[(size_t)v * N + base + ni];
                        x = x + (Fb[(size_t)ii * nv + v] - Fb[(size_t)(ii + 1) * nv + v]) / o.w[ii];
                    }
                }
            }
        }
    } catch (std::domain_error&) {
        domain_failed = 1;
    }
    if (domain_failed) throw std::domain_error("DomainError in the hybrid surface terms");
}

// ------------------------------------------------------------------------------------
// rhs! (Equations/Hyperbolic.jl:31-69).  Q, dQ: (ndofs, nv) column-major.
template <class T>
struct Work {
    std::vector<T> Qf, Fn;  // [face][side][f][v]
};

// Exterior state of one boundary face node (applyBCs!, Interfaces.jl:25-49, calling
// bc(Qin, x, frame, t, eq)): `fixed` is the per-node state of a GenericBC (DIRICHLET) or
// the constant Qext of an EulerInflowBC; fr is the face frame (n, t, b) at the node.
template <class T>
void bc_exterior(int nd, int nv, int type, const double* fixed, const T* Qin, const double* fr, T* Qext) {
    if (type == BC_DIRICHLET || type == BC_INFLOW) {
        // GenericBC with a time-independent Qext(x); EulerInflowBC (Equations/Euler.jl:69-80)
        for (int v = 0; v < nv; ++v) Qext[v] = T(fixed[v]);
    } else if (type == BC_OUTFLOW) {  // EulerOutflowBC (Equations/Euler.jl:82-86)
        for (int v = 0; v < nv; ++v) Qext[v] = Qin[v];
    } else {  // EulerSlipBC (Equations/Euler.jl:88-94)
        T R[5];
        rotate2face(nd, Qin, fr, R);
        R[1] = -R[1];
        rotate2phys(nd, R, fr, Qext);
    }
}

template <class T>
void apply_bc(const Disc& d, const T* Qin, long face, int fnode, T* Qext) {
    const int nv = d.nv;
    const int ib = d.f_bd[face];
    const int type = d.bctype.at(ib);
    const double* fixed = type == BC_DIRICHLET ? &d.f_bcstate[((size_t)face * d.npf + fnode) * nv]
                          : type == BC_INFLOW  ? d.bcdata[ib].data()
                                               : nullptr;
    bc_exterior(d.nd, nv, type, fixed, Qin, d.frame(face, fnode), Qext);
}

template <class T>
void rhs(const Disc& d, const T* Q, T* dQ, Work<T>& wk) {
    const int nd = d.nd, n = d.n, nv = d.nv, npe = d.npe, npf = d.npf;
    const long ne = d.ne, nf = d.nf, N = d.ndofs();
    const Ops& o = d.ops;
    wk.Qf.assign((size_t)nf * 2 * npf * nv, T(0.0));
    wk.Fn.assign((size_t)nf * 2 * npf * nv, T(0.0));
    T* Qf = wk.Qf.data();
    T* Fn = wk.Fn.data();
    auto QF = [&](long face, int side, int f) { return Qf + (((size_t)face * 2 + side) * npf + f) * nv; };
    auto FN = [&](long face, int side, int f) { return Fn + (((size_t)face * 2 + side) * npf + f) * nv; };

    // fill!(dQ, 0)
    for (long i = 0; i < N * nv; ++i) dQ[i] = T(0.0);

    // project2faces! (Interfaces.jl:51-109)
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; ++e) {
        const long base = e * npe;
        for (int dir = 0; dir < nd; ++dir) {
            for (int lr = 0; lr < 2; ++lr) {
                const int pos = 2 * dir + lr;
                const long face = d.e_face[e * 2 * nd + pos];
                const int side = d.e_side[e * 2 * nd + pos];
                const double* l = lr == 0 ? o.lm.data() : o.lp.data();
                for (int f = 0; f < npf; ++f) {
                    T* q = QF(face, side, f);
                    for (int v = 0; v < nv; ++v) {
                        T s = T(0.0);
                        for (int a = 0; a < n; ++a)
                            s = s + l[a] * Q[(size_t)v * N + base + line_node(nd, n, dir, f, a)];
                        q[v] = s;
                    }
                }
            }
        }
    }

    // Domain errors (sqrt/log of a negative) cannot leave an OpenMP region: record and
    // rethrow after the loop.
    std::atomic<int> domain_failed(0);

    // Nodes without boundaries (Gauss): l and r are not unit vectors
    bool gauss = false;
    for (int a = 0; a < n; ++a)
        if (o.lm[a] != (a == 0 ? 1.0 : 0.0) || o.lp[a] != (a == n - 1 ? 1.0 : 0.0)) gauss = true;
    // HybridDivOperator on Gauss nodes has no volume term: all of it is in the surface
    // contribution (OpDivergence.jl:474-491, 614-730)
    const bool hybrid_nb = d.volop == VOL_HYBRID && gauss;

    // volume_contribution! (OpDivergence.jl:109-160, 200-282)
#pragma omp parallel for schedule(static)
    for (long e = 0; e < (hybrid_nb ? 0 : ne); ++e) try {
        const long base = e * npe;
        std::vector<T> Qe((size_t)npe * nv), Ft((size_t)npe * nd * nv), acc((size_t)npe * nv);
        for (int i = 0; i < npe; ++i)
            for (int v = 0; v < nv; ++v) Qe[i * nv + v] = Q[(size_t)v * N + base + i];
        for (int i = 0; i < npe; ++i)
            for (int v = 0; v < nv; ++v) acc[i * nv + v] = dQ[(size_t)v * N + base + i];
        for (int i = 0; i < npe; ++i)
            for (int dir = 0; dir < nd; ++dir)
                contravariant_flux(d, &Qe[i * nv], d.metric(e, i), dir, &Ft[((size_t)i * nd + dir) * nv]);
        std::vector<T> We, Fb;
        if (d.volop == VOL_HYBRID) {
            // entropy variables of the element, then the telescopic subcell form
            // (OpDivergence.jl:500-620): dQ[i] += (F̄[i] - F̄[i+1]) / ω[i]
            We.resize((size_t)npe * nv);
            Fb.assign((size_t)(n + 1) * nv, T(0.0));
            for (int i = 0; i < npe; ++i) cons2entropy(&Qe[i * nv], nd, d.gamma, &We[i * nv]);
        }
        const int nlines = npf;
        for (int dir = 0; dir < nd; ++dir) {
            for (int f = 0; f < nlines; ++f) {
                if (d.volop == VOL_HYBRID) {
                    hybrid_line_fluxes(d, Qe.data(), We.data(), e, dir, f, Fb.data());
                    for (int i = 0; i < n; ++i) {
                        const int ni = line_node(nd, n, dir, f, i);
                        for (int v = 0; v < nv; ++v) {
                            const T l = i == 0 ? T(0.0) : Fb[(size_t)i * nv + v];
                            const T r = i == n - 1 ? T(0.0) : Fb[(size_t)(i + 1) * nv + v];
                            acc[ni * nv + v] = acc[ni * nv + v] + (l - r) / o.w[i];
                        }
                    }
                } else if (d.volop == VOL_STRONG) {
                    // mul!(dQ[line], Ds, F̃[line, dir], -1, 1): k outer, i inner
                    for (int k = 0; k < n; ++k) {
                        const int nk = line_node(nd, n, dir, f, k);
                        const T* Fk = &Ft[((size_t)nk * nd + dir) * nv];
                        for (int i = 0; i < n; ++i) {
                            const int ni = line_node(nd, n, dir, f, i);
                            const double Dik = o.Ds[i * n + k];
                            for (int v = 0; v < nv; ++v)
                                acc[ni * nv + v] = acc[ni * nv + v] + Dik * (-Fk[v]);
                        }
                    }
                } else {
                    // F♯ table for this line (pairs computed once, first argument = lower position)
                    std::vector<T> Fs((size_t)n * n * nv);
                    for (int a = 0; a < n; ++a) {
                        const int na = line_node(nd, n, dir, f, a);
                        for (int v = 0; v < nv; ++v)
                            Fs[((size_t)a * n + a) * nv + v] = Ft[((size_t)na * nd + dir) * nv + v];
                        for (int b = a + 1; b < n; ++b) {
                            const int nb = line_node(nd, n, dir, f, b);
                            T F[5];
                            twopointflux(d, &Qe[na * nv], &Qe[nb * nv], d.metric(e, na) + nd * dir,
                                         d.metric(e, nb) + nd * dir, F);
                            for (int v = 0; v < nv; ++v) {
                                Fs[((size_t)a * n + b) * nv + v] = F[v];
                                Fs[((size_t)b * n + a) * nv + v] = F[v];
                            }
                        }
                    }
                    // dQ[i] -= D♯[ii, ij] * F♯[ii, j]: j outer, i inner
                    for (int j = 0; j < n; ++j)
                        for (int i = 0; i < n; ++i) {
                            const int ni = line_node(nd, n, dir, f, i);
                            const double Dij = o.Dsh[i * n + j];
                            for (int v = 0; v < nv; ++v)
                                acc[ni * nv + v] = acc[ni * nv + v] - Dij * Fs[((size_t)i * n + j) * nv + v];
                        }
                }
            }
        }
        for (int i = 0; i < npe; ++i)
            for (int v = 0; v < nv; ++v) dQ[(size_t)v * N + base + i] = acc[i * nv + v];
    } catch (std::domain_error&) {
        domain_failed = 1;
    }
    if (domain_failed) throw std::domain_error("DomainError in the volume term");

    // applyBCs! (Interfaces.jl:25-49)
    for (long face = 0; face < nf; ++face) {
        if (d.f_elem[face * 2 + 1] >= 0) continue;
        for (int f = 0; f < npf; ++f) apply_bc(d, QF(face, 0, f), face, f, QF(face, 1, f));
    }

    // interface_fluxes! (Interfaces.jl:111-136): master node f meets slave node
    // j = master2slave(f, orientation); the frame and jac are the master's
#pragma omp parallel for schedule(static)
    for (long face = 0; face < nf; ++face) try {
        const int orient = d.f_orient[face];
        for (int f = 0; f < npf; ++f) {
            const int j = master2slave(nd, n, f, orient);
            const double* fr = d.frame(face, f);
            const double jac = d.facejac(face, f);
            const T* Ql = QF(face, 0, f);
            const T* Qr = QF(face, 1, j);
            T F[5];
            if (d.eq == EQ_ADVECTION) {
                // LinearAdvection.jl:27-41 (rotations are the identity)
                double an = d.a[0] * fr[0];
                for (int k = 1; k < nd; ++k) an = an + d.a[k] * fr[k];
                F[0] = (an * (Ql[0] + Qr[0])) / 2.0;
                if (d.sdiss == DISS_LXF)
                    F[0] = F[0] + ((std::fabs(an) * (Ql[0] - Qr[0])) / 2.0) * d.intensity;
            } else {
                T Rl[5], Rr[5], Fr[5];
                rotate2face(nd, Ql, fr, Rl);
                rotate2face(nd, Qr, fr, Rr);
                euler_numericalflux(d, Rl, Rr, Fr);
                rotate2phys(nd, Fr, fr, F);
            }
            T* F1 = FN(face, 0, f);
            T* F2 = FN(face, 1, j);
            for (int v = 0; v < nv; ++v) {
                F1[v] = F[v] * jac;
                F2[v] = -F1[v];
            }
        }
    } catch (std::domain_error&) {
        domain_failed = 1;
    }
    if (domain_failed) throw std::domain_error("DomainError in the interface fluxes");

    // Split form on nodes without boundaries (Gauss): the entropy-projected surface terms
    // (_splitdiv_nb_surface_contribution!, OpDivergence.jl:300-436) replace the FR lift.
    if (hybrid_nb) {
        hybrid_nb_surface(d, Q, dQ, FN);
    } else if (d.volop == VOL_SPLIT && gauss && d.eq == EQ_EULER) {
#pragma omp parallel for schedule(static)
        for (long e = 0; e < ne; ++e) try {
            const long base = e * npe;
            std::vector<T> Qe((size_t)npe * nv), We((size_t)npe * nv), Fl((size_t)n * nv), Fr((size_t)n * nv);
            for (int i = 0; i < npe; ++i)
                for (int v = 0; v < nv; ++v) Qe[i * nv + v] = Q[(size_t)v * N + base + i];
            for (int i = 0; i < npe; ++i) cons2entropy(&Qe[i * nv], nd, d.gamma, &We[i * nv]);
            for (int dir = 0; dir < nd; ++dir) {
                double fr[9], sjac, nvec[3];
                subgrid_frame(d, dir, fr, sjac);
                for (int k = 0; k < nd; ++k) nvec[k] = fr[k] * sjac;
                const long facel = d.e_face[e * 2 * nd + 2 * dir];
                const int sidel = d.e_side[e * 2 * nd + 2 * dir];
                const long facer = d.e_face[e * 2 * nd + 2 * dir + 1];
                const int sider = d.e_side[e * 2 * nd + 2 * dir + 1];
                for (int f = 0; f < npf; ++f) {
                    // entropy-projected end states
                    T Wl[5], Wr[5], Ql[5], Qr[5];
                    for (int v = 0; v < nv; ++v) {
                        const int n0 = line_node(nd, n, dir, f, 0);
                        T sl = o.lm[0] * We[n0 * nv + v], sr = o.lp[0] * We[n0 * nv + v];
                        for (int a = 1; a < n; ++a) {
                            const int na = line_node(nd, n, dir, f, a);
                            sl = sl + o.lm[a] * We[na * nv + v];
                            sr = sr + o.lp[a] * We[na * nv + v];
                        }
                        Wl[v] = sl;
                        Wr[v] = sr;
                    }
                    entropy2cons(Wl, nd, d.gamma, Ql);
                    entropy2cons(Wr, nd, d.gamma, Qr);
                    for (int a = 0; a < n; ++a) {
                        const int na = line_node(nd, n, dir, f, a);
                        twopointflux(d, &Qe[na * nv], Ql, d.metric(e, na) + nd * dir, nvec, &Fl[a * nv]);
                        twopointflux(d, &Qe[na * nv], Qr, d.metric(e, na) + nd * dir, nvec, &Fr[a * nv]);
                    }
                    const T* fnl = FN(facel, sidel, f);
                    const T* fnr = FN(facer, sider, f);
                    for (int v = 0; v < nv; ++v) {
                        T lF = o.lm[0] * Fl[v], rF = o.lp[0] * Fr[v];
                        for (int a = 1; a < n; ++a) {
                            lF = lF + o.lm[a] * Fl[a * nv + v];
                            rF = rF + o.lp[a] * Fr[a * nv + v];
                        }
                        for (int a = 0; a < n; ++a) {
                            const int na = line_node(nd, n, dir, f, a);
                            const T fl = Fl[a * nv + v] - (lF + fnl[v]);
                            const T frr = Fr[a * nv + v] - (rF - fnr[v]);
                            T& x = dQ[(size_t)v * N + base + na];
                            x = x + (o.gm[a] * fl - o.gp[a] * frr);
                        }
                    }
                }
            }
        } catch (std::domain_error&) {
            domain_failed = 1;
        }
        if (domain_failed) throw std::domain_error("DomainError in the entropy-projected surface terms");
    } else {
    // surface_contribution! (OpDivergence.jl:42-100)
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; ++e) {
        const long base = e * npe;
        for (int dir = 0; dir < nd; ++dir) {
            const long facel = d.e_face[e * 2 * nd + 2 * dir];
            const int sidel = d.e_side[e * 2 * nd + 2 * dir];
            const long facer = d.e_face[e * 2 * nd + 2 * dir + 1];
            const int sider = d.e_side[e * 2 * nd + 2 * dir + 1];
            for (int f = 0; f < npf; ++f) {
                const T* fl = FN(facel, sidel, f);
                const T* frr = FN(facer, sider, f);
                for (int a = 0; a < n; ++a) {
                    const int ni = line_node(nd, n, dir, f, a);
                    for (int v = 0; v < nv; ++v) {
                        T& x = dQ[(size_t)v * N + base + ni];
                        x = x - (o.gm[a] * fl[v] + o.gp[a] * frr[v]);
                    }
                }
            }
        }
    }
    }

    // apply_massmatrix! (MultielementDiscontinuous.jl:132-137): ldiv!(Diagonal(J), dQ)
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; ++e) {
        for (int i = 0; i < npe; ++i) {
            const double J = d.jac(e, i);
            for (int v = 0; v < nv; ++v) {
                T& x = dQ[(size_t)v * N + e * npe + i];
                x = x / J;
            }
        }
    }
}

// get_max_dt (MultielementDiscontinuous.jl:162-178; FlouCommon/Euler.jl:116-135;
// LinearAdvection.jl:46-48)
template <class T>
double max_dt(const Disc& d, const T* Q, double cfl) {
    const int nd = d.nd, nv = d.nv, npe = d.npe;
    const long N = d.ndofs();
    double dt = INFINITY;
    for (long e = 0; e < d.ne; ++e) {
        double dx = d.e_vol[e] / npe;
        dx = nd == 1 ? dx : (nd == 2 ? std::sqrt(dx) : std::cbrt(dx));
        for (int i = 0; i < npe; ++i) {
            double q[5];
            for (int v = 0; v < nv; ++v) q[v] = val(Q[(size_t)v * N + e * npe + i]);
            double loc;
            if (d.eq == EQ_ADVECTION) {
                double a2 = d.a[0] * d.a[0];
                for (int k = 1; k < nd; ++k) a2 = a2 + d.a[k] * d.a[k];
                loc = cfl * dx / std::sqrt(a2);
            } else {
                double p = pressure(q, nd, d.gamma);
                double c = osqrt(d.gamma * p / q[0]);
                double u[3];
                for (int k = 0; k < nd; ++k) u[k] = q[1 + k] / q[0];
                double um;
                if (nd == 1) um = std::fabs(u[0]);
                else {
                    double s = u[0] * u[0];
                    for (int k = 1; k < nd; ++k) s = s + u[k] * u[k];
                    um = std::sqrt(s);
                }
                loc = cfl * dx / (um + c);
            }
            dt = omin(dt, loc);
        }
    }
    return dt;
}

// tensor-product quadrature weight of node i of the standard element
inline double node_weight(const Disc& d, int i) {
    double wi[3] = {1.0, 1.0, 1.0};
    int idx = i;
    for (int k = 0; k < d.nd; ++k) { wi[k] = d.ops.w[idx % d.n]; idx /= d.n; }
    double w = wi[0];
    for (int k = 1; k < d.nd; ++k) w = w * wi[k];
    return w;
}

// integrate (MultielementDiscontinuous.jl:148-160; PhysicalRegions.jl:366-368)
template <class T>
double integrate(const Disc& d, const T* Q, int ivar) {
    const long N = d.ndofs();
    double total = 0.0;
    for (long e = 0; e < d.ne; ++e) {
        double s = 0.0;
        for (int i = 0; i < d.npe; ++i) {
            // Jω = jac .* ω with the tensor-product ω of the standard element
            double w = 1.0;
            int idx = i;
            double wi[3];
            for (int k = 0; k < d.nd; ++k) { wi[k] = d.ops.w[idx % d.n]; idx /= d.n; }
            w = wi[0];
            for (int k = 1; k < d.nd; ++k) w = w * wi[k];
            s = s + (d.jac(e, i) * w) * val(Q[(size_t)ivar * N + e * d.npe + i]);
        }
        total = total + s;
    }
    return total;
}

// ------------------------------------------------------------------------------------
// Explicit RK (OrdinaryDiffEq 6.49.1 perform_step! forms; tableaux from the published
// methods: Shu & Osher 1988; Carpenter & Kennedy, NASA TM-109112 (1994); Allampalli,
// Hixon, Nallasamy & Sawyer, JCP 228 (2009)).
struct Tableau2N {
    int s;
    double A[5], B[5], c[5];
};
Tableau2N tableau_2n(int scheme);

// Monitors (Equations/Euler.jl:541-590): Σ_e Jω_e' · s(Q_e) with s the kinetic energy
// |ρu|²/(2ρ) (FlouCommon/Euler.jl:162-175) or the mathematical entropy -ρ s/(γ-1),
// s = log p - γ log ρ (FlouCommon/Euler.jl:202-217).
enum { MON_KINETIC_ENERGY = 0, MON_ENTROPY = 1 };
inline double monitor(const Disc& d, const double* Q, int kind) {
    const int nd = d.nd, npe = d.npe;
    const long N = d.ndofs();
    double total = 0.0;
    for (long e = 0; e < d.ne; ++e) {
        double s = 0.0;
        for (int i = 0; i < npe; ++i) {
            double q[5];
            for (int v = 0; v < d.nv; ++v) q[v] = Q[(size_t)v * N + e * npe + i];
            double m2 = q[1] * q[1];
            for (int k = 1; k < nd; ++k) m2 = m2 + q[1 + k] * q[1 + k];
            double val;
            if (kind == MON_KINETIC_ENERGY) {
                val = m2 / (2.0 * q[0]);
            } else {
                const double p = pressure(q, nd, d.gamma);
                const double ent = olog(p) - d.gamma * olog(q[0]);
                val = -(q[0] * ent) / (d.gamma - 1.0);
            }
            s = s + (d.jac(e, i) * node_weight(d, i)) * val;
        }
        total = total + s;
    }
    return total;
}

// Zhang & Shu positivity limiter as Flou states it (Equations/Euler.jl:596-654), per
// element: means Q̄ = Jω'Q / vol; density pulled towards ρ̄ by θ = |(ρ̄-m)/(ρ̄-ρmin)|,
// m = min(minval, ρ̄), when θ <= 1; then the pressure of the updated nodes: p̄ = Jω'p / vol,
// θ = |(p̄-m)/(p̄-pmin)|, m = min(minval, p̄), all variables pulled towards Q̄ when θ <= 1.
inline void zhang_shu(const Disc& d, double* Q, double minval) {
    const int nd = d.nd, npe = d.npe, nv = d.nv;
    const long N = d.ndofs();
    for (long e = 0; e < d.ne; ++e) {
        double qm[5] = {0, 0, 0, 0, 0};
        for (int i = 0; i < npe; ++i) {
            const double jw = d.jac(e, i) * node_weight(d, i);
            for (int v = 0; v < nv; ++v) qm[v] = qm[v] + jw * Q[(size_t)v * N + e * npe + i];
        }
        for (int v = 0; v < nv; ++v) qm[v] = qm[v] / d.e_vol[e];
        double* rho = &Q[e * npe];
        double rmin = rho[0];
        for (int i = 1; i < npe; ++i) rmin = std::min(rmin, rho[i]);
        double m = std::min(minval, qm[0]);
        double th = std::fabs((qm[0] - m) / (qm[0] - rmin));
        if (th <= 1.0)
            for (int i = 0; i < npe; ++i) rho[i] = th * (rho[i] - qm[0]) + qm[0];
        std::vector<double> p(npe);
        double pm = 0.0;
        for (int i = 0; i < npe; ++i) {
            double q[5];
            for (int v = 0; v < nv; ++v) q[v] = Q[(size_t)v * N + e * npe + i];
            p[i] = pressure(q, nd, d.gamma);
            pm = pm + (d.jac(e, i) * node_weight(d, i)) * p[i];
        }
        pm = pm / d.e_vol[e];
        double pmin = p[0];
        for (int i = 1; i < npe; ++i) pmin = std::min(pmin, p[i]);
        m = std::min(minval, pm);
        th = std::fabs((pm - m) / (pm - pmin));
        if (th <= 1.0)
            for (int i = 0; i < npe; ++i)
                for (int v = 0; v < nv; ++v) {
                    double& x = Q[(size_t)v * N + e * npe + i];
                    x = th * (x - qm[v]) + qm[v];
                }
    }
}

// Stage limiter hook of rk_run (the OrdinaryDiffEq stage_limiter! after every stage); only
// for plain doubles.
inline void stage_limit(const Disc& d, double* u) {
    if (d.limiter_on) zhang_shu(d, u, d.limiter_minval);
}
template <int K>
inline void stage_limit(const Disc&, Dual<K>*) {}

template <class T>
void rk_run(const Disc& d, int scheme, T* u, double dt, int nsteps, Work<T>& wk) {
    const long M = d.ndofs() * d.nv;
    std::vector<T> k((size_t)M), tmp((size_t)M), u0((size_t)M), u1((size_t)M);
    for (int step = 0; step < nsteps; ++step) {
        if (scheme == RK_SSPRK33) {
            for (long i = 0; i < M; ++i) u0[i] = u[i];
            rhs(d, u0.data(), k.data(), wk);
            for (long i = 0; i < M; ++i) u1[i] = u0[i] + dt * k[i];
            stage_limit(d, u1.data());
            rhs(d, u1.data(), k.data(), wk);
            for (long i = 0; i < M; ++i) u1[i] = ((3.0 * u0[i] + u1[i]) + dt * k[i]) / 4.0;
            stage_limit(d, u1.data());
            rhs(d, u1.data(), k.data(), wk);
            for (long i = 0; i < M; ++i) u[i] = ((u0[i] + 2.0 * u1[i]) + (2.0 * dt) * k[i]) / 3.0;
            stage_limit(d, u);
        } else {
            Tableau2N tb = tableau_2n(scheme);
            for (int s = 0; s < tb.s; ++s) {
                rhs(d, u, k.data(), wk);
                if (s == 0) {
                    for (long i = 0; i < M; ++i) tmp[i] = dt * k[i];
                } else {
                    for (long i = 0; i < M; ++i) tmp[i] = tb.A[s] * tmp[i] + dt * k[i];
                }
                for (long i = 0; i < M; ++i) u[i] = u[i] + tb.B[s] * tmp[i];
                stage_limit(d, u);
            }
        }
    }
}

}  // namespace oracle
