// A C++ host of libflouhip.so that uses only include/flouhip.h: it loads the library with
// dlopen and resolves every entry point with dlsym, as Julia's ccall does, and builds its
// own 1-D periodic discretisation (operators, connectivity, geometry) without any of the
// Python host code.  Test infrastructure, run by tests/test_capi_driver.py.
//
//   capi_driver LIB            full run on the GPU (exit status 0 when every check passes)
//   capi_driver LIB --no-gpu   load, version and the argument checks of flou_hip_create only
//
// Checks of the full run:
//   * Advection1D (test/tests.jl:16-51): a = 2, Gauss nodes n = 5, 20 elements on [0, 1],
//     strong form + LxF, ORK256 with dt = 1e-3 for 500 steps returns to the initial state
//     within the reference's rtol 1e-3 (test/runtests.jl:26); the integral is conserved;
//   * 1-D Euler on GLL nodes, split Chandrasekhar + LxF: a constant state has a zero rhs, a
//     smooth periodic state keeps its three integrals over 20 SSPRK33 steps, get_max_dt is
//     positive and finite, and a state with negative pressure is reported as a domain error.
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../include/flouhip.h"

namespace {

typedef long double LD;

struct Lib {
    void* so = nullptr;
#define FN(name) decltype(&::name) name = nullptr
    FN(flou_hip_version);
    FN(flou_hip_create);
    FN(flou_hip_destroy);
    FN(flou_hip_last_error);
    FN(flou_hip_create_error);
    FN(flou_hip_ndofs);
    FN(flou_hip_state_bytes);
    FN(flou_hip_alloc);
    FN(flou_hip_free);
    FN(flou_hip_upload);
    FN(flou_hip_download);
    FN(flou_hip_synchronize);
    FN(flou_hip_rhs);
    FN(flou_hip_timeintegrate);
    FN(flou_hip_max_dt);
    FN(flou_hip_integrate);
    FN(flou_hip_check);
#undef FN
    bool load(const char* path) {
        so = dlopen(path, RTLD_NOW | RTLD_LOCAL);
        if (!so) {
            std::fprintf(stderr, "dlopen(%s): %s\n", path, dlerror());
            return false;
        }
        bool ok = true;
#define SYM(name)                                                                  \
    do {                                                                           \
        name = reinterpret_cast<decltype(name)>(dlsym(so, #name));                 \
        if (!name) { std::fprintf(stderr, "missing symbol %s\n", #name); ok = false; } \
    } while (0)
        SYM(flou_hip_version);
        SYM(flou_hip_create);
        SYM(flou_hip_destroy);
        SYM(flou_hip_last_error);
        SYM(flou_hip_create_error);
        SYM(flou_hip_ndofs);
        SYM(flou_hip_state_bytes);
        SYM(flou_hip_alloc);
        SYM(flou_hip_free);
        SYM(flou_hip_upload);
        SYM(flou_hip_download);
        SYM(flou_hip_synchronize);
        SYM(flou_hip_rhs);
        SYM(flou_hip_timeintegrate);
        SYM(flou_hip_max_dt);
        SYM(flou_hip_integrate);
        SYM(flou_hip_check);
#undef SYM
        return ok;
    }
};

int failures = 0;

void check(bool ok, const char* what, double value) {
    std::printf("%-64s %-4s %.3e\n", what, ok ? "ok" : "FAIL", value);
    if (!ok) ++failures;
}

// Legendre P_N and P_N' by the three-term recurrence
void legendre(int N, LD x, LD& P, LD& dP) {
    LD p0 = 1, p1 = x;
    if (N == 0) { P = 1; dP = 0; return; }
    for (int k = 2; k <= N; ++k) {
        LD p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    P = p1;
    dP = N * (x * p1 - p0) / (x * x - 1);
}

// 1-D nodal operators of flou_hip_desc from Gauss (gll = false) or Gauss-Lobatto nodes
struct Ops {
    int n;
    std::vector<double> xi, w, D, Ds, Dsh, lm, lp, gm, gp;
};

Ops make_ops(int n, bool gll) {
    const LD pi = 3.141592653589793238462643383279502884L;
    std::vector<LD> x(n), w(n);
    if (gll) {
        const int N = n - 1;
        x[0] = -1;
        x[N] = 1;
        for (int i = 1; i < N; ++i) {
            LD t = -std::cos(pi * i / N);
            for (int it = 0; it < 100; ++it) {  // Newton on P'_N
                LD P, dP;
                legendre(N, t, P, dP);
                LD d2P = (2 * t * dP - N * (N + 1) * P) / (1 - t * t);
                LD dx = dP / d2P;
                t -= dx;
                if (std::fabs((double)dx) < 1e-22) break;
            }
            x[i] = t;
        }
        for (int i = 0; i < n; ++i) {
            LD P, dP;
            legendre(N, x[i], P, dP);
            w[i] = 2.0L / (N * (N + 1) * P * P);
        }
    } else {
        for (int i = 0; i < n; ++i) {
            LD t = -std::cos(pi * (i + 0.75L) / (n + 0.5L));
            for (int it = 0; it < 100; ++it) {  // Newton on P_n
                LD P, dP;
                legendre(n, t, P, dP);
                LD dx = P / dP;
                t -= dx;
                if (std::fabs((double)dx) < 1e-22) break;
            }
            LD P, dP;
            legendre(n, t, P, dP);
            x[i] = t;
            w[i] = 2.0L / ((1 - t * t) * dP * dP);
        }
    }
    std::vector<LD> bw(n, 1);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k)
            if (k != j) bw[j] /= (x[j] - x[k]);
    Ops o;
    o.n = n;
    o.xi.resize(n);
    o.w.resize(n);
    o.D.assign(n * n, 0);
    o.lm.resize(n);
    o.lp.resize(n);
    for (int i = 0; i < n; ++i) { o.xi[i] = (double)x[i]; o.w[i] = (double)w[i]; }
    for (int i = 0; i < n; ++i) {
        LD diag = 0;
        for (int j = 0; j < n; ++j)
            if (j != i) {
                LD d = (bw[j] / bw[i]) / (x[i] - x[j]);
                o.D[i * n + j] = (double)d;
                diag -= d;
            }
        o.D[i * n + i] = (double)diag;
    }
    for (int side = 0; side < 2; ++side) {
        const LD t = side ? 1 : -1;
        std::vector<double>& l = side ? o.lp : o.lm;
        int hit = -1;
        for (int j = 0; j < n; ++j)
            if (x[j] == t) hit = j;
        LD den = 0;
        if (hit < 0)
            for (int j = 0; j < n; ++j) den += bw[j] / (t - x[j]);
        for (int j = 0; j < n; ++j) l[j] = hit >= 0 ? (j == hit ? 1.0 : 0.0) : (double)((bw[j] / (t - x[j])) / den);
    }
    // DGSEM reconstruction ∂g = (l(-1)/ω, l(+1)/ω); B = ∂g_r l_r' - ∂g_l l_l'
    o.gm.resize(n);
    o.gp.resize(n);
    for (int j = 0; j < n; ++j) {
        o.gm[j] = o.lm[j] / o.w[j];
        o.gp[j] = o.lp[j] / o.w[j];
    }
    o.Ds.resize(n * n);
    o.Dsh.resize(n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double B = o.gp[i] * o.lp[j] - o.gm[i] * o.lm[j];
            o.Ds[i * n + j] = o.D[i * n + j] - B;
            o.Dsh[i * n + j] = 2.0 * o.D[i * n + j] - B;
        }
    return o;
}

// A periodic 1-D mesh of ne elements on [0, L] and the descriptor pointing at it
struct Mesh1D {
    int ne;
    double L;
    std::vector<int64_t> ef, fe;
    std::vector<int8_t> es, fp;
    std::vector<int32_t> fb;
    std::vector<double> ejac, emet, evol, ffr, fjac;
};

Mesh1D periodic_mesh(int ne, double L, const Ops& o) {
    Mesh1D m;
    m.ne = ne;
    m.L = L;
    const double dx = L / ne;
    // face f joins element f-1 (its right side, the master, normal +1) and element f (its
    // left side); face 0 is the left side of element 0 (master, normal -1), paired with the
    // right side of element ne-1
    for (int e = 0; e < ne; ++e) {
        m.ef.push_back(e);
        m.ef.push_back(e + 1 == ne ? 0 : e + 1);
        m.es.push_back(e == 0 ? 0 : 1);
        m.es.push_back(e + 1 == ne ? 1 : 0);
    }
    for (int f = 0; f < ne; ++f) {
        if (f == 0) {
            m.fe.push_back(0); m.fe.push_back(ne - 1);
            m.fp.push_back(0); m.fp.push_back(1);
            m.ffr.insert(m.ffr.end(), {-1.0, 0.0, 0.0});
        } else {
            m.fe.push_back(f - 1); m.fe.push_back(f);
            m.fp.push_back(1); m.fp.push_back(0);
            m.ffr.insert(m.ffr.end(), {1.0, 0.0, 0.0});
        }
        m.fb.push_back(-1);
        m.fjac.push_back(1.0);
    }
    const double jac = dx / 2;
    double vol = 0;
    for (double wi : o.w) vol = vol + jac * wi;
    m.ejac.assign(ne, jac);
    m.emet.assign(ne, 1.0);
    m.evol.assign(ne, vol);
    return m;
}

flou_hip_desc describe(const Ops& o, const Mesh1D& m, int eq, int vol, int tp) {
    flou_hip_desc d;
    std::memset(&d, 0, sizeof d);
    d.nd = 1;
    d.n = o.n;
    d.nvars = eq == FLOU_HIP_EULER ? 3 : 1;
    d.precision = FLOU_HIP_FP64;
    d.nelements = m.ne;
    d.nfaces = m.ne;
    d.nboundaries = 0;
    d.equation = eq;
    d.gamma = 1.4;
    d.velocity[0] = 2.0;
    d.volume_op = vol;
    d.tp_flux = tp;
    d.surface_avg = FLOU_HIP_AVG_STD;
    d.surface_diss = FLOU_HIP_DISS_LXF;
    d.intensity = 1.0;
    d.D = o.D.data();
    d.Ds = o.Ds.data();
    d.Dsharp = o.Dsh.data();
    d.l_minus = o.lm.data();
    d.l_plus = o.lp.data();
    d.dg_minus = o.gm.data();
    d.dg_plus = o.gp.data();
    d.omega = o.w.data();
    d.elem_faceinds = m.ef.data();
    d.elem_facepos = m.es.data();
    d.face_eleminds = m.fe.data();
    d.face_elempos = m.fp.data();
    d.face_orientation = nullptr;
    d.face_boundary = m.fb.data();
    d.geometry = FLOU_HIP_GEOM_AFFINE;
    d.elem_jac = m.ejac.data();
    d.elem_metric = m.emet.data();
    d.elem_volume = m.evol.data();
    d.face_frames = m.ffr.data();
    d.face_jac = m.fjac.data();
    d.blend = 1.0;
    return d;
}

std::vector<double> node_x(const Ops& o, const Mesh1D& m) {
    std::vector<double> x;
    const double dx = m.L / m.ne;
    for (int e = 0; e < m.ne; ++e)
        for (int i = 0; i < o.n; ++i) x.push_back(dx * e + (o.xi[i] + 1.0) / 2.0 * dx);
    return x;
}

bool no_gpu_checks(Lib& L) {
    std::printf("library %s\n", L.flou_hip_version());
    Ops o = make_ops(4, true);
    Mesh1D m = periodic_mesh(8, 1.0, o);
    flou_hip_desc d = describe(o, m, FLOU_HIP_EULER, FLOU_HIP_SPLIT, FLOU_HIP_TP_CHANDRASEKHAR);
    d.gamma = 0.5;  // γ <= 1 is rejected before any device work
    flou_hip_handle* h = nullptr;
    int rc = L.flou_hip_create(&d, &h);
    check(rc == FLOU_HIP_EINVAL && !h, "create rejects gamma <= 1", rc);
    std::printf("  create error: %s\n", L.flou_hip_create_error());
    d = describe(o, m, FLOU_HIP_EULER, 7, FLOU_HIP_TP_CHANDRASEKHAR);
    rc = L.flou_hip_create(&d, &h);
    check(rc == FLOU_HIP_EINVAL && !h, "create rejects an unknown volume operator", rc);
    return failures == 0;
}

void advection(Lib& L) {
    Ops o = make_ops(5, false);
    Mesh1D m = periodic_mesh(20, 1.0, o);
    flou_hip_desc d = describe(o, m, FLOU_HIP_ADVECTION, FLOU_HIP_STRONG, 0);
    flou_hip_handle* h = nullptr;
    if (L.flou_hip_create(&d, &h) != FLOU_HIP_OK) {
        check(false, "advection: create", 0);
        std::printf("  %s\n", L.flou_hip_create_error());
        return;
    }
    std::vector<double> x = node_x(o, m), q0(x.size());
    for (size_t i = 0; i < x.size(); ++i) q0[i] = std::exp(-(x[i] - 0.5) * (x[i] - 0.5) / (2 * 0.1 * 0.1));
    const size_t bytes = L.flou_hip_state_bytes(h);
    void* Q = nullptr;
    L.flou_hip_alloc(h, &Q, bytes);
    L.flou_hip_upload(h, Q, q0.data(), bytes);
    double i0 = 0, i1 = 0, t = 0;
    L.flou_hip_integrate(h, Q, 0, &i0);
    int rc = L.flou_hip_timeintegrate(h, FLOU_HIP_ORK256, Q, 1e-3, 500, 0.0, &t);
    L.flou_hip_integrate(h, Q, 0, &i1);
    std::vector<double> q1(x.size());
    L.flou_hip_download(h, q1.data(), Q, bytes);
    L.flou_hip_synchronize(h);
    double n2 = 0, a2 = 0, b2 = 0;
    for (size_t i = 0; i < x.size(); ++i) {
        n2 += (q1[i] - q0[i]) * (q1[i] - q0[i]);
        a2 += q1[i] * q1[i];
        b2 += q0[i] * q0[i];
    }
    const double ret = std::sqrt(n2) / std::sqrt(std::fmax(a2, b2));
    check(rc == FLOU_HIP_OK && std::fabs(t - 0.5) < 1e-12, "advection: 500 ORK256 steps reach t = 0.5", t);
    check(ret <= 1e-3, "advection: full period returns within rtol 1e-3 (runtests.jl:26)", ret);
    check(std::fabs(i1 - i0) <= 1e-14 * std::fabs(i0), "advection: integral conserved", std::fabs(i1 - i0));
    L.flou_hip_free(h, Q);
    L.flou_hip_destroy(h);
}

void euler(Lib& L) {
    Ops o = make_ops(6, true);
    Mesh1D m = periodic_mesh(16, 1.0, o);
    flou_hip_desc d = describe(o, m, FLOU_HIP_EULER, FLOU_HIP_SPLIT, FLOU_HIP_TP_CHANDRASEKHAR);
    flou_hip_handle* h = nullptr;
    if (L.flou_hip_create(&d, &h) != FLOU_HIP_OK) {
        check(false, "euler: create", 0);
        std::printf("  %s\n", L.flou_hip_create_error());
        return;
    }
    const int64_t N = L.flou_hip_ndofs(h);
    const size_t bytes = L.flou_hip_state_bytes(h);
    std::vector<double> x = node_x(o, m), q((size_t)N * 3), r((size_t)N * 3);
    void *Q = nullptr, *dQ = nullptr;
    L.flou_hip_alloc(h, &Q, bytes);
    L.flou_hip_alloc(h, &dQ, bytes);
    // free stream: ρ = 1.2, u = 0.3, p = 2 -> rhs = 0
    for (int64_t i = 0; i < N; ++i) {
        q[i] = 1.2;
        q[N + i] = 1.2 * 0.3;
        q[2 * N + i] = 2.0 / 0.4 + 0.5 * 1.2 * 0.09;
    }
    L.flou_hip_upload(h, Q, q.data(), bytes);
    L.flou_hip_rhs(h, dQ, Q, 0.0);
    L.flou_hip_download(h, r.data(), dQ, bytes);
    L.flou_hip_synchronize(h);
    double fs = 0;
    for (double v : r) fs = std::fmax(fs, std::fabs(v));
    check(L.flou_hip_check(h) == FLOU_HIP_OK && fs <= 1e-12, "euler: constant state has a zero rhs", fs);
    // smooth periodic state: integrals over 20 SSPRK33 steps
    const double pi = 3.141592653589793;
    for (int64_t i = 0; i < N; ++i) {
        const double rho = 1 + 0.2 * std::sin(2 * pi * x[i]), u = 0.3 + 0.1 * std::cos(2 * pi * x[i]);
        const double p = 1 + 0.1 * std::sin(4 * pi * x[i] + 0.5);
        q[i] = rho;
        q[N + i] = rho * u;
        q[2 * N + i] = p / 0.4 + 0.5 * rho * u * u;
    }
    L.flou_hip_upload(h, Q, q.data(), bytes);
    double dt = 0;
    int rc = L.flou_hip_max_dt(h, Q, 0.3, &dt);
    check(rc == FLOU_HIP_OK && dt > 0 && std::isfinite(dt), "euler: get_max_dt(cfl 0.3) positive", dt);
    double a[3], b[3], t = 0;
    for (int v = 0; v < 3; ++v) L.flou_hip_integrate(h, Q, v, &a[v]);
    rc = L.flou_hip_timeintegrate(h, FLOU_HIP_SSPRK33, Q, dt, 20, 0.0, &t);
    for (int v = 0; v < 3; ++v) L.flou_hip_integrate(h, Q, v, &b[v]);
    double drift = 0;
    for (int v = 0; v < 3; ++v) drift = std::fmax(drift, std::fabs(b[v] - a[v]) / std::fabs(a[v]));
    check(rc == FLOU_HIP_OK && drift <= 1e-13, "euler: 20 SSPRK33 steps conserve mass, momentum, energy", drift);
    // negative pressure -> DomainError
    for (int64_t i = 0; i < N; ++i) q[2 * N + i] = 0.5 * q[N + i] * q[N + i] / q[i] - 0.1;
    L.flou_hip_upload(h, Q, q.data(), bytes);
    rc = L.flou_hip_timeintegrate(h, FLOU_HIP_SSPRK33, Q, dt, 1, 0.0, &t);
    const std::string msg = L.flou_hip_last_error(h);
    check(rc == FLOU_HIP_EDOMAIN && msg.find("DomainError") != std::string::npos,
          "euler: negative pressure is a domain error", rc);
    L.flou_hip_free(h, Q);
    L.flou_hip_free(h, dQ);
    L.flou_hip_destroy(h);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s LIBFLOUHIP [--no-gpu]\n", argv[0]);
        return 2;
    }
    Lib L;
    if (!L.load(argv[1])) return 1;
    const bool gpu = !(argc > 2 && std::strcmp(argv[2], "--no-gpu") == 0);
    if (!no_gpu_checks(L)) return 1;
    if (gpu) {
        advection(L);
        euler(L);
    }
    std::printf("%s: %d failure(s)\n", gpu ? "C ABI driver" : "C ABI driver (no GPU)", failures);
    return failures ? 1 : 0;
}
