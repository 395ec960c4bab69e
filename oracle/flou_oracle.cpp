// Flou.jl DGSEM oracle: setup restatements and the C ABI used by the tests.
//
// TEST INFRASTRUCTURE ONLY (see flou_oracle.hpp).
#include "flou_oracle.hpp"

#include <map>

#include <omp.h>

namespace oracle {

// ------------------------------------------------------------------------------------
// Operators
typedef long double LD;

static void legendre(int N, LD x, LD& P, LD& dP) {
    // P_N(x) and P_N'(x) by the three-term recurrence
    LD p0 = 1.0L, p1 = x;
    if (N == 0) { P = 1.0L; dP = 0.0L; return; }
    for (int k = 2; k <= N; ++k) {
        LD p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
        p0 = p1; p1 = p2;
    }
    P = p1;
    dP = (std::fabs((double)x) == 1.0) ? 0.5L * N * (N + 1) * ((x > 0) ? 1.0L : ((N % 2) ? 1.0L : -1.0L))
                                       : N * (x * p1 - p0) / (x * x - 1.0L);
}

Ops build_ops(int nodetype, int n) {
    if (n < 2 || n > 16) throw std::invalid_argument("n must be in [2, 16]");
    std::vector<LD> x(n), w(n);
    const LD pi = 3.141592653589793238462643383279502884L;
    if (nodetype == NODES_GLL) {
        // FastGaussQuadrature.gausslobatto: ±1 and the roots of P'_{n-1}
        const int N = n - 1;
        x[0] = -1.0L; x[N] = 1.0L;
        for (int i = 1; i < N; ++i) {
            LD xi = -std::cos(pi * (LD)i / (LD)N);  // Chebyshev-Gauss-Lobatto guess
            for (int it = 0; it < 100; ++it) {
                // Newton on q(x) = P'_N(x); q' = (2x P'_N - N(N+1) P_N)/(1-x^2)
                LD P, dP;
                legendre(N, xi, P, dP);
                LD d2P = (2.0L * xi * dP - N * (N + 1.0L) * P) / (1.0L - xi * xi);
                LD dx = dP / d2P;
                xi -= dx;
                if (std::fabs((double)dx) < 1e-22) break;
            }
            x[i] = xi;
        }
        for (int i = 0; i < n; ++i) {
            LD P, dP;
            legendre(N, x[i], P, dP);
            w[i] = 2.0L / (N * (N + 1.0L) * P * P);
        }
    } else if (nodetype == NODES_GL) {
        for (int i = 0; i < n; ++i) {
            LD xi = -std::cos(pi * (i + 0.75L) / (n + 0.5L));
            for (int it = 0; it < 100; ++it) {
                LD P, dP;
                legendre(n, xi, P, dP);
                LD dx = P / dP;
                xi -= dx;
                if (std::fabs((double)dx) < 1e-22) break;
            }
            x[i] = xi;
            LD P, dP;
            legendre(n, xi, P, dP);
            w[i] = 2.0L / ((1.0L - xi * xi) * dP * dP);
        }
    } else {
        throw std::invalid_argument("unknown node type");
    }
    // symmetrise (the nodes are symmetric about 0)
    for (int i = 0; i < n / 2; ++i) {
        LD m = 0.5L * (x[n - 1 - i] - x[i]);
        x[i] = -m; x[n - 1 - i] = m;
        LD ww = 0.5L * (w[i] + w[n - 1 - i]);
        w[i] = ww; w[n - 1 - i] = ww;
    }
    if (n % 2) x[n / 2] = 0.0L;
    // barycentric weights
    std::vector<LD> bw(n, 1.0L);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k)
            if (k != j) bw[j] /= (x[j] - x[k]);
    Ops o;
    o.n = n;
    o.xl = x;
    o.bw = bw;
    o.xi.resize(n); o.w.resize(n); o.D.resize(n * n);
    o.lm.resize(n); o.lp.resize(n); o.gm.resize(n); o.gp.resize(n);
    o.Ds.resize(n * n); o.Dsh.resize(n * n);
    for (int i = 0; i < n; ++i) { o.xi[i] = (double)x[i]; o.w[i] = (double)w[i]; }
    for (int i = 0; i < n; ++i) {
        LD diag = 0.0L;
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            LD Dij = (bw[j] / bw[i]) / (x[i] - x[j]);
            o.D[i * n + j] = (double)Dij;
            diag -= Dij;
        }
        o.D[i * n + i] = (double)diag;
    }
    // l = interpolation at ξ = ∓1
    for (int side = 0; side < 2; ++side) {
        const LD xe = side == 0 ? -1.0L : 1.0L;
        std::vector<double>& l = side == 0 ? o.lm : o.lp;
        int hit = -1;
        for (int j = 0; j < n; ++j) if (x[j] == xe) hit = j;
        if (hit >= 0) {
            for (int j = 0; j < n; ++j) l[j] = (j == hit) ? 1.0 : 0.0;
        } else {
            LD den = 0.0L;
            for (int j = 0; j < n; ++j) den += bw[j] / (xe - x[j]);
            for (int j = 0; j < n; ++j) l[j] = (double)((bw[j] / (xe - x[j])) / den);
        }
    }
    // DGSEMrec (Reconstruction.jl:125-134) with the sign flip of StdSegment.jl:84-85:
    // ∂g = (l(-1) ./ ω, l(+1) ./ ω)
    for (int j = 0; j < n; ++j) {
        o.gm[j] = -(-o.lm[j] / o.w[j]);
        o.gp[j] = o.lp[j] / o.w[j];
    }
    // B = ∂g[2] l[2]' - ∂g[1] l[1]';  Ds = D - B;  D♯ = 2D - B   (StdSegment.jl:87-89)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double B = o.gp[i] * o.lp[j] - o.gm[i] * o.lm[j];
            o.Ds[i * n + j] = o.D[i * n + j] - B;
            o.Dsh[i * n + j] = 2.0 * o.D[i * n + j] - B;
        }
    return o;
}

// ------------------------------------------------------------------------------------
// Cartesian mesh (CartesianMesh.jl:38-500) and periodic rewiring (Mesh.jl:236-322).
// Built with the reference's 1-based numbering, converted to 0-based at the end.
struct MFace { std::vector<long> nodes; long elem[2]; int pos[2]; };
struct MElem { std::vector<long> nodes; std::vector<long> faces; std::vector<int> side; };

static double range_point(double a, double b, int i, int n) {
    // range(a, b, n+1)[i+1], rounded once from extended precision
    return (double)((LD)a + ((LD)b - (LD)a) * (LD)i / (LD)n);
}

void build_cartesian(Disc& d, int nd, const int* nel, const double* start, const double* finish,
                     const int* periodic) {
    d.nd = nd;
    for (int k = 0; k < 3; ++k) { d.nel[k] = 1; d.start[k] = 0; d.finish[k] = 0; d.dx[k] = 0; }
    for (int k = 0; k < nd; ++k) {
        d.nel[k] = nel[k]; d.start[k] = start[k]; d.finish[k] = finish[k];
        if (!(start[k] < finish[k])) throw std::invalid_argument("start must be < finish");
        d.dx[k] = (finish[k] - start[k]) / nel[k];
    }
    const long nx = d.nel[0], ny = d.nel[1], nz = d.nel[2];
    const long npx = nx + 1, npy = ny + 1;
    // vertices
    d.vtx.clear();
    long nvtx = (nd >= 1 ? npx : 1) * (nd >= 2 ? npy : 1) * (nd >= 3 ? nz + 1 : 1);
    d.vtx.resize((size_t)nvtx * nd);
    {
        long iv = 0;
        for (long k = 0; k < (nd >= 3 ? nz + 1 : 1); ++k)
            for (long j = 0; j < (nd >= 2 ? npy : 1); ++j)
                for (long i = 0; i < npx; ++i, ++iv) {
                    d.vtx[iv * nd + 0] = range_point(start[0], finish[0], (int)i, (int)nx);
                    if (nd >= 2) d.vtx[iv * nd + 1] = range_point(start[1], finish[1], (int)j, (int)ny);
                    if (nd >= 3) d.vtx[iv * nd + 2] = range_point(start[2], finish[2], (int)k, (int)nz);
                }
    }
    std::vector<MElem> E;
    std::vector<MFace> F;
    std::vector<std::vector<long>> bdfaces(2 * nd);
    std::vector<long> intfaces;
    if (nd == 1) {
        E.resize(nx);
        for (long ie = 1; ie <= nx; ++ie) {
            MElem& m = E[ie - 1];
            m.nodes = {ie, ie + 1};
            m.faces = {ie, ie + 1};
            m.side = {(ie == 1) ? 1 : 2, 1};
        }
        long nfc = nx + 1;
        F.resize(nfc);
        for (long f = 1; f <= nfc; ++f) {
            MFace& fc = F[f - 1];
            if (f == 1) { bdfaces[0].push_back(f); fc.elem[0] = 1; fc.elem[1] = 0; fc.pos[0] = 1; fc.pos[1] = 0; }
            else if (f == nfc) { bdfaces[1].push_back(f); fc.elem[0] = nx; fc.elem[1] = 0; fc.pos[0] = 2; fc.pos[1] = 0; }
            else { intfaces.push_back(f); fc.elem[0] = f - 1; fc.elem[1] = f; fc.pos[0] = 2; fc.pos[1] = 1; }
            fc.nodes = {f};
        }
    } else if (nd == 2) {
        E.resize(nx * ny);
        for (long j = 1; j <= ny; ++j)
            for (long i = 1; i <= nx; ++i) {
                long ie = (j - 1) * nx + i;
                MElem& m = E[ie - 1];
                m.nodes = {(j - 1) * npx + i, (j - 1) * npx + i + 1, (j - 1) * npx + i + 1 + npx,
                           (j - 1) * npx + i + npx};
                m.faces = {(j - 1) * npx + i, (j - 1) * npx + i + 1, npx * ny + (j - 1) * nx + i,
                           npx * ny + (j - 1) * nx + i + nx};
                m.side = {(i == 1) ? 1 : 2, 1, (j == 1) ? 1 : 2, 1};
            }
        long nfc = npx * ny + npy * nx;
        F.resize(nfc);
        for (long j = 1; j <= ny; ++j)
            for (long i = 1; i <= npx; ++i) {
                long f = (j - 1) * npx + i;
                MFace& fc = F[f - 1];
                if (i == 1) { bdfaces[0].push_back(f); fc.elem[0] = (j - 1) * nx + 1; fc.elem[1] = 0; fc.pos[0] = 1; fc.pos[1] = 0; }
                else if (i == npx) { bdfaces[1].push_back(f); fc.elem[0] = (j - 1) * nx + nx; fc.elem[1] = 0; fc.pos[0] = 2; fc.pos[1] = 0; }
                else { intfaces.push_back(f); fc.elem[0] = (j - 1) * nx + i - 1; fc.elem[1] = (j - 1) * nx + i; fc.pos[0] = 2; fc.pos[1] = 1; }
                fc.nodes = {(j - 1) * npx + i, (j - 1) * npx + i + npx};
            }
        for (long j = 1; j <= npy; ++j)
            for (long i = 1; i <= nx; ++i) {
                long f = npx * ny + (j - 1) * nx + i;
                MFace& fc = F[f - 1];
                if (j == 1) { bdfaces[2].push_back(f); fc.elem[0] = i; fc.elem[1] = 0; fc.pos[0] = 3; fc.pos[1] = 0; }
                else if (j == npy) { bdfaces[3].push_back(f); fc.elem[0] = (ny - 1) * nx + i; fc.elem[1] = 0; fc.pos[0] = 4; fc.pos[1] = 0; }
                else { intfaces.push_back(f); fc.elem[0] = (j - 2) * nx + i; fc.elem[1] = (j - 2) * nx + i + nx; fc.pos[0] = 4; fc.pos[1] = 3; }
                fc.nodes = {(j - 1) * npx + i, (j - 1) * npx + i + 1};
            }
    } else {
        E.resize(nx * ny * nz);
        for (long k = 1; k <= nz; ++k)
            for (long j = 1; j <= ny; ++j)
                for (long i = 1; i <= nx; ++i) {
                    long ie = (k - 1) * nx * ny + (j - 1) * nx + i;
                    MElem& m = E[ie - 1];
                    long b = (k - 1) * npx * npy + (j - 1) * npx + i;
                    m.nodes = {b, b + 1, b + 1 + npx, b + npx, npx * npy + b, npx * npy + b + 1,
                               npx * npy + b + 1 + npx, npx * npy + b + npx};
                    m.faces = {(k - 1) * npx * ny + (j - 1) * npx + i,
                               (k - 1) * npx * ny + (j - 1) * npx + i + 1,
                               npx * ny * nz + (k - 1) * nx * npy + (j - 1) * nx + i,
                               npx * ny * nz + (k - 1) * nx * npy + (j - 1) * nx + i + nx,
                               npx * ny * nz + nx * npy * nz + (k - 1) * nx * ny + (j - 1) * nx + i,
                               npx * ny * nz + nx * npy * nz + (k - 1) * nx * ny + (j - 1) * nx + i + nx * ny};
                    m.side = {(i == 1) ? 1 : 2, 1, (j == 1) ? 1 : 2, 1, (k == 1) ? 1 : 2, 1};
                }
        const long npz = nz + 1;
        long nfc = npx * ny * nz + npy * nx * nz + npz * nx * ny;
        F.resize(nfc);
        for (long k = 1; k <= nz; ++k)
            for (long j = 1; j <= ny; ++j)
                for (long i = 1; i <= npx; ++i) {
                    long f = (k - 1) * npx * ny + (j - 1) * npx + i;
                    MFace& fc = F[f - 1];
                    long eb = (k - 1) * nx * ny + (j - 1) * nx;
                    if (i == 1) { bdfaces[0].push_back(f); fc.elem[0] = eb + 1; fc.elem[1] = 0; fc.pos[0] = 1; fc.pos[1] = 0; }
                    else if (i == npx) { bdfaces[1].push_back(f); fc.elem[0] = eb + nx; fc.elem[1] = 0; fc.pos[0] = 2; fc.pos[1] = 0; }
                    else { intfaces.push_back(f); fc.elem[0] = eb + i - 1; fc.elem[1] = eb + i; fc.pos[0] = 2; fc.pos[1] = 1; }
                    long b = (k - 1) * npx * npy + (j - 1) * npx + i;
                    fc.nodes = {b, b + npx, npx * npy + b + npx, npx * npy + b};
                }
        for (long k = 1; k <= nz; ++k)
            for (long j = 1; j <= npy; ++j)
                for (long i = 1; i <= nx; ++i) {
                    long f = npx * ny * nz + (k - 1) * nx * npy + (j - 1) * nx + i;
                    MFace& fc = F[f - 1];
                    if (j == 1) { bdfaces[2].push_back(f); fc.elem[0] = (k - 1) * nx * ny + i; fc.elem[1] = 0; fc.pos[0] = 3; fc.pos[1] = 0; }
                    else if (j == npy) { bdfaces[3].push_back(f); fc.elem[0] = (k - 1) * nx * ny + (ny - 1) * nx + i; fc.elem[1] = 0; fc.pos[0] = 4; fc.pos[1] = 0; }
                    else { intfaces.push_back(f); fc.elem[0] = (k - 1) * nx * ny + (j - 2) * nx + i; fc.elem[1] = (k - 1) * nx * ny + (j - 2) * nx + i + nx; fc.pos[0] = 4; fc.pos[1] = 3; }
                    long b = (k - 1) * npx * npy + (j - 1) * npx + i;
                    fc.nodes = {b, b + 1, npx * npy + b + 1, npx * npy + b};
                }
        for (long k = 1; k <= npz; ++k)
            for (long j = 1; j <= ny; ++j)
                for (long i = 1; i <= nx; ++i) {
                    long f = npx * ny * nz + npy * nx * nz + (k - 1) * nx * ny + (j - 1) * nx + i;
                    MFace& fc = F[f - 1];
                    if (k == 1) { bdfaces[4].push_back(f); fc.elem[0] = (j - 1) * nx + i; fc.elem[1] = 0; fc.pos[0] = 5; fc.pos[1] = 0; }
                    else if (k == npz) { bdfaces[5].push_back(f); fc.elem[0] = (nz - 1) * nx * ny + (j - 1) * nx + i; fc.elem[1] = 0; fc.pos[0] = 6; fc.pos[1] = 0; }
                    else { intfaces.push_back(f); fc.elem[0] = (k - 2) * nx * ny + (j - 1) * nx + i; fc.elem[1] = (k - 2) * nx * ny + (j - 1) * nx + i + nx * ny; fc.pos[0] = 6; fc.pos[1] = 5; }
                    long b = (k - 1) * npx * npy + (j - 1) * npx + i;
                    fc.nodes = {b, b + 1, npx + b + 1, npx + b};
                }
    }

    // Periodic rewiring (Mesh.jl:236-322).  Boundary ids keep their 1-based names "1".."2nd".
    std::vector<int> bdalive(2 * nd, 1);
    std::vector<long> faces2del;
    for (int k = 0; k < nd; ++k) {
        if (!periodic[k]) continue;
        const int b1 = 2 * k, b2 = 2 * k + 1;
        for (size_t q = 0; q < bdfaces[b1].size(); ++q) {
            long if1 = bdfaces[b1][q], if2 = bdfaces[b2][q];
            MFace& f1 = F[if1 - 1];
            MFace& f2 = F[if2 - 1];
            long elmind = f2.elem[0];
            int elmpos = f2.pos[0];
            f1.elem[1] = elmind;
            f1.pos[1] = elmpos;
            MElem& el = E[elmind - 1];
            el.faces[elmpos - 1] = if1;
            el.side[elmpos - 1] = 2;
        }
        intfaces.insert(intfaces.end(), bdfaces[b1].begin(), bdfaces[b1].end());
        faces2del.insert(faces2del.end(), bdfaces[b2].begin(), bdfaces[b2].end());
        bdalive[b1] = bdalive[b2] = 0;
    }
    std::sort(faces2del.begin(), faces2del.end());
    // new index of each surviving face = old index - #deleted faces before it
    std::vector<long> newidx(F.size() + 1, 0);
    {
        size_t q = 0;
        long shift = 0;
        for (long f = 1; f <= (long)F.size(); ++f) {
            while (q < faces2del.size() && faces2del[q] < f) { ++q; ++shift; }
            bool del = (q < faces2del.size() && faces2del[q] == f);
            newidx[f] = del ? 0 : f - shift;
        }
    }
    std::vector<MFace> F2;
    std::vector<int> bdof(F.size() + 1, -1);
    {
        int ib = 0;
        for (int b = 0; b < 2 * nd; ++b) {
            if (!bdalive[b]) continue;
            for (long f : bdfaces[b]) bdof[f] = ib;
            d.bdnames.push_back(std::to_string(b + 1));
            ++ib;
        }
    }
    std::vector<int> fbd;
    for (long f = 1; f <= (long)F.size(); ++f)
        if (newidx[f]) { F2.push_back(F[f - 1]); fbd.push_back(bdof[f]); }
    for (long f = 1; f <= (long)F2.size(); ++f) {
        MFace& fc = F2[f - 1];
        E[fc.elem[0] - 1].faces[fc.pos[0] - 1] = f;
        if (fc.elem[1] != 0) E[fc.elem[1] - 1].faces[fc.pos[1] - 1] = f;
    }

    // 0-based flat arrays
    d.ne = (long)E.size();
    d.nf = (long)F2.size();
    d.e_face.assign(d.ne * 2 * nd, 0);
    d.e_side.assign(d.ne * 2 * nd, 0);
    for (long e = 0; e < d.ne; ++e)
        for (int p = 0; p < 2 * nd; ++p) {
            d.e_face[e * 2 * nd + p] = E[e].faces[p] - 1;
            d.e_side[e * 2 * nd + p] = E[e].side[p] - 1;
        }
    d.f_elem.assign(d.nf * 2, -1);
    d.f_pos.assign(d.nf * 2, -1);
    d.f_orient.assign(d.nf, 0);
    d.f_bd = fbd;
    for (long f = 0; f < d.nf; ++f)
        for (int s = 0; s < 2; ++s) {
            d.f_elem[f * 2 + s] = F2[f].elem[s] - 1;
            d.f_pos[f * 2 + s] = F2[f].elem[s] ? F2[f].pos[s] - 1 : -1;
        }
    d.bctype.assign(d.bdnames.size(), BC_OUTFLOW);
    d.bcdata.assign(d.bdnames.size(), std::vector<double>());

    // ---------------- geometry (PhysicalRegions.jl:370-408, 541-696)
    const int n = d.n;
    d.npe = 1; d.npf = 1;
    for (int k = 0; k < nd; ++k) d.npe *= n;
    for (int k = 0; k < nd - 1; ++k) d.npf *= n;
    double pdx = d.dx[0];
    for (int k = 1; k < nd; ++k) pdx = pdx * d.dx[k];
    const double jac = pdx / (double)(1 << nd);
    d.e_jac.assign(d.ne, jac);
    d.e_metric.assign(d.ne * nd * nd, 0.0);
    d.e_vol.assign(d.ne, 0.0);
    for (long e = 0; e < d.ne; ++e) {
        double* M = &d.e_metric[e * nd * nd];
        if (nd == 1) M[0] = 1.0;
        else if (nd == 2) { M[0] = d.dx[1] / 2.0; M[3] = d.dx[0] / 2.0; }
        else { M[0] = d.dx[1] * d.dx[2] / 4.0; M[4] = d.dx[0] * d.dx[2] / 4.0; M[8] = d.dx[0] * d.dx[1] / 4.0; }
        double vol = 0.0;
        for (int i = 0; i < d.npe; ++i) {
            int idx = i;
            double w = 1.0;
            double wi[3];
            for (int k = 0; k < nd; ++k) { wi[k] = d.ops.w[idx % n]; idx /= n; }
            w = wi[0];
            for (int k = 1; k < nd; ++k) w = w * wi[k];
            vol = vol + jac * w;
        }
        d.e_vol[e] = vol;
    }
    d.f_frame.assign(d.nf * 3 * nd, 0.0);
    d.f_jac.assign(d.nf, 0.0);
    for (long f = 0; f < d.nf; ++f) {
        const int pos = d.f_pos[f * 2];  // master position, 0-based
        double* fr = &d.f_frame[f * 3 * nd];
        double* nn = fr; double* t = fr + nd; double* b = fr + 2 * nd;
        if (nd == 1) {
            d.f_jac[f] = 1.0;
            nn[0] = pos == 0 ? -1.0 : 1.0;
        } else if (nd == 2) {
            if (pos <= 1) {
                d.f_jac[f] = d.dx[1] / 2.0;
                if (pos == 0) { nn[0] = -1; t[1] = -1; } else { nn[0] = 1; t[1] = 1; }
            } else {
                d.f_jac[f] = d.dx[0] / 2.0;
                if (pos == 2) { nn[1] = -1; t[0] = 1; } else { nn[1] = 1; t[0] = -1; }
            }
        } else {
            if (pos <= 1) {
                d.f_jac[f] = d.dx[1] * d.dx[2] / 4.0;
                if (pos == 0) { nn[0] = -1; t[1] = -1; b[2] = 1; } else { nn[0] = 1; t[1] = 1; b[2] = 1; }
            } else if (pos <= 3) {
                d.f_jac[f] = d.dx[0] * d.dx[2] / 4.0;
                if (pos == 2) { nn[1] = -1; t[2] = -1; b[0] = 1; } else { nn[1] = 1; t[2] = 1; b[0] = 1; }
            } else {
                d.f_jac[f] = d.dx[0] * d.dx[1] / 4.0;
                if (pos == 4) { nn[2] = -1; t[0] = -1; b[1] = 1; } else { nn[2] = 1; t[0] = 1; b[1] = 1; }
            }
        }
    }
    // node coordinates (Mesh.jl:380-446 linear maps)
    d.coords.assign((size_t)d.ne * d.npe * nd, 0.0);
    for (long e = 0; e < d.ne; ++e) {
        const std::vector<long>& vn = E[e].nodes;
        for (int i = 0; i < d.npe; ++i) {
            int idx = i;
            double r[3];
            for (int k = 0; k < nd; ++k) { r[k] = (d.ops.xi[idx % n] + 1.0) / 2.0; idx /= n; }
            for (int c = 0; c < nd; ++c) {
                double x;
                auto V = [&](int q) { return d.vtx[(vn[q] - 1) * nd + c]; };
                if (nd == 1) x = V(0) * (1.0 - r[0]) + V(1) * r[0];
                else if (nd == 2)
                    x = ((V(0) * (1 - r[0]) * (1 - r[1]) + V(1) * r[0] * (1 - r[1])) + V(2) * r[0] * r[1]) + V(3) * (1 - r[0]) * r[1];
                else {
                    x = V(0) * (1 - r[0]) * (1 - r[1]) * (1 - r[2]);
                    x = x + V(1) * r[0] * (1 - r[1]) * (1 - r[2]);
                    x = x + V(2) * r[0] * r[1] * (1 - r[2]);
                    x = x + V(3) * (1 - r[0]) * r[1] * (1 - r[2]);
                    x = x + V(4) * (1 - r[0]) * (1 - r[1]) * r[2];
                    x = x + V(5) * r[0] * (1 - r[1]) * r[2];
                    x = x + V(6) * r[0] * r[1] * r[2];
                    x = x + V(7) * (1 - r[0]) * r[1] * r[2];
                }
                d.coords[((size_t)e * d.npe + i) * nd + c] = x;
            }
        }
    }
    // face node coordinates (face mappings: Point / Segment / Quad linear)
    d.f_coords.assign((size_t)d.nf * d.npf * nd, 0.0);
    for (long f = 0; f < d.nf; ++f) {
        const std::vector<long>& vn = F2[f].nodes;
        for (int i = 0; i < d.npf; ++i) {
            int idx = i;
            double r[2] = {0, 0};
            for (int k = 0; k < nd - 1; ++k) { r[k] = (d.ops.xi[idx % n] + 1.0) / 2.0; idx /= n; }
            for (int c = 0; c < nd; ++c) {
                auto V = [&](int q) { return d.vtx[(vn[q] - 1) * nd + c]; };
                double x;
                if (nd == 1) x = V(0);
                else if (nd == 2) x = V(0) * (1.0 - r[0]) + V(1) * r[0];
                else x = ((V(0) * (1 - r[0]) * (1 - r[1]) + V(1) * r[0] * (1 - r[1])) + V(2) * r[0] * r[1]) + V(3) * (1 - r[0]) * r[1];
                d.f_coords[((size_t)f * d.npf + i) * nd + c] = x;
            }
        }
    }
}

Tableau2N tableau_2n(int scheme) {
    Tableau2N t;
    if (scheme == RK_CK2N54) {
        t.s = 5;
        t.A[0] = 0.0;
        t.A[1] = -567301805773.0 / 1357537059087.0;
        t.A[2] = -2404267990393.0 / 2016746695238.0;
        t.A[3] = -3550918686646.0 / 2091501179385.0;
        t.A[4] = -1275806237668.0 / 842570457699.0;
        t.B[0] = 1432997174477.0 / 9575080441755.0;
        t.B[1] = 5161836677717.0 / 13612068292357.0;
        t.B[2] = 1720146321549.0 / 2090206949498.0;
        t.B[3] = 3134564353537.0 / 4481467310338.0;
        t.B[4] = 2277821191437.0 / 14882151754819.0;
        t.c[0] = 0.0;
        t.c[1] = 1432997174477.0 / 9575080441755.0;
        t.c[2] = 2526269341429.0 / 6820363962896.0;
        t.c[3] = 2006345519317.0 / 3224310063776.0;
        t.c[4] = 2802321613138.0 / 2924317926251.0;
    } else if (scheme == RK_ORK256) {
        t.s = 5;
        t.A[0] = 0.0; t.A[1] = -1.0; t.A[2] = -1.55798; t.A[3] = -1.0; t.A[4] = -0.45031;
        t.B[0] = 0.2; t.B[1] = 0.83204; t.B[2] = 0.6; t.B[3] = 0.35394; t.B[4] = 0.2;
        t.c[0] = 0.0; t.c[1] = 0.2; t.c[2] = 0.2; t.c[3] = 0.8; t.c[4] = 0.8;
    } else {
        throw std::invalid_argument("not a 2N scheme");
    }
    return t;
}

}  // namespace oracle

// ======================================================================================
// C ABI for ctypes (tests and the CPU baseline only)
using namespace oracle;

struct oracle_handle {
    Disc d;
    Work<double> wk;
    std::string err;
};

// ------------------------------------------------------------------------------------
// Output to equispaced points

void oracle::interp_matrix(const Ops& o, int m, const double* xe, double* A) {
    const int n = o.n;
    for (int i = 0; i < m; ++i) {
        const LD t = (LD)xe[i];
        int hit = -1;
        for (int j = 0; j < n; ++j) if (o.xl[j] == t) hit = j;
        if (hit >= 0) {
            for (int j = 0; j < n; ++j) A[i * n + j] = (j == hit) ? 1.0 : 0.0;
            continue;
        }
        LD den = 0.0L;
        for (int j = 0; j < n; ++j) den += o.bw[j] / (t - o.xl[j]);
        for (int j = 0; j < n; ++j) A[i * n + j] = (double)((o.bw[j] / (t - o.xl[j])) / den);
    }
}

void oracle::equispaced(int nd, int n, int m, const double* A, int64_t ne, int ncols, const double* Q,
                int64_t ldq, double* Qe, int64_t lde) {
    int npe = 1, mpe = 1;
    for (int k = 0; k < nd; ++k) { npe *= n; mpe *= m; }
    // node2eq, row-major mpe x npe
    std::vector<double> K((size_t)mpe * npe);
    for (int i = 0; i < mpe; ++i)
        for (int j = 0; j < npe; ++j) {
            const int i1 = i % m, j1 = j % n;
            if (nd == 1) { K[(size_t)i * npe + j] = A[i1 * n + j1]; continue; }
            const int i2 = (i / m) % m, j2 = (j / n) % n;
            if (nd == 2) { K[(size_t)i * npe + j] = A[i2 * n + j2] * A[i1 * n + j1]; continue; }
            const int i3 = i / (m * m), j3 = j / (n * n);
            K[(size_t)i * npe + j] = (A[i3 * n + j3] * A[i2 * n + j2]) * A[i1 * n + j1];
        }
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e)
        for (int c = 0; c < ncols; ++c)
            for (int i = 0; i < mpe; ++i) {
                double acc = 0.0;
                for (int j = 0; j < npe; ++j) acc += K[(size_t)i * npe + j] * Q[c * ldq + e * npe + j];
                Qe[c * lde + e * mpe + i] = acc;
            }
}

std::vector<int64_t> oracle::vtk_connectivities(int nd, int n) {
    // li(i, j, k): 1-based lattice position -> 0-based linear index (LinearIndices)
    auto li = [n](int i, int j = 1, int k = 1) -> int64_t { return (i - 1) + (int64_t)n * ((j - 1) + (int64_t)n * (k - 1)); };
    std::vector<int64_t> c;
    if (nd == 1) {
        // [1, n, 2..n-1] .- 1
        c.push_back(li(1)); c.push_back(li(n));
        for (int i = 2; i <= n - 1; ++i) c.push_back(li(i));
    } else if (nd == 2) {
        // corners (1,1) (n,1) (n,n) (1,n); edges y=1, x=n, y=n, x=1; interior x fastest
        c = {li(1, 1), li(n, 1), li(n, n), li(1, n)};
        for (int i = 2; i <= n - 1; ++i) c.push_back(li(i, 1));
        for (int j = 2; j <= n - 1; ++j) c.push_back(li(n, j));
        for (int i = 2; i <= n - 1; ++i) c.push_back(li(i, n));
        for (int j = 2; j <= n - 1; ++j) c.push_back(li(1, j));
        for (int j = 2; j <= n - 1; ++j)
            for (int i = 2; i <= n - 1; ++i) c.push_back(li(i, j));
    } else {
        c = {li(1, 1, 1), li(n, 1, 1), li(n, n, 1), li(1, n, 1),
             li(1, 1, n), li(n, 1, n), li(n, n, n), li(1, n, n)};
        auto run_i = [&](int j, int k) { for (int i = 2; i <= n - 1; ++i) c.push_back(li(i, j, k)); };
        auto run_j = [&](int i, int k) { for (int j = 2; j <= n - 1; ++j) c.push_back(li(i, j, k)); };
        auto run_k = [&](int i, int j) { for (int k = 2; k <= n - 1; ++k) c.push_back(li(i, j, k)); };
        // edges: the four of z=1, the four of z=n, then the four along z
        run_i(1, 1); run_j(n, 1); run_i(n, 1); run_j(1, 1);
        run_i(1, n); run_j(n, n); run_i(n, n); run_j(1, n);
        run_k(1, 1); run_k(n, 1); run_k(n, n); run_k(1, n);
        // faces x=1, x=n (y fastest), y=1, y=n (x fastest), z=1, z=n (x fastest)
        for (int xf : {1, n})
            for (int k = 2; k <= n - 1; ++k)
                for (int j = 2; j <= n - 1; ++j) c.push_back(li(xf, j, k));
        for (int yf : {1, n})
            for (int k = 2; k <= n - 1; ++k)
                for (int i = 2; i <= n - 1; ++i) c.push_back(li(i, yf, k));
        for (int zf : {1, n})
            for (int j = 2; j <= n - 1; ++j)
                for (int i = 2; i <= n - 1; ++i) c.push_back(li(i, j, zf));
        for (int k = 2; k <= n - 1; ++k)
            for (int j = 2; j <= n - 1; ++j)
                for (int i = 2; i <= n - 1; ++i) c.push_back(li(i, j, k));
    }
    return c;
}

extern "C" {

int oracle_interp_matrix(int nodetype, int n, int m, const double* xe, double* A) {
    try {
        Ops o = build_ops(nodetype, n);
        interp_matrix(o, m, xe, A);
        return 0;
    } catch (...) {
        return -1;
    }
}

int oracle_equispaced(int nd, int n, int m, const double* A, int64_t ne, int ncols, const double* Q,
                      double* Qe) {
    if (nd < 1 || nd > 3 || n < 1 || m < 1 || ne < 0 || ncols < 1) return -1;
    int64_t npe = 1, mpe = 1;
    for (int k = 0; k < nd; ++k) { npe *= n; mpe *= m; }
    equispaced(nd, n, m, A, ne, ncols, Q, ne * npe, Qe, ne * mpe);
    return 0;
}

// writes n^nd ids into out; returns their count
int64_t oracle_vtk_connectivities(int nd, int n, int64_t* out) {
    if (nd < 1 || nd > 3 || n < 2) return -1;
    std::vector<int64_t> c = vtk_connectivities(nd, n);
    std::copy(c.begin(), c.end(), out);
    return (int64_t)c.size();
}

int oracle_build_ops(int nodetype, int n, double* xi, double* w, double* D, double* Ds,
                     double* Dsh, double* lm, double* lp, double* gm, double* gp) {
    try {
        Ops o = build_ops(nodetype, n);
        std::memcpy(xi, o.xi.data(), n * sizeof(double));
        std::memcpy(w, o.w.data(), n * sizeof(double));
        std::memcpy(D, o.D.data(), n * n * sizeof(double));
        std::memcpy(Ds, o.Ds.data(), n * n * sizeof(double));
        std::memcpy(Dsh, o.Dsh.data(), n * n * sizeof(double));
        std::memcpy(lm, o.lm.data(), n * sizeof(double));
        std::memcpy(lp, o.lp.data(), n * sizeof(double));
        std::memcpy(gm, o.gm.data(), n * sizeof(double));
        std::memcpy(gp, o.gp.data(), n * sizeof(double));
        return 0;
    } catch (...) {
        return -1;
    }
}

// ops: packed [xi(n), w(n), D(n*n), Ds(n*n), Dsh(n*n), lm(n), lp(n), gm(n), gp(n)]
oracle_handle* oracle_create(int nd, int n, const int* nel, const double* start,
                             const double* finish, const int* periodic, const double* ops,
                             int eq, const double* eqparams, int volop, int tpflux, int savg,
                             int sdiss, double intensity) {
    oracle_handle* h = new oracle_handle();
    try {
        Disc& d = h->d;
        d.n = n;
        Ops& o = d.ops;
        o.n = n;
        const double* p = ops;
        o.xi.assign(p, p + n); p += n;
        o.w.assign(p, p + n); p += n;
        o.D.assign(p, p + n * n); p += n * n;
        o.Ds.assign(p, p + n * n); p += n * n;
        o.Dsh.assign(p, p + n * n); p += n * n;
        o.lm.assign(p, p + n); p += n;
        o.lp.assign(p, p + n); p += n;
        o.gm.assign(p, p + n); p += n;
        o.gp.assign(p, p + n); p += n;
        d.eq = eq;
        if (eq == EQ_EULER) { d.gamma = eqparams[0]; d.nv = nd + 2; }
        else { for (int k = 0; k < nd; ++k) d.a[k] = eqparams[k]; d.nv = 1; }
        d.volop = volop; d.tpflux = tpflux; d.savg = savg; d.sdiss = sdiss;
        d.intensity = intensity;
        build_cartesian(d, nd, nel, start, finish, periodic);
    } catch (std::exception& ex) {
        delete h;
        return nullptr;
    }
    return h;
}

void oracle_destroy(oracle_handle* h) { delete h; }

// Per-node geometry of a mapped (curved) mesh, computed by the caller: node_metric
// (ndofs*nd*nd, column-major per node), node_jac (ndofs), face_frames (nf*npf*3*nd, per
// face node of the master side), face_jac (nf*npf), elem_volume (ne).
void oracle_set_curved(oracle_handle* h, const double* node_metric, const double* node_jac,
                       const double* face_frames, const double* face_jac, const double* elem_volume) {
    Disc& d = h->d;
    const size_t N = (size_t)d.ndofs();
    d.curved = true;
    d.n_metric.assign(node_metric, node_metric + N * d.nd * d.nd);
    d.n_jac.assign(node_jac, node_jac + N);
    d.fn_frame.assign(face_frames, face_frames + (size_t)d.nf * d.npf * 3 * d.nd);
    d.fn_jac.assign(face_jac, face_jac + (size_t)d.nf * d.npf);
    d.e_vol.assign(elem_volume, elem_volume + d.ne);
}

// HybridDivOperator blending constant (the operator itself is volop = VOL_HYBRID).
void oracle_set_blend(oracle_handle* h, double blend) { h->d.blend = blend; }

void oracle_set_limiter(oracle_handle* h, int on, double minval) {
    h->d.limiter_on = on != 0;
    h->d.limiter_minval = minval;
}

// Zhang-Shu limiter applied once to Q (ndofs x nv, column-major), in place.
int oracle_limit(oracle_handle* h, double* Q, double minval) {
    zhang_shu(h->d, Q, minval);
    return 0;
}

int oracle_monitor(oracle_handle* h, const double* Q, int kind, double* out) {
    try {
        *out = monitor(h->d, Q, kind);
    } catch (std::domain_error& e) {
        h->err = e.what();
        return -5;
    }
    return 0;
}

void oracle_set_option(oracle_handle* h, int which, int value) {
    if (which == 0) h->d.chandrasekhar_surface_typo = value;
}

void oracle_sizes(oracle_handle* h, long* out) {
    // ne, nf, npe, npf, nv, nboundaries
    out[0] = h->d.ne; out[1] = h->d.nf; out[2] = h->d.npe; out[3] = h->d.npf;
    out[4] = h->d.nv; out[5] = (long)h->d.bdnames.size();
}

void oracle_connectivity(oracle_handle* h, long* e_face, int* e_side, long* f_elem, int* f_pos,
                         int* f_bd) {
    const Disc& d = h->d;
    std::memcpy(e_face, d.e_face.data(), d.e_face.size() * sizeof(long));
    std::memcpy(e_side, d.e_side.data(), d.e_side.size() * sizeof(int));
    std::memcpy(f_elem, d.f_elem.data(), d.f_elem.size() * sizeof(long));
    std::memcpy(f_pos, d.f_pos.data(), d.f_pos.size() * sizeof(int));
    std::memcpy(f_bd, d.f_bd.data(), d.f_bd.size() * sizeof(int));
}

void oracle_geometry(oracle_handle* h, double* e_jac, double* e_metric, double* e_vol,
                     double* f_frame, double* f_jac) {
    const Disc& d = h->d;
    std::memcpy(e_jac, d.e_jac.data(), d.e_jac.size() * sizeof(double));
    std::memcpy(e_metric, d.e_metric.data(), d.e_metric.size() * sizeof(double));
    std::memcpy(e_vol, d.e_vol.data(), d.e_vol.size() * sizeof(double));
    std::memcpy(f_frame, d.f_frame.data(), d.f_frame.size() * sizeof(double));
    std::memcpy(f_jac, d.f_jac.data(), d.f_jac.size() * sizeof(double));
}

// Replace the element/face connectivity of the same numbers of elements and faces (a mesh
// whose elements number their nodes differently: the rhs then reads face orientations).
void oracle_set_connectivity(oracle_handle* h, const long* e_face, const int* e_side, const long* f_elem,
                             const int* f_pos, const int* f_orient) {
    Disc& d = h->d;
    std::memcpy(d.e_face.data(), e_face, d.e_face.size() * sizeof(long));
    std::memcpy(d.e_side.data(), e_side, d.e_side.size() * sizeof(int));
    std::memcpy(d.f_elem.data(), f_elem, d.f_elem.size() * sizeof(long));
    std::memcpy(d.f_pos.data(), f_pos, d.f_pos.size() * sizeof(int));
    std::memcpy(d.f_orient.data(), f_orient, d.f_orient.size() * sizeof(int));
}

int oracle_master2slave(int nd, int n, int i, int o) { return master2slave(nd, n, i, o); }
int oracle_slave2master(int nd, int n, int i, int o) { return slave2master(nd, n, i, o); }

void oracle_coords(oracle_handle* h, double* coords, double* f_coords) {
    const Disc& d = h->d;
    if (coords) std::memcpy(coords, d.coords.data(), d.coords.size() * sizeof(double));
    if (f_coords) std::memcpy(f_coords, d.f_coords.data(), d.f_coords.size() * sizeof(double));
}

// Boundary condition on boundary index ib.  For BC_DIRICHLET, `data` holds one state per
// face node of every mesh face (nf * npf * nv, only boundary faces of ib are read).
int oracle_set_bc(oracle_handle* h, int ib, int type, const double* data) {
    Disc& d = h->d;
    if (ib < 0 || ib >= (int)d.bdnames.size()) return -1;
    d.bctype[ib] = type;
    if (type == BC_INFLOW) d.bcdata[ib].assign(data, data + d.nv);
    if (type == BC_DIRICHLET) {
        if (d.f_bcstate.empty()) d.f_bcstate.assign((size_t)d.nf * d.npf * d.nv, 0.0);
        for (long f = 0; f < d.nf; ++f)
            if (d.f_bd[f] == ib)
                for (int q = 0; q < d.npf * d.nv; ++q)
                    d.f_bcstate[(size_t)f * d.npf * d.nv + q] = data[(size_t)f * d.npf * d.nv + q];
    }
    return 0;
}

const char* oracle_last_error(oracle_handle* h) { return h->err.c_str(); }

int oracle_rhs(oracle_handle* h, const double* Q, double* dQ) {
    try {
        rhs(h->d, Q, dQ, h->wk);
    } catch (std::domain_error& e) {
        h->err = e.what();
        return -5;
    }
    return 0;
}

double oracle_max_dt(oracle_handle* h, const double* Q, double cfl) {
    try {
        return max_dt(h->d, Q, cfl);
    } catch (std::domain_error& e) {
        h->err = e.what();
        return -1.0;
    }
}

double oracle_integrate(oracle_handle* h, const double* Q, int ivar) {
    return integrate(h->d, Q, ivar);
}

int oracle_rk_run(oracle_handle* h, int scheme, double* u, double dt, int nsteps) {
    try {
        rk_run(h->d, scheme, u, dt, nsteps, h->wk);
    } catch (std::domain_error& e) {
        h->err = e.what();
        return -5;
    }
    return 0;
}

void oracle_tableau_2n(int scheme, double* A, double* B, double* c) {
    Tableau2N t = tableau_2n(scheme);
    for (int s = 0; s < t.s; ++s) { A[s] = t.A[s]; B[s] = t.B[s]; c[s] = t.c[s]; }
}

// Directional derivative (JVP) of rhs! with dual numbers: out = rhs(Q), dout = J(Q) dir.
int oracle_rhs_jvp(oracle_handle* h, const double* Q, const double* dir, double* out,
                   double* dout) {
    const Disc& d = h->d;
    const long M = d.ndofs() * d.nv;
    std::vector<Dual<1>> q((size_t)M), r((size_t)M);
    for (long i = 0; i < M; ++i) { q[i].v = Q[i]; q[i].d[0] = dir[i]; }
    Work<Dual<1>> wk;
    try {
        rhs(d, q.data(), r.data(), wk);
    } catch (std::domain_error& e) {
        h->err = e.what();
        return -5;
    }
    for (long i = 0; i < M; ++i) { out[i] = r[i].v; dout[i] = r[i].d[0]; }
    return 0;
}

// Directional derivative of a whole RK run: u <- Φ(u), du <- Φ'(u) du.
int oracle_rk_run_jvp(oracle_handle* h, int scheme, double* u, double* du, double dt,
                      int nsteps) {
    const Disc& d = h->d;
    const long M = d.ndofs() * d.nv;
    std::vector<Dual<1>> q((size_t)M);
    for (long i = 0; i < M; ++i) { q[i].v = u[i]; q[i].d[0] = du[i]; }
    Work<Dual<1>> wk;
    try {
        rk_run(d, scheme, q.data(), dt, nsteps, wk);
    } catch (std::domain_error& e) {
        h->err = e.what();
        return -5;
    }
    for (long i = 0; i < M; ++i) { u[i] = q[i].v; du[i] = q[i].d[0]; }
    return 0;
}

// OpenMP team size of the calling thread's later parallel regions (a per-thread setting:
// callers that run several oracle handles from their own threads set 1).
void oracle_set_threads(int n) { omp_set_num_threads(n < 1 ? 1 : n); }

// Point-wise exports for closed-form checks of the flux formulas.
double oracle_logmean(double a, double b) { return logarithmic_mean(a, b); }

int oracle_twopoint(oracle_handle* h, const double* Q1, const double* Q2, const double* Ja1,
                    const double* Ja2, double* F) {
    try {
        twopointflux(h->d, Q1, Q2, Ja1, Ja2, F);
    } catch (std::domain_error& e) {
        h->err = e.what();
        return -5;
    }
    return 0;
}

// Numerical flux through a face with frame (n, t, b): rotate, flux, rotate back (no jac).
int oracle_numflux(oracle_handle* h, const double* Ql, const double* Qr, const double* frame,
                   double* F) {
    const Disc& d = h->d;
    try {
        if (d.eq == EQ_ADVECTION) {
            double an = d.a[0] * frame[0];
            for (int k = 1; k < d.nd; ++k) an = an + d.a[k] * frame[k];
            F[0] = (an * (Ql[0] + Qr[0])) / 2.0;
            if (d.sdiss == DISS_LXF) F[0] = F[0] + ((std::fabs(an) * (Ql[0] - Qr[0])) / 2.0) * d.intensity;
        } else {
            double Rl[5], Rr[5], Fr[5];
            rotate2face(d.nd, Ql, frame, Rl);
            rotate2face(d.nd, Qr, frame, Rr);
            euler_numericalflux(d, Rl, Rr, Fr);
            rotate2phys(d.nd, Fr, frame, F);
        }
    } catch (std::domain_error& e) {
        h->err = e.what();
        return -5;
    }
    return 0;
}

// Exterior state that boundary `ib` gives an interior trace Qin at a face node with
// frame (n, t, b) (the point form of applyBCs!, Interfaces.jl:25-49).  GenericBC
// (per-node data) has no point form here: -1.
int oracle_bc_point(oracle_handle* h, int ib, const double* Qin, const double* frame, double* Qext) {
    const Disc& d = h->d;
    if (ib < 0 || ib >= (int)d.bdnames.size()) return -1;
    const int type = d.bctype[ib];
    if (type == BC_DIRICHLET) return -1;
    bc_exterior(d.nd, d.nv, type, type == BC_INFLOW ? d.bcdata[ib].data() : nullptr, Qin, frame, Qext);
    return 0;
}

void oracle_cons2entropy(int nd, double gamma, const double* Q, double* W) {
    cons2entropy(Q, nd, gamma, W);
}

int oracle_entropy2cons(int nd, double gamma, const double* W, double* Q) {
    try {
        entropy2cons(W, nd, gamma, Q);
    } catch (std::domain_error&) {
        return -5;
    }
    return 0;
}

}  // extern "C"
